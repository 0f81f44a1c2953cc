"""D3Q15 step rate against lattice size (BGK + Smagorinsky, or MRT).

Each size is a periodic cube holding a 64^3 decaying-turbulence field
(k0 4, u_rms 0.05, seed 7) tiled to fill it; after 60 untimed calls, three
windows of 40 calls are timed with HIP events and the best is reported.
The addressing form of the step follows the library's defaults
(LBMC_FORCE_WIDE=1 in the environment selects the 64-bit-offset kernels,
which address slot rows from wave-uniform bases when Lx % 64 == 0).

    python lbm-c_amd/size_sweep.py --sizes 256 512 640 768 --storage fp32
    python lbm-c_amd/size_sweep.py --sizes 256 --collision mrt

MRT uses the corrected D3Q15 M^-1 (LBMC_FIX_D3Q15_MINV; the reference's
table is all zeros, so its MRT relaxes nothing), tau 0.6: rates 1/tau on
the stress moments and 1.2-1.6 on the others.

One process per size keeps the device memory of one lattice at a time.
"""
import argparse
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
P = 64


# D3Q15 MRT rates in the moment order of M (rho, e, eps, jx, qx, jy, qy, jz,
# qz, 3pxx, pww, pxy, pyz, pxz, mxyz): stress moments at 1/tau
def mrt_rates(tau):
    s = 1.0 / tau
    return [0.0, 1.6, 1.2, 0.0, 1.6, 0.0, 1.6, 0.0, 1.6, s, s, s, s, s, 1.2]


def one(L, storage, collision):
    sys.path.insert(0, HERE)
    import numpy as np

    import lbmc
    import workloads

    st = {"fp64": lbmc.FP64, "fp32": lbmc.FP32, "fp16s": lbmc.FP16S, "bf16s": lbmc.BF16S}[storage]
    bpn = {"fp64": 240, "fp32": 120, "fp16s": 60, "bf16s": 60}[storage]
    u_small = workloads.decaying_turbulence(P, k0=4.0, u_rms=0.05, seed=7)
    n = L ** 3
    reps = (L // P,) * 3
    u = [np.tile(v, reps).reshape(-1) for v in u_small]
    if collision == "mrt":
        lat = lbmc.Lattice(lbmc.D3Q15, (L, L, L), lbmc.MRT, tau=0.6, tau_mrt=mrt_rates(0.6), storage=st,
                           quirks=lbmc.FIX_D3Q15_MINV)
    else:
        lat = lbmc.Lattice(lbmc.D3Q15, (L, L, L), lbmc.BGK, tau=0.51, c_smag=0.17, storage=st)
    lat.set_fields(geometry=np.zeros(n))
    lat.init_equilibrium(np.ones(n), u)
    del u
    lat.time_steps(0, 60, 0)
    windows, t = [], 60
    for _ in range(3):
        windows.append(lat.time_steps(t, 40, 0) / 40)
        t += 40
    lat.close()
    ms = min(windows)
    return {"L": L, "storage": storage, "collision": collision, "ms_per_step": [round(w, 4) for w in windows],
            "MLUPS": round(n / ms / 1e3, 1), "GBs": round(bpn * n / (ms * 1e-3) / 1e9, 1),
            "force_wide": os.environ.get("LBMC_FORCE_WIDE", "0")}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--sizes", type=int, nargs="+", default=[256, 512, 640, 768])
    ap.add_argument("--storage", default="fp32", choices=["fp64", "fp32", "fp16s", "bf16s"])
    ap.add_argument("--collision", default="bgk", choices=["bgk", "mrt"])
    ap.add_argument("--one", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.one:
        print(json.dumps(one(args.one, args.storage, args.collision)), flush=True)
        return
    for L in args.sizes:
        if L % P:
            sys.exit(f"size {L}: must be a multiple of {P}")
        r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--one", str(L), "--storage", args.storage,
                            "--collision", args.collision], check=False)
        if r.returncode != 0:
            sys.exit(r.returncode)


if __name__ == "__main__":
    main()
