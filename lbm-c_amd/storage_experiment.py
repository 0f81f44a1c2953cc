#!/usr/bin/env python3
"""Storage-format accuracy on D3Q15 decaying turbulence (BGK + Smagorinsky
C_s = 0.17, tau = 0.51, 256^3, seeded solenoidal field with
E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, u_rms = 0.05): FP32, FP16S and BF16S
populations against FP64 populations, whose kernels match the FP64 CPU
restatement of the reference bit for bit.

At each checkpoint: relative velocity error ||u - u_64|| / ||u_64|| (L2 over
all nodes) and max |u - u_64| / max |u_64|, the relative kinetic-energy
difference, and the mean |log10(E(k)/E_64(k))| over k = 1..L/4 (the
statistical comparison that remains meaningful once the chaotic
trajectories separate).  Also the measured MLUPS of each format.

Writes one JSON document (--out).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import lbmc  # noqa: E402
import workloads  # noqa: E402


def log(msg):
    print(f"[storage {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--checkpoints", type=str, default="10,50,100,1000,5000,20000")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", type=str, default="storage_experiment.json")
    args = ap.parse_args()
    L = args.size
    N = L ** 3
    cps = [int(c) for c in args.checkpoints.split(",") if int(c) <= args.steps]
    fields = workloads.decaying_turbulence(L, k0=4.0, u_rms=0.05, seed=args.seed)
    fields = tuple(v.reshape(-1) for v in fields)
    res = {"size": L, "steps": args.steps, "checkpoints": cps, "seed": args.seed}
    snaps = {}
    perf = {}
    for st in (lbmc.FP64, lbmc.FP32, lbmc.FP16S, lbmc.BF16S):
        name = lbmc.STORAGE_NAMES[st]
        lat = lbmc.Lattice(lbmc.D3Q15, (L, L, L), lbmc.BGK, tau=0.51, c_smag=0.17, storage=st)
        lat.set_fields(geometry=np.zeros(N))
        lat.init_equilibrium(np.ones(N), fields)
        t, ms = 0, 0.0
        snaps[name] = {}
        for c in cps:
            ms += lat.time_steps(t, c - 1 - t, 0)
            lat.step(c - 1, True)
            t = c
            rho, u = lat.get_macros()
            snaps[name][c] = (rho.astype(np.float32), u.astype(np.float32))
        lat.close()
        perf[name] = round(N * (t - len(cps)) / (ms * 1e-3) / 1e6, 1)
        log(f"{name}: {t} steps, {perf[name]} MLUPS")
    res["MLUPS"] = perf
    out = {}
    kmax = L // 4
    for name in ("f32", "f16s", "bf16s"):
        rows = []
        for c in cps:
            r64, u64 = (a.astype(np.float64) for a in snaps["f64"][c])
            r, u = (a.astype(np.float64) for a in snaps[name][c])
            e64 = workloads.energy_spectrum(*(v.reshape(L, L, L) for v in u64))
            e = workloads.energy_spectrum(*(v.reshape(L, L, L) for v in u))
            # shells that carry energy (the initial spectrum falls off as
            # exp(-2 (k/k0)^2): far shells hold only rounding noise)
            k = np.arange(1, kmax + 1)
            k = k[e64[k] > 1e-8 * e64.max()]
            rows.append({
                "step": c,
                "u_rel_l2": float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
                "u_rel_max": float(np.max(np.abs(u - u64)) / np.max(np.abs(u64))),
                "rho_max_abs": float(np.max(np.abs(r - r64))),
                "energy_rel": float(abs(np.sum(u * u) - np.sum(u64 * u64)) / np.sum(u64 * u64)),
                "spectrum_log_err": float(np.mean(np.abs(np.log10(e[k] / e64[k])))),
            })
            log(f"{name} step {c}: " + ", ".join(f"{k_}={v:.3e}" for k_, v in rows[-1].items() if k_ != "step"))
        out[name] = rows
    res["vs_f64"] = out
    with open(args.out, "w") as fh:
        json.dump(res, fh, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
