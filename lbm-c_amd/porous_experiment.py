#!/usr/bin/env python3
"""D2Q9 MRT flow through a seeded porous medium (2048 x 1024, discs of
radius 8..24 to a solid fraction of 0.30, Zou-He pressure inlet rho_in and
outlet rho_out, walls top and bottom, Lallemand-Luo rates with shear rates
1/tau): permeability per storage format and step throughput.

Permeability (Darcy, lattice units): k = nu <u_x> L_x / (c_s^2 (rho_in - rho_out)),
c_s^2 = 1/3, <u_x> the mean over all nodes (solids count as zero), at
checkpoints up to --steps calls.  FP64 storage runs the kernels that match
the FP64 CPU restatement of the reference bit for bit; the other formats
are compared with it.  Note: the FP32 lattice (76 MB) fits in the 256 MB
last-level cache, so its throughput is not an HBM roofline figure.

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
    print(f"[porous {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--lx", type=int, default=2048)
    ap.add_argument("--ly", type=int, default=1024)
    ap.add_argument("--tau", type=float, default=0.6)
    ap.add_argument("--rho-in", type=float, default=1.001)
    ap.add_argument("--rho-out", type=float, default=1.000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--checkpoints", type=int, default=5)
    ap.add_argument("--out", type=str, default="porous_experiment.json")
    args = ap.parse_args()

    Lx, Ly, tau = args.lx, args.ly, args.tau
    geo, discs = workloads.porous_discs(Lx, Ly, 8, 24, 0.30, seed=args.seed)
    macro = np.zeros((Ly, Lx), dtype=np.int32)
    micro = np.zeros((Ly, Lx), dtype=np.int32)
    macro[1:-1, 0] = 8
    macro[1:-1, -1] = 8
    micro[1:-1, 0] = 1
    micro[1:-1, -1] = 2
    rho = np.ones((Ly, Lx))
    rho[:, 0] = args.rho_in
    rho[:, -1] = args.rho_out
    s = 1.0 / tau
    rates = [0.0, 1.64, 1.54, 0.0, 1.9, 0.0, 1.9, s, s]
    nu = (tau - 0.5) / 3.0
    N = Lx * Ly
    res = {"dims": [Lx, Ly], "tau": tau, "rho_in": args.rho_in, "rho_out": args.rho_out,
           "solid_fraction": float(geo.mean()), "discs": len(discs), "steps": args.steps}
    log(f"{Lx}x{Ly}, {len(discs)} discs, solid fraction {geo.mean():.4f}")
    per = args.steps // args.checkpoints
    perms = {}
    for st in (lbmc.FP64, lbmc.FP32, lbmc.FP16S, lbmc.BF16S):
        name = lbmc.STORAGE_NAMES[st]
        lat = lbmc.Lattice(lbmc.D2Q9, (Lx, Ly), lbmc.MRT, macro_bc=True, micro_bc=True, tau=tau,
                           tau_mrt=rates, storage=st)
        lat.set_fields(geometry=geo.reshape(-1), macro_bc=macro.reshape(-1), micro_bc=micro.reshape(-1),
                       rho=rho.reshape(-1))
        lat.init_static()
        ks, t, ms_total = [], 0, 0.0
        for c in range(args.checkpoints):
            ms_total += lat.time_steps(t, per - 1, 0)
            lat.step(t + per - 1, True)
            t += per
            _, u = lat.get_macros()
            k = nu * float(np.mean(u[0])) * Lx / ((args.rho_in - args.rho_out) / 3.0)
            ks.append({"step": t, "k": k, "finite": bool(np.all(np.isfinite(u)))})
        lat.close()
        steps_timed = args.checkpoints * (per - 1)
        perms[name] = {"k": ks, "MLUPS": round(N * steps_timed / (ms_total * 1e-3) / 1e6, 1),
                       "ms_per_step": round(ms_total / steps_timed, 5)}
        log(f"{name}: k = {ks[-1]['k']:.6g} at step {t}, {perms[name]['MLUPS']} MLUPS")
    ref = perms["f64"]["k"][-1]["k"]
    for name in ("f32", "f16s", "bf16s"):
        perms[name]["k_rel_err_vs_f64"] = abs(perms[name]["k"][-1]["k"] - ref) / abs(ref)
    res["runs"] = perms
    with open(args.out, "w") as fh:
        json.dump(res, fh, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
