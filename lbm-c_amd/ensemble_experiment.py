#!/usr/bin/env python3
"""Batched ensemble of uneven D2Q9 lattices (workloads.ensemble_d2q9):
aggregate throughput of the batched step and each member's error against a
reference ensemble.

Members are MRT or BGK; in the "learned" mix the BGK members use the learned
relaxation-time closure (one shared network trained on the ensemble's own
box-filtered refined runs, workloads.ensemble_closure / ensemble_training.py,
shifted to each member's tau; fallback band [0.5005, 2.0]), evaluated every
call or every 8th call.

  1. build the members, static start; autotune the launch shape on a
     throw-away copy, then time K steps: the plain mix at FP32, FP16S,
     BF16S and FP64 storage, the learned mix at FP32, FP16S and BF16S;
  2. from a developed start (the plane Poiseuille profile at each member's
     mean speed; --start static: f = w, as LBM-C's Init 0), run the plain
     mix N steps (default 10,000) at every storage and compare
     u per member against FP64 storage (whose kernels match the FP64 CPU
     restatement of the reference bit for bit): relative L2 error over fluid
     nodes, and the member's Reynolds number from its mean velocity;
  3. run the learned mix N steps at FP32, FP16S and BF16S and compare the
     16-bit formats against FP32 (the closure needs FP32 arithmetic, so
     there is no FP64 run of it).

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
    print(f"[ensemble {time.strftime('%H:%M:%S')}] {msg}", flush=True)


# the shared network (trained: workloads.ensemble_closure); a member of
# relaxation time tau takes the shift tau - TAU_REF
_P, _MU, _ISD, _OUT_A, _OUT_B, TAU_REF = workloads.ensemble_closure()


def make_model():
    return lbmc.Mlp(_P, mu=_MU, isd=_ISD, out_a=_OUT_A, out_b=_OUT_B)


def build(specs, storage, model=None, every=1, developed=False):
    """The members; with a model, BGK members use the learned closure; a
    static start (f = w) or a developed one (workloads.ensemble_start)."""
    lats = []
    for m in specs:
        Lx, Ly = m["length"]
        coll = lbmc.MRT if m["collision"] == "MRT" else lbmc.BGK
        lat = lbmc.Lattice(lbmc.D2Q9, (Lx, Ly), coll, forcing=True, tau=m["tau"], tau_mrt=m["rates"],
                           storage=storage)
        n = Lx * Ly
        lat.set_fields(geometry=m["geometry"].reshape(-1), force=[np.full(n, m["force"]), np.zeros(n)])
        if developed:
            rho, u = workloads.ensemble_start(m)
            lat.init_equilibrium(rho, u)
        else:
            lat.init_static()
        if model is not None and coll == lbmc.BGK:
            lat.set_closure(model, 0.5005, 2.0)
            lat.set_closure_shift(m["tau"] - TAU_REF)
            lat.set_closure_interval(every)
        lats.append(lat)
    return lats


def member_errors(specs, got, ref):
    e = []
    for k, m in enumerate(specs):
        fluid = m["geometry"].reshape(-1) < 1
        d = got[k][:, fluid] - ref[k][:, fluid]
        e.append(float(np.linalg.norm(d) / max(np.linalg.norm(ref[k][:, fluid]), 1e-300)))
    return {"max": max(e), "median": float(np.median(e)), "per_member": e}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--steps", type=int, default=10000, help="steps of the error study")
    ap.add_argument("--time-steps", type=int, default=1000, help="timed steps of the throughput runs")
    ap.add_argument("--start", choices=["developed", "static"], default="developed",
                    help="initial state of the error study: the plane Poiseuille profile at each member's "
                         "mean speed (the flows start at their target Reynolds number) or f = w")
    ap.add_argument("--out", type=str, default="ensemble_experiment.json")
    ap.add_argument("--plain-fp32-only", action="store_true",
                    help="only the plain-mix FP32 throughput run (short; for kernel traces)")
    args = ap.parse_args()

    specs = workloads.ensemble_d2q9(args.count, args.seed)
    N = sum(m["length"][0] * m["length"][1] for m in specs)
    res = {"members": args.count, "seed": args.seed, "nodes": N, "steps": args.steps, "start": args.start,
           "sizes": [list(m["length"]) for m in specs], "Re_target": [m["Re"] for m in specs],
           "tau": [m["tau"] for m in specs], "collision": [m["collision"] for m in specs]}
    log(f"{args.count} members, {N} nodes")

    model = make_model()

    # ---- throughput (autotuned launch shape)
    def timed(st, learned_every=0):
        lats = build(specs, st, model if learned_every else None, learned_every or 1)
        b = lbmc.Batch(lats)
        grid, _ = b.autotune(0, steps_per_trial=20)
        shape = b.shape()
        b.close()
        for lat in lats:
            lat.init_static()
        b = lbmc.Batch(lats)
        b.set_shape(*shape)
        b.time_steps(0, 64, 0)  # warm-up (graph build)
        ms = b.time_steps(64, args.time_steps, 0)
        info = b.info()
        b.close()
        for lat in lats:
            lat.close()
        return {"MLUPS": round(N * args.time_steps / (ms * 1e-3) / 1e6, 1),
                "ms_per_step": round(ms / args.time_steps, 5), "grid": grid, "chunks": info["chunks"],
                "chunk_nodes": info["chunk_nodes"], "blocks_per_cu": info["blocks_per_cu"]}

    thr = {}
    for st in (lbmc.FP32, lbmc.FP16S, lbmc.BF16S, lbmc.FP64):
        name = lbmc.STORAGE_NAMES[st]
        thr[name] = timed(st)
        log(f"plain {name}: {thr[name]}")
        if args.plain_fp32_only:
            model.close()
            return
    res["throughput"] = thr
    thr_l = {}
    for st in (lbmc.FP32, lbmc.FP16S, lbmc.BF16S):
        name = lbmc.STORAGE_NAMES[st]
        for every in (1, 8):
            thr_l[f"{name}_every_{every}"] = timed(st, every)
            log(f"learned {name} every {every}: {thr_l[f'{name}_every_{every}']}")
    res["throughput_learned"] = thr_l

    # ---- accuracy after the full run, against FP64
    finals = {}
    for st in (lbmc.FP64, lbmc.FP32, lbmc.FP16S, lbmc.BF16S):
        name = lbmc.STORAGE_NAMES[st]
        lats = build(specs, st, developed=args.start == "developed")
        b = lbmc.Batch(lats)
        t0 = time.perf_counter()
        b.run(0, args.steps, args.steps)
        el = time.perf_counter() - t0
        finals[name] = [lat.get_macros()[1] for lat in lats]
        b.close()
        for lat in lats:
            lat.close()
        log(f"{name}: {args.steps} steps in {el:.2f} s")
    errs = {}
    ref = finals["f64"]
    for name in ("f32", "f16s", "bf16s"):
        errs[name] = member_errors(specs, finals[name], ref)
        log(f"{name} vs f64: max rel L2 {errs[name]['max']:.3e}, median {errs[name]['median']:.3e}")
    res["error_vs_f64"] = errs

    # ---- the learned mix (network every call): 16-bit formats against FP32
    finals_l = {}
    for st in (lbmc.FP32, lbmc.FP16S, lbmc.BF16S):
        name = lbmc.STORAGE_NAMES[st]
        lats = build(specs, st, model, 1, developed=args.start == "developed")
        b = lbmc.Batch(lats)
        b.run(0, args.steps, args.steps)
        finals_l[name] = [lat.get_macros()[1] for lat in lats]
        b.close()
        for lat in lats:
            lat.close()
    errs_l = {}
    for name in ("f16s", "bf16s"):
        errs_l[name] = member_errors(specs, finals_l[name], finals_l["f32"])
        log(f"learned mix {name} vs f32: max rel L2 {errs_l[name]['max']:.3e}, median {errs_l[name]['median']:.3e}")
    res["learned_error_vs_f32"] = errs_l
    model.close()
    re_meas = []
    for k, m in enumerate(specs):
        fluid = m["geometry"].reshape(-1) < 1
        U = float(np.mean(ref[k][0][fluid]))
        D = 2.0 * float(np.mean([d[2] for d in m["discs"]]))
        re_meas.append(U * D / ((m["tau"] - 0.5) / 3.0))
    res["Re_measured_f64"] = re_meas
    with open(args.out, "w") as fh:
        json.dump(res, fh, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
