#!/usr/bin/env python3
"""Benchmark of the fused D3Q15 time step on one MI355X.

Workload (the configuration the D3Q15 MLUPS metric is quoted on): D3Q15
BGK + Smagorinsky (C_s = 0.17), tau = 0.51, 256^3 periodic box, decaying
turbulence from a seeded solenoidal random-phase velocity field with
E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, per-component u_rms = 0.05, f = f_eq
(rho = 1, u) at t = 0; FP32 populations.  A step is one call of the fused
kernel over all 256^3 nodes, with the populations resident in HBM.

Timing: after set-up, ~50 ms of a plain device copy (the roofline's
bandwidth reference) and 100 calls of an identical twin lattice (the GPU's
clocks leave the transient that follows any change of load; the measured
lattice itself starts at t = 0), W untimed warm-up steps, then exactly K
steps between HIP events
recorded on the library's stream (the stream the kernel runs on), with a
stream synchronisation before and after.  One process, one GPU.  Five
further windows of K/5 steps (at least 10) on the same lattice give the
run-to-run spread ("repeats"); the reported value is the K-step window,
"median_value" the MLUPS of the median repeat window.

Printed: one JSON line with MLUPS, the roofline of the step kernel
(120 algorithmic bytes per node per step: 15 FP32 populations read and
written; the periodic box has no boundary nodes, so the step reads no
node mask), and the FP64 CPU oracle timed on a bounded sample of the same
workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lbm-c_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_d3q15_f32_256_traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_lattice(lbmc, L, storage, fields):
    import numpy as np

    ux, uy, uz = fields
    lat = lbmc.Lattice(lbmc.D3Q15, (L, L, L), lbmc.BGK, tau=0.51, c_smag=0.17, storage=storage)
    lat.set_fields(geometry=np.zeros(L ** 3))
    lat.init_equilibrium(np.ones(L ** 3), (ux, uy, uz))
    return lat


def cpu_baseline(L, fields, seconds, threads):
    """FP64 oracle (the CPU restatement of the reference step) on the same
    workload: whole lattice, as many steps as fit in `seconds` (at least 2)."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    pyoracle.set_threads(threads)
    o = pyoracle.Oracle(pyoracle.D3Q15, (L, L, L), pyoracle.BGK, tau=0.51, c_smag=0.17)
    o.set_fields(geometry=np.zeros(L ** 3))
    o.init_equilibrium(np.ones(L ** 3), fields)
    o.step(0, False)  # warm-up

    def timed(t, min_seconds):
        steps, t0 = 0, time.perf_counter()
        while True:
            o.step(t + steps, False)
            steps += 1
            el = time.perf_counter() - t0
            if steps >= 2 and el >= min_seconds:
                return steps, el

    steps, el = timed(1, seconds)
    # one thread, a short sample of the same lattice
    pyoracle.set_threads(1)
    steps1, el1 = timed(1 + steps, 1.0)
    o.close()
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(L ** 3 * steps / el / 1e6, 3), "unit": "MLUPS",
            "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, same D3Q15 BGK+Smagorinsky {L}^3 lattice and initial state, "
                      f"{steps} steps in {el:.1f} s on {threads} OpenMP threads",
            "one_core_MLUPS": round(L ** 3 * steps1 / el1 / 1e6, 3),
            "one_core_sample": f"{steps1} steps in {el1:.1f} s on 1 thread",
            "nproc": os.cpu_count(), "cpu_model": model}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-variants", action="store_true", help="skip the FP16S/BF16S/FP64 lines")
    ap.add_argument("--no-closure", action="store_true", help="skip the learned-closure step line")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble line")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--trace-calls", type=int, default=0, metavar="N",
                    help="diagnostic instead of the benchmark: per-call step times (HIP events) of "
                         "N calls of the workload, FP32 and FP16S: a fresh start after a 2 s idle "
                         "gap, a fresh start after the copy preload, a settled lattice after a 2 s "
                         "idle gap; prints one JSON line")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the time step "
              "has no multi-GPU path in this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import closure
    import lbmc
    import workloads

    L = args.size
    N = L ** 3
    log(f"bench: generating 256^3 turbulence field (seed 2024)" if L == 256 else
        f"bench: generating {L}^3 turbulence field (seed 2024)")
    fields = workloads.decaying_turbulence(L, k0=4.0, u_rms=0.05, seed=2024)
    fields = tuple(v.reshape(-1) for v in fields)

    def preload():
        """A plain 1 GiB copy (reads + writes 1 GiB each) for ~50 ms: the
        bandwidth reference of the roofline.  It also ends the idle that
        host-side set-up leaves.  After an idle gap the shader clock dips for
        ~10-40 calls of load (per-dispatch counter samples: ~2.2 GHz, down to
        ~1.8 GHz, then ~2.45 GHz), whatever the lattice holds; a 20-step
        window started from idle times that transient, not the step."""
        nbytes, reps = 1 << 30, 150
        ms_ = lbmc.copy_bandwidth(nbytes, reps)
        return 2.0 * nbytes * reps / (ms_ * 1e-3) / 1e9

    if args.trace_calls > 0:
        # call-by-call times of (1) a fresh start after the GPU idled for 2 s,
        # (2) a fresh start right after the preload, (3) a lattice settled by
        # 300 calls, after the GPU idled for 2 s: if (1) and (3) slow down
        # alike and (2) does not, the slowdown follows the idle gap, not the
        # lattice's state
        out = {"trace_calls": args.trace_calls, "unit": "us per call", "workload": f"{L}^3"}
        for st in (lbmc.FP32, lbmc.FP16S):
            for start in ("idle", "preload", "settled_idle"):
                tl = make_lattice(lbmc, L, st, fields)
                t0 = 0
                if start == "settled_idle":
                    tl.run(0, 300, 0)
                    tl.sync()
                    t0 = 300
                if start == "preload":
                    preload()
                else:
                    time.sleep(2.0)
                per = [round(tl.time_steps(t0 + t, 1, 0) * 1e3, 1) for t in range(args.trace_calls)]
                tl.close()
                out[f"{lbmc.STORAGE_NAMES[st]}_{start}"] = per
                log(f"bench: {lbmc.STORAGE_NAMES[st]} {start}: " + " ".join(f"{v:.0f}" for v in per))
        print(json.dumps(out), flush=True)
        return

    def steady(factory):
        """A fresh lattice from factory() (t = 0, its state untouched) and an
        identical twin that has just run 100 calls: the GPU then runs the
        measured calls at the clocks of a long run of the same kernel.  A
        change of load (idle, or another kernel) is followed by a clock
        transient of ~10-40 calls that the copy preload alone only shortens
        (bench.py --trace-calls).  The caller closes both after timing."""
        lat = factory()
        twin = factory()
        twin.run(0, 100, 0)
        twin.sync()
        return lat, twin

    log(f"bench: FP32 lattice {L}^3, warmup {args.warmup}, timed {args.steps}")
    copy_gbs = preload()
    log(f"bench: copy kernel {copy_gbs:.0f} GB/s")
    lat, twin = steady(lambda: make_lattice(lbmc, L, lbmc.FP32, fields))
    lat.time_steps(0, args.warmup, 0)
    lat.sync()
    wall0 = time.perf_counter()
    ms = lat.time_steps(args.warmup, args.steps, 0)
    lat.sync()
    wall = time.perf_counter() - wall0
    # five further windows on the same lattice, for the run-to-run spread
    # (value is the K steps above)
    t_next = args.warmup + args.steps
    rwin = max(10, args.steps // 5)
    repeats = []
    for _ in range(5):
        repeats.append(lat.time_steps(t_next, rwin, 0) / rwin)
        t_next += rwin
    lat.step(t_next, True)
    rho, u = lat.get_macros()
    finite = bool(np.all(np.isfinite(rho)) and np.all(np.isfinite(u)))
    lat.close()
    twin.close()

    ms_step = ms / args.steps
    mlups = N * args.steps / (ms * 1e-3) / 1e6
    bytes_node = 2 * 15 * 4  # all-fluid lattice: no mask read
    achieved = bytes_node * N / (ms_step * 1e-3) / 1e9
    log(f"bench: {ms_step:.4f} ms/step, {mlups:.0f} MLUPS, {achieved:.0f} GB/s algorithmic, "
        f"wall {wall:.3f} s, finite={finite}")

    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    variants = {}
    if not args.no_variants:
        # FP16S / BF16S: value-shifted 16-bit slots; FP64: the reference's own
        # arithmetic, bit for bit
        for st in (lbmc.FP16S, lbmc.BF16S, lbmc.FP64):
            v, vtwin = steady(lambda: make_lattice(lbmc, L, st, fields))
            # 100 calls (~20 ms at 16 bits, ~75 ms at FP64): a steady-state
            # window rather than a few milliseconds of the clock's swings
            vsteps = max(100, args.steps)
            v.time_steps(0, args.warmup, 0)
            vms = v.time_steps(args.warmup, vsteps, 0)
            vtwin.close()
            vb = 2 * 15 * (8 if st == lbmc.FP64 else 2)
            variants[lbmc.STORAGE_NAMES[st]] = {
                "MLUPS": round(N * vsteps / (vms * 1e-3) / 1e6, 1),
                "ms_per_step": round(vms / vsteps, 5),
                "bytes_per_node": vb,
                "achieved_GBs": round(vb * N / (vms / vsteps * 1e-3) / 1e9, 1),
                "frac": round(vb * N / (vms / vsteps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            v.close()
            log(f"bench: {lbmc.STORAGE_NAMES[st]} {variants[lbmc.STORAGE_NAMES[st]]}")

    closure_line = None
    if not args.no_closure:
        # same lattice and state, relaxation time from the closure network
        # (He-initialised weights; the cost does not depend on their values
        # beyond the share of nodes that fall back to Smagorinsky)
        mu = np.zeros(7, dtype=np.float32)
        mu[6] = 1.0
        isd = np.full(7, 1e3, dtype=np.float32)
        isd[6] = 1e2
        model = lbmc.Mlp(closure.init_params(7), mu=mu, isd=isd, out_a=0.005, out_b=0.515)
        closure_line = {"mfma_flop_per_node": 2 * (16 * 64 + 64 * 64 + 64 * 16)}
        # both steps over the same calls (a multiple of 8 calls from the
        # first multiple of 8 after the warm-up), each after its own twin
        c0 = (args.warmup + 7) // 8 * 8
        csteps = max(32, (args.steps + 7) // 8 * 8)

        def closure_lattice(every):
            c = make_lattice(lbmc, L, lbmc.FP32, fields)
            c.set_closure(model)
            c.set_closure_interval(every)
            return c

        sm, smtwin = steady(lambda: make_lattice(lbmc, L, lbmc.FP32, fields))
        sm.time_steps(0, c0, 0)
        sms = sm.time_steps(c0, csteps, 0) / csteps
        sm.close()
        smtwin.close()
        closure_line["smagorinsky_ms_per_step"] = round(sms, 5)
        for every in (1, 2, 4, 8):
            c, ctwin = steady(lambda: closure_lattice(every))
            c.time_steps(0, c0, 0)
            cms = c.time_steps(c0, csteps, 0)
            c.close()
            ctwin.close()
            closure_line[f"every_{every}"] = {
                "MLUPS": round(N * csteps / (cms * 1e-3) / 1e6, 1), "ms_per_step": round(cms / csteps, 5),
                "vs_smagorinsky_step": round((cms / csteps) / sms, 3)}
        # the same network evaluated with fp16 operands (lbmc_mlp_set_eval_fp16)
        model.set_eval_fp16(True)
        c, ctwin = steady(lambda: closure_lattice(1))
        c.time_steps(0, c0, 0)
        cms = c.time_steps(c0, csteps, 0)
        c.close()
        ctwin.close()
        closure_line["every_1_fp16_operands"] = {
            "MLUPS": round(N * csteps / (cms * 1e-3) / 1e6, 1), "ms_per_step": round(cms / csteps, 5),
            "vs_smagorinsky_step": round((cms / csteps) / sms, 3)}
        model.close()
        log(f"bench: closure {closure_line}")

    ensemble_line = None
    if not args.no_ensemble:
        # 64 uneven D2Q9 channels with discs (seeded), FP32, one batched
        # launch per step, autotuned launch shape.  Members are MRT or
        # BGK + learned closure (one shared network trained on the
        # ensemble's own filtered refined runs, shifted to each member's
        # tau; fallback band [0.5005, 2]); timed with the network evaluated
        # every call and every 8th call, and with plain BGK in place of
        # BGK + learned.
        specs = workloads.ensemble_d2q9(64, 0)
        EN = sum(m["length"][0] * m["length"][1] for m in specs)
        ens_p, ens_mu, ens_isd, ens_a, ens_b, ENS_TAU_REF = workloads.ensemble_closure()
        model = lbmc.Mlp(ens_p, mu=ens_mu, isd=ens_isd, out_a=ens_a, out_b=ens_b)

        def run_ensemble(every):
            lats = []
            for m in specs:
                Lx, Ly = m["length"]
                lt = lbmc.Lattice(lbmc.D2Q9, (Lx, Ly), lbmc.MRT if m["collision"] == "MRT" else lbmc.BGK,
                                  forcing=True, tau=m["tau"], tau_mrt=m["rates"], storage=lbmc.FP32)
                lt.set_fields(geometry=m["geometry"].reshape(-1),
                              force=[np.full(Lx * Ly, m["force"]), np.zeros(Lx * Ly)])
                lt.init_static()
                if every and m["collision"] != "MRT":
                    lt.set_closure(model, 0.5005, 2.0)
                    lt.set_closure_shift(m["tau"] - ENS_TAU_REF)
                    lt.set_closure_interval(every)
                lats.append(lt)
            b = lbmc.Batch(lats)
            grid, _ = b.autotune(0, steps_per_trial=16)
            _, chunk_nodes, bpc = b.shape()
            esteps = max(64, args.steps)
            preload()
            b.time_steps(1000, 64, 0)
            ems = b.time_steps(1064, esteps, 0)
            b.close()
            for lt in lats:
                lt.close()
            return {"MLUPS": round(EN * esteps / (ems * 1e-3) / 1e6, 1), "ms_per_step": round(ems / esteps, 5),
                    "grid": grid, "chunk_nodes": chunk_nodes, "blocks_per_cu": bpc,
                    "achieved_GBs": round(73 * EN / (ems / esteps * 1e-3) / 1e9, 1),
                    "frac": round(73 * EN / (ems / esteps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

        ensemble_line = {"members": 64, "nodes": EN, "bytes_per_node": 2 * 9 * 4 + 1,
                         "members_kind": "MRT or BGK + learned closure (trained on the ensemble's filtered refined runs)",
                         "closure_every_1": run_ensemble(1), "closure_every_8": run_ensemble(8),
                         "plain_bgk_mrt": run_ensemble(0)}
        model.close()
        log(f"bench: ensemble {ensemble_line}")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"bench: CPU oracle baseline, {threads} threads, ~{args.cpu_seconds} s")
        cpu = cpu_baseline(L, fields, args.cpu_seconds, threads)
        log(f"bench: CPU {cpu}")

    out = {
        "metric": "D3Q15 MLUPS",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"D3Q15 BGK+Smagorinsky decaying turbulence {L}^3",
                   "lattice": "D3Q15", "dims": [L, L, L], "tau": 0.51, "c_smag": 0.17,
                   "storage": "f32", "init": "f_eq(1, u), solenoidal k^4 exp(-2(k/4)^2), "
                   "u_rms 0.05, seed 2024", "parallelism": "single GPU"},
        "median_value": round(N / (sorted(repeats)[2] * 1e-3) / 1e6, 1),
        "repeats": {"steps_each": rwin, "ms_per_step": [round(r, 5) for r in repeats],
                    "median_ms_per_step": round(sorted(repeats)[2], 5)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "bytes_per_node": bytes_node,
                     "copy_kernel_GBs": round(copy_gbs, 1)},
        "cpu_baseline": cpu,
        "variants": variants,
        "closure": closure_line,
        "ensemble": ensemble_line,
        "finite": finite,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
