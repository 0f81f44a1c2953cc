#!/usr/bin/env python3
"""Training the ensemble's shared closure network a priori from filtered
finer runs of the ensemble's own channels, and testing it a posteriori.

  1. every BGK member of the ensemble (workloads.ensemble_d2q9) is also run
     on its grid refined by r = 2 (each coarse node a 2 x 2 block of the
     same kind; acoustic scaling: tau_f = 1/2 + r (tau - 1/2), force F / r,
     the same Reynolds number), from the plane Poiseuille profile plus
     noise, FP32 storage.  Snapshots are box-filtered by r; the subgrid
     stress and the filtered strain rate give the pointwise least-squares
     eddy viscosity nu_t and the target correction dtau = 3 nu_t; inputs are
     what the coarse lattice sees before collision, Pi_ab / rho =
     -(2/3) (tau + dtau) S_ab and rho (fluid nodes with no solid neighbour);
  2. one network for every member is trained on the device on the pooled
     samples: tau' = tau_member + out_a y + out_b (lbmc_set_closure_shift
     carries tau_member);
  3. a posteriori, from another noise seed: each BGK member's coarse run
     with the trained closure, without a model (plain BGK at tau) and with
     Smagorinsky, against its filtered fine run with FP64 storage at the
     same physical time: relative L2 error of u over the fluid nodes;
  4. writes the trained network (--params, numpy .npz: params, mu, isd,
     out_a, out_b) and a JSON summary (--out).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import closure  # noqa: E402
import lbmc  # noqa: E402
import workloads  # noqa: E402


def log(msg):
    print(f"[ensemble-train {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def box2(a, r):
    """Mean over r x r blocks of a (Y, X) array."""
    Y, X = a.shape
    return a.reshape(Y // r, r, X // r, r).mean(axis=(1, 3))


def strain2(u):
    """Sbar = (xx, yy, xy) of u = (ux, uy) on (Y, X) by central differences
    (periodic rolls; values at solid nodes are not used)."""
    def dx(f):
        return 0.5 * (np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1))

    def dy(f):
        return 0.5 * (np.roll(f, -1, axis=0) - np.roll(f, 1, axis=0))

    return np.stack([dx(u[0]), dy(u[1]), 0.5 * (dy(u[0]) + dx(u[1]))])


def nu_t2(ts, S):
    """Pointwise least-squares nu_t >= 0 with tau_dev ~ -2 nu_t S (2-D,
    packed xx, yy, xy; deviatoric part with the 2-D trace)."""
    tr = 0.5 * (ts[0] + ts[1])
    dxx, dyy, dxy = ts[0] - tr, ts[1] - tr, ts[2]
    num = -(dxx * S[0] + dyy * S[1] + 2.0 * dxy * S[2])
    den = 2.0 * (S[0] * S[0] + S[1] * S[1] + 2.0 * S[2] * S[2])
    return np.maximum(0.0, num / np.maximum(den, 1e-300))


def refine(m, r):
    Lx, Ly = m["length"]
    return dict(m, length=(r * Lx, r * Ly), geometry=np.kron(m["geometry"], np.ones((r, r))),
                tau=0.5 + r * (m["tau"] - 0.5), force=m["force"] / r)


def fine_start(m, r, noise, rng):
    """Poiseuille profile of the coarse member evaluated at the fine nodes
    (wall half a coarse node outside the first fluid row), zero in the
    solids, plus noise in the fluid."""
    Lx, Ly = m["length"]
    H = Ly - 2
    yf = np.arange(r * Ly)
    yc = (yf + 0.5) / r - 0.5
    eta = yc - 0.5
    prof = np.clip(6.0 * m["u_mean"] * eta * (H - eta) / (H * H), 0.0, None)
    ux = np.broadcast_to(prof[:, None], (r * Ly, r * Lx)).copy()
    uy = np.zeros_like(ux)
    solid = np.kron(m["geometry"], np.ones((r, r))) >= 1.0
    ux[solid] = 0.0
    if noise > 0.0:
        ux += noise * rng.standard_normal(ux.shape) * ~solid
        uy += noise * rng.standard_normal(ux.shape) * ~solid
    return ux, uy


def lattice(m, storage, c_smag=0.0, model=None):
    Lx, Ly = m["length"]
    n = Lx * Ly
    lat = lbmc.Lattice(lbmc.D2Q9, (Lx, Ly), lbmc.BGK, forcing=True, tau=m["tau"], c_smag=c_smag,
                       storage=storage)
    lat.set_fields(geometry=m["geometry"].reshape(-1), force=[np.full(n, m["force"]), np.zeros(n)])
    if model is not None:
        lat.set_closure(model, 0.5005, 2.0)
        lat.set_closure_shift(m["tau"])
    return lat


def interior(m):
    """Coarse fluid nodes with no solid node among their 8 neighbours."""
    g = m["geometry"] >= 1.0
    near = g.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            near |= np.roll(np.roll(g, dy, 0), dx, 1)
    return ~near


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0, help="ensemble seed (workloads.ensemble_d2q9)")
    ap.add_argument("--ratio", type=int, default=2)
    ap.add_argument("--steps", type=int, default=2000, help="coarse steps of every run")
    ap.add_argument("--snapshots", type=int, default=4, help="training snapshots per member, evenly spaced")
    ap.add_argument("--noise", type=float, default=2e-3)
    ap.add_argument("--train-seed", type=int, default=1)
    ap.add_argument("--test-seed", type=int, default=2)
    ap.add_argument("--per-snapshot", type=int, default=20000, help="samples kept per member and snapshot")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--c-smag", type=float, default=0.1, help="Smagorinsky constant of the baseline runs")
    ap.add_argument("--params", type=str, default=os.path.join(HERE, "data", "ensemble_closure.npz"))
    ap.add_argument("--out", type=str, default="ensemble_training.json")
    args = ap.parse_args()

    r = args.ratio
    specs = workloads.ensemble_d2q9(args.count, args.seed)
    bgk = [k for k, m in enumerate(specs) if m["collision"] == "BGK"]
    res = {"ratio": r, "steps": args.steps, "members": bgk, "noise": args.noise}
    log(f"{len(bgk)} BGK members, refinement {r}, {args.steps} coarse steps")

    # ---- 1. training samples from the refined runs
    rng = np.random.default_rng(args.train_seed)
    X, Y = [], []
    t0 = time.perf_counter()
    for k in bgk:
        m = specs[k]
        mf = refine(m, r)
        ux, uy = fine_start(m, r, args.noise, rng)
        lat = lattice(mf, lbmc.FP32)
        lat.init_equilibrium(np.ones(ux.size), [ux.reshape(-1), uy.reshape(-1)])
        keep = interior(m).reshape(-1)
        Lxf, Lyf = mf["length"]
        t, every = 0, r * args.steps // args.snapshots
        for _ in range(args.snapshots):
            lat.run(t, every, every)
            t += every
            rho, u = lat.get_macros()
            rho = rho.reshape(Lyf, Lxf)
            u = u.reshape(2, Lyf, Lxf)
            rb = box2(rho, r)
            ub = np.stack([box2(u[0], r), box2(u[1], r)])
            ts = np.stack([box2(u[0] * u[0], r) - ub[0] * ub[0], box2(u[1] * u[1], r) - ub[1] * ub[1],
                           box2(u[0] * u[1], r) - ub[0] * ub[1]])
            S = strain2(ub)
            dtau = 3.0 * nu_t2(ts, S)
            tp = m["tau"] + dtau
            f = np.zeros(rb.shape + (7,), dtype=np.float32)
            f[..., 0] = -(2.0 / 3.0) * tp * S[0]
            f[..., 1] = -(2.0 / 3.0) * tp * S[1]
            f[..., 3] = -(2.0 / 3.0) * tp * S[2]
            f[..., 6] = rb
            f = f.reshape(-1, 7)[keep]
            d = dtau.reshape(-1)[keep].astype(np.float32)
            pick = rng.choice(f.shape[0], size=min(args.per_snapshot, f.shape[0]), replace=False)
            X.append(f[pick])
            Y.append(d[pick])
        lat.close()
    X = np.concatenate(X)
    Y = np.concatenate(Y)
    res["train_samples"] = int(X.shape[0])
    res["fine_train_seconds"] = round(time.perf_counter() - t0, 1)
    res["target"] = {"mean": float(Y.mean()), "p99": float(np.percentile(Y, 99)), "share_positive": float(np.mean(Y > 0))}
    log(f"{X.shape[0]} samples: dtau mean {Y.mean():.3e}, p99 {np.percentile(Y, 99):.3e}, "
        f"share > 0 {np.mean(Y > 0):.3f}")

    # ---- 2. one network for every member
    mu, isd, out_a, out_b = closure.fit_scaling(X, Y)
    params0 = closure.init_params(args.train_seed)
    model = lbmc.Mlp(params0, mu=mu, isd=isd, out_a=out_a, out_b=out_b)
    model.set_data(X, (Y - out_b) / out_a)
    t0 = time.perf_counter()
    hist = model.train(args.epochs, batch=args.batch, lr=args.lr, seed=args.train_seed)
    res["train_seconds"] = round(time.perf_counter() - t0, 2)
    res["train_loss"] = [float(v) for v in hist]
    log(f"trained: loss {hist[0]:.4f} -> {hist[-1]:.4f} (normalised target variance 1)")
    params = model.get_params()
    os.makedirs(os.path.dirname(os.path.abspath(args.params)), exist_ok=True)
    np.savez(args.params, params=params, mu=mu, isd=isd, out_a=np.float64(out_a), out_b=np.float64(out_b))
    log(f"wrote {args.params}")

    # ---- 3. a posteriori against the filtered fine runs
    rng = np.random.default_rng(args.test_seed)
    errs = {"learned": [], "no_model": [], f"smagorinsky_cs{args.c_smag:g}": []}
    for k in bgk:
        m = specs[k]
        mf = refine(m, r)
        ux, uy = fine_start(m, r, args.noise, rng)
        ref = lattice(mf, lbmc.FP64)
        ref.init_equilibrium(np.ones(ux.size), [ux.reshape(-1), uy.reshape(-1)])
        ref.run(0, r * args.steps, r * args.steps)
        Lxf, Lyf = mf["length"]
        _, u = ref.get_macros()
        ref.close()
        u = u.reshape(2, Lyf, Lxf)
        uref = np.stack([box2(u[0], r), box2(u[1], r)]).reshape(2, -1)
        u0 = [box2(ux, r).reshape(-1), box2(uy, r).reshape(-1)]
        fluid = (m["geometry"] < 1.0).reshape(-1)
        for name in errs:
            if name == "learned":
                lat = lattice(m, lbmc.FP32, model=model)
            elif name == "no_model":
                lat = lattice(m, lbmc.FP32)
            else:
                lat = lattice(m, lbmc.FP32, c_smag=args.c_smag)
            lat.init_equilibrium(np.ones(u0[0].size), u0)
            lat.run(0, args.steps, args.steps)
            _, uc = lat.get_macros()
            lat.close()
            d = uc[:, fluid] - uref[:, fluid]
            e = float(np.linalg.norm(d) / np.linalg.norm(uref[:, fluid])) if np.all(np.isfinite(uc)) else float("inf")
            errs[name].append(e)
        log(f"member {k} {m['length']} Re {m['Re']:.0f} tau {m['tau']:.4f}: " +
            ", ".join(f"{n} {v[-1]:.3e}" for n, v in errs.items()))
    res["error_per_member"] = errs
    res["error_median"] = {n: float(np.median(v)) for n, v in errs.items()}
    res["learned_better_than_no_model"] = int(np.sum(np.array(errs["learned"]) < np.array(errs["no_model"])))
    res["Re"] = [specs[k]["Re"] for k in bgk]
    log("median relative L2 error: " + ", ".join(f"{n} {v:.3e}" for n, v in res["error_median"].items()) +
        f"; learned better than no model on {res['learned_better_than_no_model']} of {len(bgk)} members")
    model.close()
    with open(args.out, "w") as fh:
        json.dump(res, fh, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
