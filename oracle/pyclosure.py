"""FP64 reference of the step with the learned relaxation-time closure.

TEST INFRASTRUCTURE ONLY: the checker of the HIP closure step; the product
never imports it.

Scope: periodic lattices of fluid and solid nodes, BGK collision with
optional uniform Guo force, the reference's call semantics
(LBM-C/solver.cu:10-136): call 0 collides in place, every later call first
pulls f_i(x) <- f_i(x - e_i) and then collides; solids bounce back
(LBM-C/collision.cu:423-443).  At fluid nodes tau comes from the network
(oracle/pymlp.py, the evaluation's bf16 -- or with fp16=True fp16 --
quantisation points emulated) as out_a * y + out_b
+ shift (shift: the lattice's offset, 0 by default; evaluated on calls t % every == 0 and the first call, the recorded tau
reused in between), or from the Smagorinsky formula (LBM-C/collision.cu:445-463, restated in
closed form: tau' = (tau + sqrt(tau^2 + 36 C^2 (sum|Pi_aa| + 2 sum_{a<b}
|Pi_ab|) / rho)) / 2) when that leaves [lo, hi].  The Guo terms follow
LBM-C/collision.cu:69-111 with the base tau in the source.  Lattice tables
come from the C oracle (oracle/lbmc_oracle.c), whose tables are pinned by
tests/test_oracle_tables.py.
"""
from __future__ import annotations

import numpy as np

import pymlp
import pyoracle


class ClosureOracle:
    def __init__(self, lattice, length, tau, c_smag, params, mu, isd, out_a, out_b, lo, hi,
                 force=None, solid=None, quant=True, every=1, shift=0.0, fp16=False):
        self.lattice = lattice
        self.dim = 2 if lattice == pyoracle.D2Q9 else 3
        e, w, opp, _, _ = pyoracle.tables(lattice)
        self.e = e[: self.dim].astype(np.float64)
        self.w = w
        self.opp = opp
        self.q = w.size
        self.shape = tuple(reversed(length[: self.dim]))  # (Lz,) Ly, Lx
        self.tau, self.c_smag = float(tau), float(c_smag)
        self.p = params
        self.mu = np.asarray(mu, dtype=np.float32)
        self.isd = np.asarray(isd, dtype=np.float32)
        self.out_a, self.out_b = float(out_a), float(out_b)
        self.shift = float(shift)
        self.lo, self.hi = float(lo), float(hi)
        self.F = np.zeros(self.dim) if force is None else np.asarray(force, dtype=np.float64)[: self.dim]
        self.guo = force is not None
        n = int(np.prod(self.shape))
        self.solid = np.zeros(n, dtype=bool) if solid is None else np.asarray(solid, dtype=bool).reshape(-1)
        self.quant = ("fp16" if fp16 else True) if quant else False
        self.every = int(every)
        self.record = None   # tau of the last evaluation per node, NaN where it fell back
        self.f = None
        self.calls = 0
        self.fallback = None

    def init_equilibrium(self, rho, u):
        self.f = self.equilibrium(np.asarray(rho, dtype=np.float64).reshape(-1),
                                  np.asarray(u, dtype=np.float64).reshape(self.dim, -1))

    def equilibrium(self, rho, u):
        eu = 3.0 * (self.e.T @ u)                  # [Q, N]
        usq = 1.5 * np.sum(u * u, axis=0)
        return self.w[:, None] * rho[None, :] * (1.0 + eu + 0.5 * eu * eu - usq[None, :])

    def stream(self):
        axes = {0: self.dim - 1, 1: self.dim - 2, 2: 0}   # x, y, z -> array axis
        out = np.empty_like(self.f)
        for i in range(self.q):
            a = self.f[i].reshape(self.shape)
            for d in range(self.dim):
                s = int(self.e[d, i])
                if s:
                    a = np.roll(a, s, axis=axes[d])
            out[i] = a.reshape(-1)
        self.f = out

    def features(self, f):
        rho = f.sum(axis=0)
        j = self.e @ f
        u = j / rho
        if self.guo:
            u = u + self.F[:, None] / (2.0 * rho[None, :])
        feq = self.equilibrium(rho, u)
        fn = f - feq
        pi = {}
        for a, b, k in ((0, 0, 0), (1, 1, 1), (2, 2, 2), (0, 1, 3), (1, 2, 4), (0, 2, 5)):
            if a < self.dim and b < self.dim:
                pi[k] = np.sum(self.e[a][:, None] * self.e[b][:, None] * fn, axis=0)
            else:
                pi[k] = np.zeros_like(rho)
        x = np.stack([pi[k] / rho for k in range(6)] + [rho], axis=1)
        return rho, u, feq, pi, x

    def smagorinsky(self, rho, pi):
        sd = np.abs(pi[0]) + np.abs(pi[1]) + np.abs(pi[2])
        so = np.abs(pi[3]) + np.abs(pi[4]) + np.abs(pi[5])
        A = 36.0 * self.c_smag ** 2 * (sd + 2.0 * so) / rho
        return 0.5 * (self.tau + np.sqrt(self.tau ** 2 + A))

    def collide(self):
        f = self.f
        rho, u, feq, pi, x = self.features(f)
        if self.every <= 1 or self.record is None or self.calls % self.every == 0:
            xn = ((x.astype(np.float32) - self.mu) * self.isd).astype(np.float64)
            y, _ = pymlp.forward(self.p, xn, quant=self.quant)
            t = self.out_a * y + self.out_b + self.shift
            self.tau_net = t
            bad = ~((t >= self.lo) & (t <= self.hi))
            self.record = np.where(bad, np.nan, t)
        else:
            # between evaluations: the recorded tau, Smagorinsky of the current
            # state where the record fell back
            t = np.nan_to_num(self.record, nan=0.0)
            bad = np.isnan(self.record)
        if self.c_smag > 0:
            t = np.where(bad, self.smagorinsky(rho, pi), t)
        else:
            t = np.where(bad, self.tau, t)
        self.fallback = bad & ~self.solid
        post = f - (f - feq) / t[None, :]
        if self.guo:
            c = 1.0 - 0.5 / self.tau
            eF = self.e.T @ self.F                       # [Q]
            eu = self.e.T @ u                            # [Q, N]
            uF = self.F @ u                              # [N]
            S = c * self.w[:, None] * (3.0 * (eF[:, None] - uF[None, :]) + 9.0 * eu * eF[:, None])
            post = post + S
        sol = self.solid
        if sol.any():
            post[:, sol] = f[self.opp][:, sol]
            rho = np.where(sol, 0.0, rho)
            u = np.where(sol[None, :], 0.0, u)
        self.f = post
        self.rho, self.u = rho, u

    def step(self):
        if self.calls > 0:
            self.stream()
        self.collide()
        self.calls += 1

    def run(self, n):
        for _ in range(n):
            self.step()
        return self.rho, self.u
