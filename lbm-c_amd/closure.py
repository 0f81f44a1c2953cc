"""Host side of the learned relaxation-time closure: parameter layout and
initialisation, and the a-priori training pipeline from a resolved run.

Pipeline (decaying turbulence, coarse grid = fine grid / r per axis):
  1. a resolved fine run (BGK, tau_f) provides rho and u at chosen steps;
  2. box-filtering by r^3 gives the coarse fields ubar and the subgrid
     stress tau_ab = bar(u_a u_b) - ubar_a ubar_b; the filtered strain rate
     Sbar comes from central differences on the coarse grid;
  3. the local eddy viscosity that best matches the subgrid stress,
     nu_t = max(0, -tau^dev : Sbar / (2 Sbar : Sbar)), gives the target
     relaxation time tau' = tau_c + 3 nu_t (coarse lattice units; acoustic
     scaling keeps the lattice velocity, so nu_c = nu_f / r and
     tau_c = 1/2 + (tau_f - 1/2) / r);
  4. the network's inputs are what the coarse lattice sees before a
     collision: Pi_ab / rho = -(2/3) tau' Sbar_ab (Chapman-Enskog, pre-
     collision non-equilibrium flux with c_s^2 = 1/3), and rho.
The network is trained on (inputs, (tau' - out_b) / out_a) with the
library's own training kernels (lbmc.Mlp.train).
"""
from __future__ import annotations

import numpy as np

NPARAM = 4737
W1, W2, B2, W3, B3 = 0, 512, 4608, 4672, 4736


def init_params(seed=0):
    """He-initialised flat parameters (w1[64][8] with zero bias column, w2, b2 = 0, w3, b3 = 0)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    w1 = rng.standard_normal((64, 8)) * np.sqrt(2.0 / 7.0)
    w1[:, 7] = 0.0
    p[W1:W2] = w1.reshape(-1)
    p[W2:B2] = (rng.standard_normal((64, 64)) * np.sqrt(2.0 / 64.0)).reshape(-1)
    p[W3:B3] = rng.standard_normal(64) * np.sqrt(1.0 / 64.0)
    return p


def box_filter(a, r):
    """Mean over r^3 blocks of a (Z, Y, X) array."""
    Z, Y, X = a.shape
    return a.reshape(Z // r, r, Y // r, r, X // r, r).mean(axis=(1, 3, 5))


def strain_rate(u):
    """Sbar_ab = (d_a u_b + d_b u_a) / 2 by periodic central differences;
    u = (ux, uy, uz) of shape (Z, Y, X).  Returned packed xx, yy, zz, xy, yz, xz."""
    axis = {0: 2, 1: 1, 2: 0}

    def d(f, a):
        return 0.5 * (np.roll(f, -1, axis=axis[a]) - np.roll(f, 1, axis=axis[a]))

    g = [[d(u[b], a) for b in range(3)] for a in range(3)]  # g[a][b] = d_a u_b
    return np.stack([g[0][0], g[1][1], g[2][2],
                     0.5 * (g[0][1] + g[1][0]), 0.5 * (g[1][2] + g[2][1]), 0.5 * (g[0][2] + g[2][0])])


def sgs_targets(rho, u, r, tau_c):
    """From fine fields rho (Z, Y, X) and u (3, Z, Y, X): coarse rho, coarse
    u, Sbar (6, ...), the subgrid stress (6, ...) and the target tau'."""
    rb = box_filter(rho, r)
    ub = np.stack([box_filter(u[a], r) for a in range(3)])
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2))
    tsgs = np.stack([box_filter(u[a] * u[b], r) - ub[a] * ub[b] for a, b in pairs])
    S = strain_rate(ub)
    return rb, ub, S, tsgs, tau_c + 3.0 * eddy_viscosity(tsgs, S)


def eddy_viscosity(tsgs, S):
    """Pointwise least-squares nu_t >= 0 with tau^dev ~ -2 nu_t Sbar (packed
    symmetric tensors: xx, yy, zz, xy, yz, xz)."""
    tr = (tsgs[0] + tsgs[1] + tsgs[2]) / 3.0
    dev = tsgs.copy()
    dev[0:3] -= tr
    w = np.array([1, 1, 1, 2, 2, 2], dtype=np.float64).reshape((6,) + (1,) * (S.ndim - 1))
    num = -np.sum(w * dev * S, axis=0)
    den = 2.0 * np.sum(w * S * S, axis=0)
    return np.maximum(0.0, num / np.maximum(den, 1e-300))


def features(rho, S, tau_t):
    """Closure inputs per coarse node: Pi_ab / rho = -(2/3) tau' Sbar_ab, then rho."""
    f = np.empty(rho.shape + (7,), dtype=np.float32)
    for k in range(6):
        f[..., k] = -(2.0 / 3.0) * tau_t * S[k]
    f[..., 6] = rho
    return f.reshape(-1, 7)


def fit_scaling(feat, target):
    """Input normalisation (mu, isd) and output map (out_a, out_b) from data."""
    mu = feat.mean(axis=0).astype(np.float32)
    sd = feat.std(axis=0)
    isd = (1.0 / np.where(sd > 0, sd, 1.0)).astype(np.float32)
    out_b = float(np.mean(target))
    out_a = float(np.std(target)) or 1.0
    return mu, isd, out_a, out_b
