"""Seeded synthetic inputs for the benchmark configurations.

  decaying_turbulence  periodic box, solenoidal random-phase velocity with
                       E(k) ~ k^4 exp(-2 (k/k0)^2), scaled to a per-component
                       rms velocity (the D3Q15 turbulence workloads)
  channel              body-force channel with solid walls at y=0 and y=Ly-1
  porous_discs         non-overlapping discs with radii in [rmin, rmax] placed
                       until a solid fraction is reached, walls top/bottom,
                       inlet/outlet columns kept clear (the D2Q9 porous case)
  ensemble_d2q9        seeded ensemble of uneven D2Q9 channels with 1-6 discs
                       each, Reynolds number set through tau and the body
                       force (the batched-ensemble case)
  ensemble_start       a member's developed start: the plane Poiseuille profile
                       at its mean speed, zero in the solids
  ensemble_closure     the ensemble's shared closure network: trained on the
                       ensemble's filtered refined runs (data/ensemble_closure.npz),
                       or untrained with a scaling that keeps its tau' near
                       each member's tau

All generators take an explicit seed and use numpy's PCG64 generator, so the
fields (and the node masks built from them) are reproducible bit for bit.
"""
from __future__ import annotations

import os

import numpy as np


def decaying_turbulence(L, k0=4.0, u_rms=0.05, seed=0):
    """Returns (ux, uy, uz) float64 arrays of shape (L, L, L) (z, y, x order),
    divergence-free in the spectral sense, with per-component rms u_rms."""
    rng = np.random.default_rng(seed)
    shape = (L, L, L)
    k1 = np.fft.fftfreq(L, d=1.0 / L)
    kr = np.fft.rfftfreq(L, d=1.0 / L)
    kz, ky, kx = np.meshgrid(k1, k1, kr, indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    k = np.sqrt(k2)
    # random phases: transform of white noise in each component
    comps = [np.fft.rfftn(rng.standard_normal(shape)) for _ in range(3)]
    # solenoidal projection u_hat - k (k . u_hat)/|k|^2
    with np.errstate(invalid="ignore", divide="ignore"):
        dot = (kx * comps[0] + ky * comps[1] + kz * comps[2]) / np.where(k2 > 0, k2, 1.0)
    comps = [comps[0] - kx * dot, comps[1] - ky * dot, comps[2] - kz * dot]
    # shell spectrum E(k) ~ k^4 exp(-2(k/k0)^2): amplitude ~ sqrt(E(k)/k^2)
    with np.errstate(invalid="ignore", divide="ignore"):
        amp = np.where(k > 0, np.sqrt(k ** 4 * np.exp(-2.0 * (k / k0) ** 2) / np.where(k2 > 0, k2, 1.0)), 0.0)
    u = [np.fft.irfftn(c * amp, s=shape, axes=(0, 1, 2)) for c in comps]
    scale = u_rms / np.sqrt(np.mean(u[0] ** 2 + u[1] ** 2 + u[2] ** 2) / 3.0)
    return tuple(np.ascontiguousarray(v * scale) for v in u)


def energy_spectrum(ux, uy, uz):
    """Shell-averaged E(k) of a periodic (L, L, L) field; k = 0..L/2."""
    L = ux.shape[0]
    k1 = np.fft.fftfreq(L, d=1.0 / L)
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.rint(np.sqrt(kx * kx + ky * ky + kz * kz)).astype(int)
    e = np.zeros(L // 2 + 1)
    for c in (ux, uy, uz):
        p = np.abs(np.fft.fftn(c) / c.size) ** 2
        e += np.bincount(kk.reshape(-1), weights=0.5 * p.reshape(-1), minlength=kk.max() + 1)[: L // 2 + 1]
    return e


def channel(Lx, Ly):
    """Porosity field of a channel: solid rows y=0 and y=Ly-1."""
    geo = np.zeros((Ly, Lx))
    geo[0, :] = 1.0
    geo[-1, :] = 1.0
    return geo


def porous_discs(Lx, Ly, rmin=8, rmax=24, solid_fraction=0.30, seed=0, margin=None,
                 max_tries=200000):
    """Porosity field with seeded non-overlapping discs (radius uniform in
    [rmin, rmax]) added until the solid fraction (walls included) reaches the
    target.  Discs stay `margin` columns (default rmax) clear of the inlet and
    outlet and one node clear of each other and of the walls."""
    rng = np.random.default_rng(seed)
    margin = rmax if margin is None else margin
    geo = channel(Lx, Ly)
    discs = []
    yy, xx = np.ogrid[0:Ly, 0:Lx]
    tries = 0
    while geo.mean() < solid_fraction and tries < max_tries:
        tries += 1
        r = rng.uniform(rmin, rmax)
        cx = rng.uniform(margin + r, Lx - 1 - margin - r)
        cy = rng.uniform(1 + r + 1, Ly - 2 - r - 1)
        if any((cx - x) ** 2 + (cy - y) ** 2 < (r + s + 1) ** 2 for x, y, s in discs):
            continue
        discs.append((cx, cy, r))
        x0, x1 = int(cx - r - 1), int(cx + r + 2)
        y0, y1 = int(cy - r - 1), int(cy + r + 2)
        sub = (xx[:, x0:x1] - cx) ** 2 + (yy[y0:y1, :] - cy) ** 2 <= r * r
        geo[y0:y1, x0:x1][sub] = 1.0
    return geo, discs


def ensemble_d2q9(count=64, seed=0, nx=(128, 1024), ny=(64, 512), re=(10.0, 1000.0), u_mean=0.04,
                  tau_min=0.505):
    """Member specifications of a batched ensemble: periodic-in-x channels
    (solid rows y = 0 and y = ny-1) of seeded size (nx and ny uniform over
    the integers of their ranges), with 1-6 seeded
    non-overlapping discs of radius in [ny/16, ny/8], driven by a uniform
    body force.  Re = u_mean D / nu with D the mean disc diameter, log-uniform
    in `re`; nu = u_mean D / Re gives tau = 1/2 + 3 nu (raised to tau_min, in
    which case u_mean drops to keep Re), and the force follows the plane
    Poiseuille estimate F = 12 nu u_mean / H^2.  Even members use MRT with
    the shear and momentum rates 1/tau (the consistent Guo forcing), odd
    members BGK."""
    rng = np.random.default_rng(seed)
    members = []
    for k in range(count):
        Lx = int(rng.integers(nx[0], nx[1] + 1))
        Ly = int(rng.integers(ny[0], ny[1] + 1))
        geo = channel(Lx, Ly)
        yy, xx = np.ogrid[0:Ly, 0:Lx]
        ndisc = int(rng.integers(1, 7))
        discs = []
        tries = 0
        while len(discs) < ndisc and tries < 10000:
            tries += 1
            r = rng.uniform(Ly / 16.0, Ly / 8.0)
            cx = rng.uniform(r + 1, Lx - r - 2)
            cy = rng.uniform(r + 3, Ly - r - 4)
            if any((cx - x) ** 2 + (cy - y) ** 2 < (r + s + 2) ** 2 for x, y, s in discs):
                continue
            discs.append((cx, cy, r))
            geo[(xx - cx) ** 2 + (yy - cy) ** 2 <= r * r] = 1.0
        D = 2.0 * float(np.mean([d[2] for d in discs]))
        Re = float(np.exp(rng.uniform(np.log(re[0]), np.log(re[1]))))
        nu = u_mean * D / Re
        tau = 0.5 + 3.0 * nu
        u = u_mean
        if tau < tau_min:
            tau = tau_min
            nu = (tau - 0.5) / 3.0
            u = Re * nu / D
        H = Ly - 2
        F = 12.0 * nu * u / (H * H)
        mrt = k % 2 == 0
        s = 1.0 / tau
        members.append(dict(
            name=f"member{k}", length=(Lx, Ly), geometry=geo, discs=discs, Re=Re, tau=tau, force=F, u_mean=u,
            collision="MRT" if mrt else "BGK",
            rates=[0.0, 1.64, 1.54, s, 1.9, s, 1.9, s, s] if mrt else None))
    return members


def ensemble_start(m, noise=0.0, rng=None):
    """Initial (rho, [ux, uy]) of an ensemble member: the plane Poiseuille
    profile u_x = 6 U eta (H - eta) / H^2 (mean U = the member's u_mean, walls
    half a node outside the fluid rows, H = ny - 2) in the fluid, zero in the
    solids, rho = 1; plus Gaussian noise of the given size in the fluid."""
    Lx, Ly = m["length"]
    H = Ly - 2
    eta = np.arange(Ly) - 0.5
    prof = np.clip(6.0 * m["u_mean"] * eta * (H - eta) / (H * H), 0.0, None)
    ux = np.broadcast_to(prof[:, None], (Ly, Lx)).copy()
    uy = np.zeros((Ly, Lx))
    solid = m["geometry"] >= 1.0
    ux[solid] = 0.0
    if noise > 0.0:
        ux += noise * rng.standard_normal((Ly, Lx)) * ~solid
        uy += noise * rng.standard_normal((Ly, Lx)) * ~solid
    return np.ones(Lx * Ly), [ux.reshape(-1), uy.reshape(-1)]


# the trained ensemble network (lbm-c_amd/ensemble_training.py writes it)
ENSEMBLE_CLOSURE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "ensemble_closure.npz")


def ensemble_closure(trained=True, seed=0):
    """The ensemble's shared closure network: (params, mu, isd, out_a, out_b,
    tau_ref); a member of relaxation time tau takes
    lbmc_set_closure_shift(tau - tau_ref), so tau' = tau - tau_ref + out_b
    + out_a y.

    trained=True: the network trained a priori on box-filtered runs of the
    ensemble's own channels refined by 2 (ensemble_training.py; it predicts
    the eddy-viscosity correction tau' - tau, tau_ref = 0).  trained=False:
    He-initialised parameters, inputs Pi_ab / rho scaled by 1e3 and rho
    around 1 by 100, tau' = 0.55 + 1e-3 y before the shift (tau_ref = 0.55):
    tau' stays within a few 1e-3 of the member's tau; it stands for the cost
    of a network, not for its physics."""
    if trained:
        d = np.load(ENSEMBLE_CLOSURE_FILE)
        return (d["params"].astype(np.float32), d["mu"].astype(np.float32), d["isd"].astype(np.float32),
                float(d["out_a"]), float(d["out_b"]), 0.0)
    import closure

    p = closure.init_params(seed)
    mu = np.array([0, 0, 0, 0, 0, 0, 1.0], dtype=np.float32)
    isd = np.array([1e3] * 6 + [100.0], dtype=np.float32)
    return p, mu, isd, 1e-3, 0.55, 0.55
