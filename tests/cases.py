"""Seeded lattice cases shared by the oracle tests and the GPU parity tests.

Each case is a dict: `desc` (constructor keywords common to lbmc.Lattice and
pyoracle.Oracle), `fields` (set_fields keywords), `init` ("static" or
("equilibrium", rho, u)) and `steps`.
"""
from __future__ import annotations

import numpy as np

D2Q9, D3Q15 = 0, 1
BGK, NTPOR, MRT, MRTPOR = 0, 1, 2, 3

# Lallemand-Luo D2Q9 rates with the shear rates from tau
def d2q9_rates(tau):
    s = 1.0 / tau
    return [0.0, 1.64, 1.54, 0.0, 1.9, 0.0, 1.9, s, s]


def d3q15_rates(tau):
    s = 1.0 / tau
    return [0.0, 1.6, 1.2, 0.0, 1.6, 0.0, 1.6, 0.0, 1.6, s, s, s, s, s, 1.2]


def random_velocity(rng, shape, amp):
    return [amp * rng.standard_normal(shape).reshape(-1) for _ in range(len(shape))]


def channel_d2q9(Lx=32, Ly=16, tau=0.8, g=1e-6, steps=200, collision=BGK, rates=None):
    """Body-force channel: solid rows y=0 and y=Ly-1, periodic in x."""
    geo = np.zeros((Ly, Lx))
    geo[0, :] = 1.0
    geo[-1, :] = 1.0
    force = [np.full(Lx * Ly, g), np.zeros(Lx * Ly)]
    return dict(
        name=f"channel_d2q9_{Lx}x{Ly}_c{collision}",
        desc=dict(lattice=D2Q9, length=(Lx, Ly), collision_type=collision, forcing=True,
                  tau=tau, tau_mrt=rates),
        fields=dict(geometry=geo.reshape(-1), force=force),
        init="static", steps=steps)


def pressure_d2q9(Lx=48, Ly=24, tau=0.6, steps=200, seed=1, collision=MRT, porous=False):
    """Zou-He pressure inlet/outlet (macro code 8 + micro 1/2), solid walls,
    a seeded disc obstacle and, optionally, partially saturated nodes."""
    rng = np.random.default_rng(seed)
    geo = np.zeros((Ly, Lx))
    geo[0, :] = 1.0
    geo[-1, :] = 1.0
    yy, xx = np.mgrid[0:Ly, 0:Lx]
    cx, cy, r = Lx // 2, Ly // 2, Ly // 6
    geo[(xx - cx) ** 2 + (yy - cy) ** 2 <= r * r] = 1.0
    if porous:
        ring = ((xx - cx) ** 2 + (yy - cy) ** 2 <= (r + 2) ** 2) & (geo < 1)
        geo[ring] = rng.uniform(0.1, 0.9, size=ring.sum())
    macro = np.zeros((Ly, Lx), dtype=np.int32)
    micro = np.zeros((Ly, Lx), dtype=np.int32)
    macro[1:-1, 0] = 8
    macro[1:-1, -1] = 8
    micro[1:-1, 0] = 1
    micro[1:-1, -1] = 2
    rho = np.ones((Ly, Lx))
    rho[:, 0] = 1.001
    rho[:, -1] = 1.000
    return dict(
        name=f"pressure_d2q9_{Lx}x{Ly}_c{collision}_p{int(porous)}",
        desc=dict(lattice=D2Q9, length=(Lx, Ly), collision_type=collision, macro_bc=True,
                  micro_bc=True, tau=tau, tau_mrt=d2q9_rates(tau)),
        fields=dict(geometry=geo.reshape(-1), macro_bc=macro.reshape(-1),
                    micro_bc=micro.reshape(-1), rho=rho.reshape(-1)),
        init="static", steps=steps)


def porous_d2q9(L=16, tau=0.7, steps=100, seed=2, collision=NTPOR, forcing=True):
    """Partially saturated (NTPOR) nodes with B in (0,1), a few solids, body force."""
    rng = np.random.default_rng(seed)
    geo = np.where(rng.random((L, L)) < 0.3, rng.uniform(0.05, 0.95, (L, L)), 0.0)
    geo[rng.random((L, L)) < 0.05] = 1.0
    n = L * L
    force = [np.full(n, 2e-5), np.full(n, -1e-5)]
    u = random_velocity(rng, (L, L), 0.02)
    return dict(
        name=f"porous_d2q9_{L}_c{collision}_g{int(forcing)}",
        desc=dict(lattice=D2Q9, length=(L, L), collision_type=collision, forcing=forcing,
                  tau=tau, tau_mrt=d2q9_rates(tau)),
        fields=dict(geometry=geo.reshape(-1), force=force if forcing else None),
        init=("equilibrium", np.ones(n) + 0.01 * rng.standard_normal(n), u), steps=steps)


def turbulence_d3q15(L=12, tau=0.51, c_smag=0.17, steps=40, seed=3, amp=0.05, forcing=False,
                     collision=BGK, quirks=0):
    """Periodic box from a seeded random velocity field, equilibrium start."""
    rng = np.random.default_rng(seed)
    n = L ** 3
    u = random_velocity(rng, (L, L, L), amp)
    rho = 1.0 + 0.01 * rng.standard_normal(n)
    force = [np.full(n, 1e-5), np.full(n, 2e-6), np.full(n, -3e-6)] if forcing else None
    return dict(
        name=f"turb_d3q15_{L}_c{collision}_s{c_smag}_g{int(forcing)}_q{quirks}",
        desc=dict(lattice=D3Q15, length=(L, L, L), collision_type=collision, forcing=forcing,
                  tau=tau, tau_mrt=d3q15_rates(tau), c_smag=c_smag, quirks=quirks),
        fields=dict(geometry=np.zeros(n), force=force),
        init=("equilibrium", rho, u), steps=steps)


def faces_d3q15(L=(10, 8, 6), tau=0.7, steps=30, seed=4, collision=BGK, quirks=0,
                with_sf=False):
    """All six Zou-He pressure faces (macro code 8 + micro 1..6) on a box with a
    solid obstacle; optionally sf faces instead on y/z."""
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = L
    n = Lx * Ly * Lz
    geo = np.zeros((Lz, Ly, Lx))
    geo[Lz // 2, Ly // 2, Lx // 2] = 1.0
    macro = np.zeros((Lz, Ly, Lx), dtype=np.int32)
    micro = np.zeros((Lz, Ly, Lx), dtype=np.int32)
    rho = np.ones((Lz, Ly, Lx))
    # faces in code order x, X, y, Y, z, Z (later faces win at edges)
    sel = [(slice(None), slice(None), 0), (slice(None), slice(None), Lx - 1),
           (slice(None), 0, slice(None)), (slice(None), Ly - 1, slice(None)),
           (0, slice(None), slice(None)), (Lz - 1, slice(None), slice(None))]
    for k, s in enumerate(sel):
        macro[s] = 8
        code = k + 1
        if with_sf and k >= 2:
            code = k + 7
        micro[s] = code
        rho[s] = 1.0 + 0.001 * (k + 1)
    u = random_velocity(rng, (Lz, Ly, Lx), 0.01)
    return dict(
        name=f"faces_d3q15_c{collision}_q{quirks}_sf{int(with_sf)}",
        desc=dict(lattice=D3Q15, length=L, collision_type=collision, macro_bc=True,
                  micro_bc=True, tau=tau, tau_mrt=d3q15_rates(tau), quirks=quirks),
        fields=dict(geometry=geo.reshape(-1), macro_bc=macro.reshape(-1),
                    micro_bc=micro.reshape(-1), rho=rho.reshape(-1)),
        init=("equilibrium", np.ones(n), u), steps=steps)


def macro_codes_d2q9(Lx=20, Ly=17, steps=20, seed=5, forcing=True):
    """Every macro code 1..15 on its own row of nodes, prescribed values seeded."""
    rng = np.random.default_rng(seed)
    n = Lx * Ly
    macro = np.zeros((Ly, Lx), dtype=np.int32)
    for code in range(1, 16):
        macro[code, 2:-2] = code
    rho = 1.0 + 0.01 * rng.standard_normal(n)
    u = [0.01 * rng.standard_normal(n), 0.01 * rng.standard_normal(n)]
    force = [np.full(n, 1e-5), np.zeros(n)] if forcing else None
    return dict(
        name=f"macro_codes_d2q9_g{int(forcing)}",
        desc=dict(lattice=D2Q9, length=(Lx, Ly), collision_type=BGK, macro_bc=True,
                  forcing=forcing, tau=0.9),
        fields=dict(geometry=np.zeros(n), macro_bc=macro.reshape(-1), u=u, rho=rho, force=force),
        init="static", steps=steps)


def sf_d2q9(Lx=24, Ly=12, steps=40, seed=6):
    """sf copy conditions on all four D2Q9 faces with a driven interior."""
    rng = np.random.default_rng(seed)
    n = Lx * Ly
    micro = np.zeros((Ly, Lx), dtype=np.int32)
    micro[:, 0] = 5
    micro[:, -1] = 6
    micro[0, 1:-1] = 7
    micro[-1, 1:-1] = 8
    u = random_velocity(rng, (Ly, Lx), 0.02)
    return dict(
        name="sf_d2q9",
        desc=dict(lattice=D2Q9, length=(Lx, Ly), collision_type=BGK, micro_bc=True, tau=0.8),
        fields=dict(geometry=np.zeros(n), micro_bc=micro.reshape(-1)),
        init=("equilibrium", np.ones(n), u), steps=steps)


def nonuniform_force_d3q15(L=(8, 6, 5), steps=20, seed=7):
    rng = np.random.default_rng(seed)
    n = int(np.prod(L))
    force = [1e-5 * rng.standard_normal(n) for _ in range(3)]
    u = random_velocity(rng, L[::-1], 0.02)
    return dict(
        name="force_field_d3q15",
        desc=dict(lattice=D3Q15, length=L, collision_type=BGK, forcing=True, tau=0.75),
        fields=dict(geometry=np.zeros(n), force=force),
        init=("equilibrium", np.ones(n), u), steps=steps)


def porous_media_d2q9(Lx=2048, Ly=1024, tau=0.6, rho_in=1.001, rho_out=1.000, seed=0, steps=0):
    """The porous-medium configuration: seeded discs (radius 8..24) to a solid
    fraction of 0.30 between solid walls, Zou-He pressure inlet (x = 0) and
    outlet (x = Lx-1) with macro code 8, MRT with Lallemand-Luo rates."""
    import workloads

    geo, _ = workloads.porous_discs(Lx, Ly, 8, 24, 0.30, seed=seed)
    macro = np.zeros((Ly, Lx), dtype=np.int32)
    micro = np.zeros((Ly, Lx), dtype=np.int32)
    macro[1:-1, 0] = 8
    macro[1:-1, -1] = 8
    micro[1:-1, 0] = 1
    micro[1:-1, -1] = 2
    rho = np.ones((Ly, Lx))
    rho[:, 0] = rho_in
    rho[:, -1] = rho_out
    return dict(
        name=f"porous_media_d2q9_{Lx}x{Ly}",
        desc=dict(lattice=D2Q9, length=(Lx, Ly), collision_type=MRT, macro_bc=True, micro_bc=True,
                  tau=tau, tau_mrt=d2q9_rates(tau)),
        fields=dict(geometry=geo.reshape(-1), macro_bc=macro.reshape(-1), micro_bc=micro.reshape(-1),
                    rho=rho.reshape(-1)),
        init="static", steps=steps)


def all_cases():
    """The parity matrix: every collision code, BC kind and quirk flag."""
    out = [
        channel_d2q9(),
        channel_d2q9(collision=MRT, rates=d2q9_rates(0.8)),
        pressure_d2q9(),
        pressure_d2q9(collision=BGK),
        pressure_d2q9(collision=MRTPOR, porous=True),
        porous_d2q9(),
        porous_d2q9(forcing=False),
        porous_d2q9(collision=MRTPOR),
        porous_d2q9(collision=MRTPOR, forcing=False),
        turbulence_d3q15(),
        turbulence_d3q15(c_smag=0.0, forcing=True),
        turbulence_d3q15(c_smag=0.17, forcing=True),
        turbulence_d3q15(collision=MRT, c_smag=0.0),
        turbulence_d3q15(collision=MRT, c_smag=0.0, forcing=True, quirks=0x1 | 0x2),
        turbulence_d3q15(collision=MRT, c_smag=0.0, quirks=0x1),
        turbulence_d3q15(collision=NTPOR, c_smag=0.1),
        faces_d3q15(),
        faces_d3q15(quirks=0x4),
        faces_d3q15(with_sf=True),
        faces_d3q15(collision=MRT, quirks=0x1 | 0x2 | 0x4),
        macro_codes_d2q9(),
        macro_codes_d2q9(forcing=False),
        sf_d2q9(),
        nonuniform_force_d3q15(),
    ]
    return out


def build(mod_cls, case, **extra):
    """Instantiate a lattice (lbmc.Lattice or pyoracle.Oracle) for a case."""
    lat = mod_cls(**case["desc"], **extra)
    lat.set_fields(**case["fields"])
    init = case["init"]
    if init == "static":
        lat.init_static()
    else:
        _, rho, u = init
        lat.init_equilibrium(rho, u)
    return lat
