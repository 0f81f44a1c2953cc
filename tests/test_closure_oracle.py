"""CPU pins of the closure-step oracle (oracle/pyclosure.py) against the C
oracle: with every node outside the band it must reduce to BGK (c_smag = 0)
or BGK + Smagorinsky, with and without Guo forcing and solids; a network
with constant output must act as BGK with that tau."""
import numpy as np
import pytest

import pyclosure
import pymlp
import pyoracle


def _turbulent_d3q15(L=(8, 6, 5), amp=0.04, seed=0):
    rng = np.random.default_rng(seed)
    n = int(np.prod(L))
    rho = 1.0 + 0.01 * rng.standard_normal(n)
    u = amp * rng.standard_normal((3, n))
    return rho, u


def _never_in_band_params():
    # output far below the band: every fluid node falls back
    p = pymlp.init_params(0)
    p["b3"] = -100.0
    return p


def _run_both(lattice, L, tau, c_smag, params, out_a, out_b, lo, hi, steps, force=None, solid=None,
              rho=None, u=None):
    dim = 2 if lattice == pyoracle.D2Q9 else 3
    n = int(np.prod(L))
    co = pyclosure.ClosureOracle(lattice, L, tau, c_smag, params, np.zeros(7), np.ones(7), out_a, out_b,
                                 lo, hi, force=force, solid=solid)
    co.init_equilibrium(rho, u)
    rc, uc = co.run(steps)
    o = pyoracle.Oracle(lattice, L, pyoracle.BGK, forcing=force is not None, tau=tau, c_smag=c_smag)
    fields = {}
    if solid is not None:
        fields["geometry"] = np.asarray(solid, dtype=np.float64)
    if force is not None:
        fields["force"] = [np.full(n, force[d]) for d in range(dim)]
    o.set_fields(**fields)
    o.init_equilibrium(rho, u)
    o.run(0, steps, 1)
    ro, uo = o.get_macros()
    return (rc, uc), (ro, uo), co


@pytest.mark.parametrize("c_smag", [0.0, 0.17])
def test_fallback_reduces_to_reference_bgk_d3q15(c_smag):
    L = (8, 6, 5)
    rho, u = _turbulent_d3q15(L)
    (rc, uc), (ro, uo), co = _run_both(pyoracle.D3Q15, L, 0.51, c_smag, _never_in_band_params(), 1.0, 0.0,
                                       0.6, 2.0, 6, rho=rho, u=u)
    assert co.fallback.all()
    assert np.max(np.abs(rc - ro)) < 1e-13
    assert np.max(np.abs(uc - uo)) < 1e-13


def test_fallback_guo_and_solids_d2q9():
    Lx, Ly = 12, 9
    n = Lx * Ly
    solid = np.zeros((Ly, Lx))
    solid[0, :] = 1
    solid[-1, :] = 1
    solid[4, 5] = 1
    rng = np.random.default_rng(2)
    rho = 1.0 + 0.01 * rng.standard_normal(n)
    u = 0.03 * rng.standard_normal((2, n))
    (rc, uc), (ro, uo), co = _run_both(pyoracle.D2Q9, (Lx, Ly), 0.7, 0.12, _never_in_band_params(), 1.0, 0.0,
                                       0.6, 2.0, 9, force=[1e-5, -2e-6], solid=solid.reshape(-1), rho=rho, u=u)
    assert np.max(np.abs(rc - ro)) < 1e-13
    assert np.max(np.abs(uc - uo)) < 1e-13


def test_constant_network_is_bgk_with_its_tau():
    # w3 = 0: y = b3 everywhere; tau' = out_a * b3 + out_b = 0.73 inside the band
    p = pymlp.init_params(1)
    p["w3"][:] = 0.0
    p["b3"] = 0.5
    L = (6, 5, 4)
    rho, u = _turbulent_d3q15(L, seed=3)
    co = pyclosure.ClosureOracle(pyoracle.D3Q15, L, 0.51, 0.17, p, np.zeros(7), np.ones(7), 0.2, 0.63,
                                 0.6, 2.0)
    co.init_equilibrium(rho, u)
    rc, uc = co.run(5)
    assert not co.fallback.any()
    o = pyoracle.Oracle(pyoracle.D3Q15, L, pyoracle.BGK, tau=0.73)
    o.set_fields()
    o.init_equilibrium(rho, u)
    o.run(0, 5, 1)
    ro, uo = o.get_macros()
    assert np.max(np.abs(rc - ro)) < 1e-13
    assert np.max(np.abs(uc - uo)) < 1e-13


def test_features_are_strain_rate_for_bgk_shear_wave():
    # pre-collision Pi = -2 rho c_s^2 tau S (Chapman-Enskog), so for a slowly
    # varying shear wave Pi_xy / rho = -(tau / 3) du_x/dy
    Lx, Ly = 4, 64
    y = np.arange(Ly)
    ux = np.tile((1e-3 * np.sin(2 * np.pi * y / Ly))[:, None], (1, Lx)).reshape(-1)
    n = Lx * Ly
    tau = 0.8
    co = pyclosure.ClosureOracle(pyoracle.D2Q9, (Lx, Ly), tau, 0.0, _never_in_band_params(), np.zeros(7),
                                 np.ones(7), 1.0, 0.0, 0.6, 2.0)
    co.init_equilibrium(np.ones(n), np.stack([ux, np.zeros(n)]))
    co.run(200)
    co.stream()
    _, u, _, _, x = co.features(co.f)
    uxf = u[0].reshape(Ly, Lx)[:, 0]
    dudy = (np.roll(uxf, -1) - np.roll(uxf, 1)) / 2.0
    pxy = x[:, 3].reshape(Ly, Lx)[:, 0]
    assert np.max(np.abs(pxy + (tau / 3.0) * dudy)) < 0.02 * np.max(np.abs(pxy))
