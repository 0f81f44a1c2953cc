"""Lattice-level pins of the oracle's step (LBM-C/solver.cu:10-136):
index-shift streaming moves population i by e_i per call, periodic BGK
conserves mass and momentum, Guo forcing adds F per node per step, shear
waves decay at nu = (tau - 1/2)/3, and the steady channel profile has the
curvature -g/nu; node masks and link sets agree with brute force."""
import numpy as np
import pytest

import cases
import pyoracle as po


def identity_collision(lat, L):
    """D3Q15 MRT with the reference's all-zero M^-1 leaves f unchanged:
    the step is pure streaming.  (D2Q9: MRT with every rate 0.)"""
    q = 9 if lat == po.D2Q9 else 15
    return po.Oracle(lat, L, po.MRT, tau_mrt=[0.0] * q)


@pytest.mark.parametrize("lat,L", [(po.D2Q9, (5, 4)), (po.D3Q15, (5, 4, 3))])
def test_streaming_moves_population_by_e(lat, L):
    """A marker written into node x0's direction-k slot at t=0 is pulled by
    node x0 + t e_k at call t (periodic), and by no other node."""
    e, w, *_ = po.tables(lat)
    dim = len(L)
    n = int(np.prod(L))
    for k in range(1, e.shape[1]):
        o = identity_collision(lat, L)
        o.set_fields(geometry=np.zeros(n))
        o.init_static()
        f = o.get_f()
        x0 = np.array([1, 2, 1][:dim])
        s0 = np.ravel_multi_index(x0[::-1], L[::-1])
        f[k, s0] += 1e-3
        o.set_f(f)
        for t in range(7):
            o.step(t, True)
            rho, _ = o.get_macros()
            hit = np.nonzero(np.abs(rho - 1.0) > 1e-9)[0]
            want = (x0 + t * e[:, k]) % np.array(L)
            assert list(hit) == [np.ravel_multi_index(want[::-1], L[::-1])], (k, t)


@pytest.mark.parametrize("lat,L", [(po.D2Q9, (7, 5)), (po.D3Q15, (6, 5, 4))])
def test_slot_map_is_a_bijection(lat, L):
    """Every slot is read by exactly one node per call: with identity collision
    a step leaves every slot value unchanged (each node writes back the slots
    it read) and the stored densities sum to the population total."""
    n = int(np.prod(L))
    o = identity_collision(lat, L)
    o.set_fields(geometry=np.zeros(n))
    rng = np.random.default_rng(5)
    f = rng.random((o.q, n))
    o.set_f(f)
    for t in (0, 1, 2, 5, 13, 10 ** 6 + 3):
        o.step(t, True)
        assert np.array_equal(o.get_f(), f)
        rho, _ = o.get_macros()
        assert abs(rho.sum() - f.sum()) < 1e-11


@pytest.mark.parametrize("case", [cases.turbulence_d3q15(L=8, c_smag=0.0, steps=30),
                                  cases.turbulence_d3q15(L=8, c_smag=0.17, steps=30),
                                  cases.turbulence_d3q15(L=8, collision=po.MRT, c_smag=0.0,
                                                         quirks=3, steps=30)],
                         ids=["bgk", "smag", "mrt"])
def test_periodic_conservation(case):
    o = cases.build(po.Oracle, case)
    e, *_ = po.tables(o.lattice)
    f0 = o.get_f()
    m0, j0 = f0.sum(), e @ f0.sum(1)
    o.run(0, case["steps"], 0)
    f1 = o.get_f()
    assert abs(f1.sum() - m0) < 1e-11 * m0
    assert np.allclose(e @ f1.sum(1), j0, atol=1e-11)


def test_uniform_force_momentum_growth():
    """Guo forcing: total momentum grows by exactly N F per call; the stored u
    includes the half-force shift."""
    L = (6, 5, 4)
    n = int(np.prod(L))
    F = np.array([1e-5, -2e-6, 3e-6])
    o = po.Oracle(po.D3Q15, L, po.BGK, forcing=True, tau=0.7)
    o.set_fields(geometry=np.zeros(n), force=[np.full(n, v) for v in F])
    o.init_static()
    e, *_ = po.tables(po.D3Q15)
    for t in range(10):
        o.step(t, True)
        j = e @ o.get_f().sum(1)
        assert np.allclose(j, (t + 1) * n * F, rtol=1e-9, atol=1e-16)
    _, u = o.get_macros()
    assert np.allclose(u.mean(1), 9.5 * F, rtol=1e-9)


@pytest.mark.parametrize("lat,coll,tau", [(po.D2Q9, po.BGK, 0.8), (po.D2Q9, po.MRT, 0.8),
                                          (po.D3Q15, po.BGK, 0.6)])
def test_shear_wave_decay(lat, coll, tau):
    """u_x = A sin(2 pi y / Ly) decays as exp(-nu k^2 t), nu = (tau - 1/2)/3."""
    Ly = 64
    L = (4, Ly) if lat == po.D2Q9 else (4, Ly, 4)
    n = int(np.prod(L))
    y = (np.arange(n) // L[0]) % Ly
    k = 2 * np.pi / Ly
    A = 1e-3
    rates = cases.d2q9_rates(tau) if lat == po.D2Q9 else None
    o = po.Oracle(lat, L, coll, tau=tau, tau_mrt=rates)
    o.set_fields(geometry=np.zeros(n))
    u = [A * np.sin(k * y), np.zeros(n)] + ([np.zeros(n)] if lat == po.D3Q15 else [])
    o.init_equilibrium(np.ones(n), u)
    T = 400
    o.run(0, T, T)
    _, uu = o.get_macros()
    amp = 2 * np.mean(uu[0] * np.sin(k * y))
    nu_meas = -np.log(amp / A) / (k * k * T)
    nu = (tau - 0.5) / 3
    assert abs(nu_meas / nu - 1) < 0.01


def test_channel_profile_curvature():
    """Steady body-force channel with full-way bounceback walls: the bulk
    profile's second difference is -g/nu (nu = (tau-1/2)/3), the profile is
    symmetric, and its maximum is within 2% of the parabola g H^2/(8 nu) with
    the walls halfway between fluid and solid rows."""
    Lx, Ly, tau, g = 4, 18, 0.8, 1e-6
    case = cases.channel_d2q9(Lx=Lx, Ly=Ly, tau=tau, g=g)
    o = cases.build(po.Oracle, case)
    o.run(0, 12000, 12000)
    _, u = o.get_macros()
    ux = u[0].reshape(Ly, Lx)[:, 0]
    nu = (tau - 0.5) / 3
    d2 = ux[2:-1] - 2 * ux[1:-2] + ux[:-3]
    assert np.allclose(d2[1:-1], -g / nu, rtol=1e-4)
    assert np.allclose(ux[1:-1], ux[1:-1][::-1], rtol=1e-10)
    H = Ly - 2
    umax = g * H * H / (8 * nu)
    assert abs(ux.max() / umax - 1) < 0.02


def brute_classes(case):
    geo = case["fields"]["geometry"]
    n = geo.size
    ct = case["desc"]["collision_type"]
    macro = case["fields"].get("macro_bc")
    micro = case["fields"].get("micro_bc")
    macro = np.zeros(n, int) if macro is None or not case["desc"].get("macro_bc") else macro
    micro = np.zeros(n, int) if micro is None or not case["desc"].get("micro_bc") else micro
    desc = (macro.astype(np.uint16) << 4) | (micro.astype(np.uint16) << 8)
    desc = np.where((ct in (1, 3)) & (geo != 0), desc | 2, desc)
    desc = np.where(geo >= 1, 1, desc).astype(np.uint16)
    uniq = np.unique(desc[desc != 0])
    table = np.concatenate([[0], uniq]).astype(np.uint16)
    mask = np.searchsorted(table, desc).astype(np.uint8)
    return mask, table


@pytest.mark.parametrize("case", cases.all_cases(), ids=[c["name"] for c in cases.all_cases()])
def test_node_classes_and_links_brute_force(case):
    o = cases.build(po.Oracle, case)
    mask, table = o.node_types()
    bm, bt = brute_classes(case)
    assert np.array_equal(table, bt)
    assert np.array_equal(mask, bm)
    # links: bit i when the periodic neighbour x + e_i is solid and x is not
    e, *_ = po.tables(o.lattice)
    L = o.length
    solid = (case["fields"]["geometry"] >= 1).reshape(L[::-1])
    links = np.zeros(solid.shape, dtype=np.uint16)
    for i in range(1, e.shape[1]):
        shift = tuple(-int(c) for c in e[::-1, i])
        nb = np.roll(solid, shift, axis=tuple(range(len(L))))
        links |= np.where(nb & ~solid, 1 << i, 0).astype(np.uint16)
    assert np.array_equal(o.links(), links.reshape(-1))


def test_kinetic_energy_sum():
    case = cases.turbulence_d3q15(L=8, steps=3)
    o = cases.build(po.Oracle, case)
    o.run(0, 3, 1)
    rho, u = o.get_macros()
    assert abs(o.kinetic_energy() - np.sum(0.5 * rho * (u * u).sum(0))) < 1e-13
