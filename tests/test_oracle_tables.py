"""Pins of the oracle's lattice tables (LBM-C/macros.cu:14-125) against what
the lattice Boltzmann method fixes: velocity-set isotropy, opposite
directions and the orthogonal moment bases of Lallemand-Luo (D2Q9) and
d'Humieres et al. (D3Q15)."""
import itertools

import numpy as np
import pytest

import pyoracle as po

LATTICES = [(po.D2Q9, 9, 2), (po.D3Q15, 15, 3)]


@pytest.mark.parametrize("lat,q,dim", LATTICES)
def test_weights_and_isotropy(lat, q, dim):
    e, w, opp, M, Minv = po.tables(lat)
    assert e.shape == (dim, q)
    assert np.isclose(w.sum(), 1.0, rtol=0, atol=1e-15)
    # odd moments vanish, second moment c_s^2 delta with c_s^2 = 1/3
    assert np.allclose(e @ w, 0, atol=1e-16)
    assert np.allclose(np.einsum("i,ai,bi->ab", w, e, e), np.eye(dim) / 3, atol=1e-15)
    assert np.allclose(np.einsum("i,ai,bi,ci->abc", w, e, e, e), 0, atol=1e-16)
    # fourth-order isotropy: sum w e_a e_b e_c e_d = (d_ab d_cd + d_ac d_bd + d_ad d_bc)/9
    d = np.eye(dim)
    want = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
            + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.allclose(np.einsum("i,ai,bi,ci,di->abcd", w, e, e, e, e), want, atol=1e-15)
    # the rest population is first and the velocity set is unique
    assert np.all(e[:, 0] == 0)
    assert len({tuple(c) for c in e.T}) == q


@pytest.mark.parametrize("lat,q,dim", LATTICES)
def test_opposites(lat, q, dim):
    e, w, opp, M, Minv = po.tables(lat)
    assert np.array_equal(e[:, opp], -e)
    assert np.array_equal(opp[opp], np.arange(q))
    assert np.array_equal(w[opp], w)


@pytest.mark.parametrize("lat,q,dim", LATTICES)
def test_moment_basis_rows(lat, q, dim):
    """M has orthogonal rows; rho, j and the energy row are the textbook ones."""
    e, w, opp, M, Minv = po.tables(lat)
    G = M @ M.T
    assert np.allclose(G - np.diag(np.diag(G)), 0)
    assert np.array_equal(M[0], np.ones(q))
    e2 = (e * e).sum(0)
    if lat == po.D2Q9:
        assert np.array_equal(M[1], 3 * e2 - 4)          # energy e
        assert np.array_equal(M[3], e[0]) and np.array_equal(M[5], e[1])
        assert np.array_equal(M[7], e[0] ** 2 - e[1] ** 2)
        assert np.array_equal(M[8], e[0] * e[1])
    else:
        assert np.array_equal(M[1], e2 - 2)               # energy e
        assert np.array_equal(M[3], e[0]) and np.array_equal(M[5], e[1])
        assert np.array_equal(M[7], e[2])
        assert np.array_equal(M[9], 3 * e[0] ** 2 - e2)   # 3 p_xx
        assert np.array_equal(M[11], e[0] * e[1])
        assert np.array_equal(M[12], e[1] * e[2])
        assert np.array_equal(M[13], e[0] * e[2])
        assert np.array_equal(M[14], e[0] * e[1] * e[2])


def test_d2q9_minv_table():
    """The reference's numeric M^-1 (macros.cu:31-43) inverts M to rounding and
    equals M^T D^-1 to within one unit in the last place."""
    e, w, opp, M, Minv = po.tables(po.D2Q9)
    assert np.max(np.abs(M @ Minv - np.eye(9))) < 1e-15
    exact = M.T / np.diag(M @ M.T)
    assert np.max(np.abs(Minv - exact)) < 2e-16


def test_d3q15_minv_reference_is_zero_and_fix_inverts():
    """macros.cu:99-113 writes M^-1 as integer fractions, all 0 in C; the fix
    flag gives the table those fractions spell."""
    _, _, _, M, Minv = po.tables(po.D3Q15)
    assert np.all(Minv == 0)
    # C integer division truncates toward zero: 1/15, -1/9, 2/45 ... are all 0
    assert int(1 / 15) == 0 and int(-1 / 9) == 0 and int(2 / 45) == 0
    _, _, _, M, Minv = po.tables(po.D3Q15, po.FIX_D3Q15_MINV)
    assert np.max(np.abs(M @ Minv - np.eye(15))) < 1e-15
    assert Minv[0, 1] == -1.0 / 9.0 and Minv[0, 2] == 2.0 / 45.0 and Minv[7, 4] == 1.0 / 40.0


@pytest.mark.parametrize("lat,q,dim", LATTICES)
def test_incompressible_equilibrium_moments(lat, q, dim):
    """M f_eq(incompressible, rho0 = 1) reproduces the moment equilibria
    meq_d2q9 / meq_d3q15 (collision.cu:465-503): the oracle's MRT with all
    rates 1 returns M^-1 meq.  (D3Q15 with the -7/3 and M^-1 fixes.)"""
    rng = np.random.default_rng(11)
    quirks = po.FIX_D3Q15_MINV | po.FIX_D3Q15_MEQ
    e, w, opp, M, Minv = po.tables(lat, quirks)
    length = (4, 4) if lat == po.D2Q9 else (4, 4, 4)
    o = po.Oracle(lat, length, po.MRT, tau_mrt=[1.0] * q, quirks=quirks)
    for _ in range(20):
        rho = 1 + 0.05 * rng.standard_normal()
        j = 0.05 * rng.standard_normal(dim)
        eu = e.T @ j
        feq = w * (rho + 3 * eu + 4.5 * eu ** 2 - 1.5 * j @ j)
        f = feq + 1e-3 * rng.standard_normal(q)
        fo, _, _, _ = o.collide(4, f, rho, j / rho)
        assert np.max(np.abs(M @ fo - M @ feq)) < 1e-14


def test_d3q15_reference_q_moment_uses_minus_two():
    """(-7/3) is integer -2 (collision.cu:492-496): the reference's q_x moment is
    -2 j_x.  Checked through the MRT with rate 1 only on the q_x row."""
    e, w, opp, M, _ = po.tables(po.D3Q15)
    _, _, _, _, Minv = po.tables(po.D3Q15, po.FIX_D3Q15_MINV)
    rates = [0.0] * 15
    rates[4] = 1.0
    for quirks, coef in ((po.FIX_D3Q15_MINV, -2.0), (po.FIX_D3Q15_MINV | po.FIX_D3Q15_MEQ, -7.0 / 3.0)):
        o = po.Oracle(po.D3Q15, (4, 4, 4), po.MRT, tau_mrt=rates, quirks=quirks)
        rho, u = 1.0, np.array([0.03, -0.01, 0.02])
        f = w.copy()
        fo, _, _, _ = o.collide(4, f, rho, u)
        m = M @ fo
        assert abs(m[4] - coef * rho * u[0]) < 1e-15
