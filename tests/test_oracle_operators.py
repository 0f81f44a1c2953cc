"""Pins of the oracle's single-node operators (LBM-C/collision.cu,
LBM-C/boundary_conditions/*) by properties the method fixes: equilibrium
moments, Guo-source moments, BGK relaxation of the non-equilibrium part,
the Smagorinsky closure's defining quadratic, Zou-He closure, the bounceback
permutation, the partially-saturated blend's limits and the macro-BC table."""
import numpy as np
import pytest

import pyoracle as po

RNG = np.random.default_rng(1234)
LATTICES = [(po.D2Q9, 9, 2, (4, 4)), (po.D3Q15, 15, 3, (4, 4, 4))]


def tables(lat, quirks=0):
    return po.tables(lat, quirks)


def feq_numpy(lat, rho, u):
    """Second-order equilibrium w_i rho (1 + 3 e.u + 9/2 (e.u)^2 - 3/2 u^2)."""
    e, w, *_ = tables(lat)
    eu = e.T @ u
    return w * rho * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * u @ u)


@pytest.mark.parametrize("lat,q,dim,L", LATTICES)
def test_equilibrium_moments(lat, q, dim, L):
    e, w, *_ = tables(lat)
    o = po.Oracle(lat, L)
    for _ in range(20):
        rho = 1 + 0.1 * RNG.standard_normal()
        u = 0.1 * RNG.standard_normal(dim)
        feq = o.equilibrium(rho, u)
        assert abs(feq.sum() - rho) < 1e-14
        assert np.allclose(e @ feq, rho * u, atol=1e-15)
        # second moment of the quadratic equilibrium: rho (u u + I/3)
        P = np.einsum("i,ai,bi->ab", feq, e, e)
        assert np.allclose(P, rho * (np.outer(u, u) + np.eye(dim) / 3), atol=1e-15)


@pytest.mark.parametrize("lat,q,dim,L", LATTICES)
def test_bgk_relaxes_nonequilibrium(lat, q, dim, L):
    """Code 0: f' - feq = (1 - 1/tau)(f - feq) for any f with the node's moments."""
    e, w, *_ = tables(lat)
    tau = 0.8
    o = po.Oracle(lat, L, tau=tau)
    for _ in range(10):
        rho = 1 + 0.05 * RNG.standard_normal()
        u = 0.05 * RNG.standard_normal(dim)
        feq = feq_numpy(lat, rho, u)
        d = 1e-3 * RNG.standard_normal(q)
        # remove mass and momentum from the perturbation
        B = np.vstack([np.ones(q), e])
        d -= B.T @ np.linalg.solve(B @ B.T, B @ d)
        fo, r2, u2, t2 = o.collide(0, feq + d, rho, u)
        assert np.allclose(fo - feq, (1 - 1 / tau) * d, atol=1e-15)
        assert t2 == tau and r2 == rho and np.array_equal(u2, u)


@pytest.mark.parametrize("lat,q,dim,L", LATTICES)
def test_guo_source_moments(lat, q, dim, L):
    """Code 1 at f = feq(u + F/2rho): the update is the Guo source S alone, with
    sum S = 0, sum e S = (1 - 1/2tau) F, sum e e S = (1 - 1/2tau)(u* F + F u*)
    where u* is the half-force-shifted velocity, which is also what the node
    stores."""
    e, w, *_ = tables(lat)
    tau = 0.7
    o = po.Oracle(lat, L, forcing=True, tau=tau)
    for _ in range(10):
        rho = 1 + 0.05 * RNG.standard_normal()
        u = 0.05 * RNG.standard_normal(dim)
        F = 1e-3 * RNG.standard_normal(dim)
        ustar = u + F / (2 * rho)
        f = feq_numpy(lat, rho, ustar)
        fo, r2, u2, _ = o.collide(1, f, rho, u, F=F)
        S = fo - f
        c = 1 - 1 / (2 * tau)
        assert abs(S.sum()) < 1e-16
        assert np.allclose(e @ S, c * F, atol=1e-16)
        P = np.einsum("i,ai,bi->ab", S, e, e)
        assert np.allclose(P, c * (np.outer(ustar, F) + np.outer(F, ustar)), atol=1e-16)
        assert np.allclose(u2, ustar, atol=1e-17)


@pytest.mark.parametrize("lat,q,dim,L", LATTICES)
def test_smagorinsky_closure(lat, q, dim, L):
    """tau' solves tau'(tau' - tau) = A/4 with A = 18 sqrt(2) C^2 Qhat / rho and
    Qhat = sum_ab sqrt(2 Pi_ab^2), Pi = sum e e (f - feq) (collision.cu:445-463);
    C = 0 and f = feq leave tau unchanged."""
    e, w, *_ = tables(lat)
    tau, C = 0.51, 0.17
    o = po.Oracle(lat, L, tau=tau, c_smag=C)
    o0 = po.Oracle(lat, L, tau=tau, c_smag=0.0)
    for _ in range(10):
        rho = 1 + 0.05 * RNG.standard_normal()
        u = 0.05 * RNG.standard_normal(dim)
        feq = feq_numpy(lat, rho, u)
        f = feq + 1e-3 * RNG.standard_normal(q)
        fo, _, _, t2 = o.collide(0, f, rho, u)
        Pi = np.einsum("ai,bi,i->ab", e, e, f - feq)
        qhat = np.sqrt(2 * Pi ** 2).sum()
        A = 18 * np.sqrt(2) * C ** 2 * qhat / rho
        assert abs(t2 * (t2 - tau) - A / 4) < 1e-14
        assert t2 > tau
        # the relaxation uses the new tau
        assert np.allclose(fo, f - (f - feq) / t2, atol=1e-15)
        # C = 0: plain BGK
        fz, _, _, tz = o0.collide(0, f, rho, u)
        assert tz == tau
        assert np.allclose(fz, f - (f - feq) / tau, atol=1e-15)
    _, _, _, teq = o.collide(0, feq_numpy(lat, 1.0, np.zeros(dim)), 1.0, np.zeros(dim))
    assert abs(teq - tau) < 1e-15


@pytest.mark.parametrize("lat,q,dim,L", LATTICES)
def test_mrt_conserves_and_reduces_to_bgk(lat, q, dim, L):
    """MRT (code 4) with the conserved rates zero conserves rho and j exactly
    up to rounding, and with every rate 1/tau equals BGK on the incompressible
    equilibrium (the equilibrium whose moments meq are)."""
    quirks = po.FIX_D3Q15_MINV | po.FIX_D3Q15_MEQ
    e, w, opp, M, Minv = tables(lat, quirks)
    tau = 0.8
    rates = [1.0 / tau] * q
    cons = [0, 3, 5] if lat == po.D2Q9 else [0, 3, 5, 7]
    rates_c = list(rates)
    for k in cons:
        rates_c[k] = 0.0
    o = po.Oracle(lat, L, po.MRT, tau_mrt=rates, quirks=quirks)
    oc = po.Oracle(lat, L, po.MRT, tau_mrt=rates_c, quirks=quirks)
    for _ in range(10):
        rho = 1 + 0.05 * RNG.standard_normal()
        u = 0.05 * RNG.standard_normal(dim)
        j = rho * u
        eu = e.T @ j
        feq_inc = w * (rho + 3 * eu + 4.5 * eu ** 2 - 1.5 * j @ j)
        f = feq_inc + 1e-3 * RNG.standard_normal(q)
        f[0] += rho - f.sum()
        fo, *_ = oc.collide(4, f, f.sum(), (e @ f) / f.sum())
        assert abs(fo.sum() - f.sum()) < 1e-15
        assert np.allclose(e @ fo, e @ f, atol=1e-16)
        r = f.sum()
        fb, *_ = o.collide(4, f, r, (e @ f) / r)
        jj = e @ f
        eu = e.T @ jj
        ref = w * (r + 3 * eu + 4.5 * eu ** 2 - 1.5 * jj @ jj)
        assert np.allclose(fb, f - (f - ref) / tau, atol=1e-15)


def test_zou_he_d2q9_closure():
    """zh_pressure_x/X (d2q9_zh_defs.cu): after the BC, sum f = rho_prescribed,
    j = (rho u_x, 0) with the u_x the BC computes."""
    e, w, *_ = tables(po.D2Q9)
    o = po.Oracle(po.D2Q9, (4, 4), micro_bc=True)
    for code in (1, 2):
        for _ in range(10):
            rho = 1 + 0.01 * RNG.standard_normal()
            f = feq_numpy(po.D2Q9, 1.0, 0.02 * RNG.standard_normal(2)) + 1e-4 * RNG.standard_normal(9)
            fo, r2, u2 = o.micro(code, f, rho, np.zeros(2))
            assert abs(fo.sum() - rho) < 1e-15
            assert np.allclose(e @ fo, [rho * u2[0], 0.0], atol=1e-16)
            assert u2[1] == 0.0
            # only the unknown populations change
            unknown = {1: [1, 5, 8], 2: [3, 6, 7]}[code]
            keep = [i for i in range(9) if i not in unknown]
            assert np.array_equal(fo[keep], f[keep])


@pytest.mark.parametrize("face", range(6))
def test_zou_he_d3q15_fixed_closure(face):
    """The corrected D3Q15 Zou-He closes mass and all momentum components."""
    e, w, *_ = tables(po.D3Q15)
    o = po.Oracle(po.D3Q15, (4, 4, 4), micro_bc=True, quirks=po.FIX_D3Q15_ZH)
    axis, side = face // 2, (1 if face % 2 == 0 else -1)
    for _ in range(10):
        rho = 1 + 0.01 * RNG.standard_normal()
        f = feq_numpy(po.D3Q15, 1.0, 0.02 * RNG.standard_normal(3)) + 1e-4 * RNG.standard_normal(15)
        fo, r2, u2 = o.micro(face + 1, f, rho, np.zeros(3))
        assert abs(fo.sum() - rho) < 1e-15
        want = np.zeros(3)
        want[axis] = rho * u2[axis]
        assert np.allclose(e @ fo, want, atol=1e-16)
        changed = np.nonzero(fo != f)[0]
        assert set(changed) <= set(np.nonzero(e[axis] == side)[0])


# Moments left by the reference's D3Q15 Zou-He faces (d3q15_zh_defs.cu:4-92),
# derived by hand from its formulas.  Notation: A = f3 - f4, B = f5 - f6,
# P = f1 - f2 (populations the face keeps); K = the sum of the four known
# diagonals (e_n = -side), Kab = their sum weighted by e_a e_b (a < b the two
# transverse axes); S = f0 + the four in-plane axis populations + 2 * (the
# five incoming known populations); u_n = side * (1 - S / rho) is the normal
# velocity the face stores.  Each face adds c = rho u_n / 12 to the four
# unknown diagonals and subtracts a quarter of a signed sum of two of A, B, P;
# the signs decide the transverse leftovers:
#   face  sum f                         j = (jx, jy, jz)      Pi_ab          Pi_aa (a: first transverse axis)
#   x     rho                           (rho u_x, 2A, 0)      2 Kab          f3+f4 + 2K + rho u_x / 3
#   X     rho                           (rho u_x, 2A, B)      2 Kab + B      f3+f4 + 2K - rho u_x / 3
#   y     rho                           (2P, rho u_y, 0)      2 Kab          f1+f2 + 2K + rho u_y / 3
#   Y     rho                           (2P, rho u_y, 0)      2 Kab          f1+f2 + 2K - rho u_y / 3
#   z     rho                           (2P, 0, rho u_z)      2 Kab          f1+f2 + 2K + rho u_z / 3
#   Z     f0+f1+f2+f3+f4 + 2(f5+f7+f8+f9+f10) = rho (1 + u_z)
#                                       (2P, 0, 0)            2 Kab          f1+f2 + 2K
# Face Z computes its rho-u terms from u[0], which it has just set to 0
# (:87-91): the unknowns copy their partners plus the transverse term only,
# so neither mass nor normal momentum closes.
ZH_REF = {
    # face: (jx, jy, jz as functions of (f, rho u_n)), extra Pi_ab, sign of the rho u_n / 3 term
    0: (lambda f, m: (m, 2 * (f[3] - f[4]), 0.0), lambda f: 0.0, 1),
    1: (lambda f, m: (m, 2 * (f[3] - f[4]), f[5] - f[6]), lambda f: f[5] - f[6], -1),
    2: (lambda f, m: (2 * (f[1] - f[2]), m, 0.0), lambda f: 0.0, 1),
    3: (lambda f, m: (2 * (f[1] - f[2]), m, 0.0), lambda f: 0.0, -1),
    4: (lambda f, m: (2 * (f[1] - f[2]), 0.0, m), lambda f: 0.0, 1),
    5: (lambda f, m: (2 * (f[1] - f[2]), 0.0, 0.0), lambda f: 0.0, 0),
}


def test_zou_he_d3q15_moment_set_fixes_the_unknowns():
    """The five functionals the reference-face pin checks (sum f, j_a, j_b,
    Pi_ab, Pi_aa) are independent on each face's five unknown populations,
    so the pin leaves no combination of the unknowns unchecked."""
    e, *_ = tables(po.D3Q15)
    for face in range(6):
        axis, side = face // 2, (1 if face % 2 == 0 else -1)
        a, b = [d for d in range(3) if d != axis]
        unk = np.nonzero(e[axis] == side)[0]
        rows = np.array([np.ones(15), e[a], e[b], e[a] * e[b], e[a] ** 2])[:, unk]
        assert len(unk) == 5 and np.linalg.matrix_rank(rows) == 5


@pytest.mark.parametrize("face", range(6))
def test_zou_he_d3q15_reference_defects(face):
    """The reference D3Q15 faces (d3q15_zh_defs.cu:4-92) against the moments
    derived by hand from them (table above): mass, all three momentum
    components, the transverse shear moment and one normal-stress moment,
    on random nodes, to FP64 rounding.  Only the unknown populations
    change, rho is kept, the transverse velocities are zero."""
    e, *_ = tables(po.D3Q15)
    o = po.Oracle(po.D3Q15, (4, 4, 4), micro_bc=True)
    axis, side = face // 2, (1 if face % 2 == 0 else -1)
    a, b = [d for d in range(3) if d != axis]
    jfun, pab_extra, usign = ZH_REF[face]
    known_diag = np.nonzero((e[axis] == -side) & (np.abs(e).sum(axis=0) == 3))[0]
    inplane = np.nonzero((e[axis] == 0) & (np.abs(e).sum(axis=0) == 1))[0]
    incoming = np.nonzero(e[axis] == -side)[0]
    for _ in range(20):
        rho = 1 + 0.01 * RNG.standard_normal()
        f = feq_numpy(po.D3Q15, 1.0, 0.03 * RNG.standard_normal(3)) + 3e-3 * RNG.standard_normal(15)
        fo, r2, u2 = o.micro(face + 1, f, rho, np.zeros(3))
        S = f[0] + f[inplane].sum() + 2 * f[incoming].sum()
        un = side * (1 - S / rho)
        assert r2 == rho
        assert abs(u2[axis] - un) < 1e-14 and u2[a] == 0.0 and u2[b] == 0.0
        m = rho * u2[axis]
        tol = 1e-15
        assert abs(fo.sum() - (rho if face < 5 else S)) < tol
        assert np.allclose(e @ fo, jfun(f, m), rtol=0, atol=tol)
        K = f[known_diag].sum()
        Kab = (e[a] * e[b] * f)[known_diag].sum()
        assert abs((e[a] * e[b] * fo).sum() - (2 * Kab + pab_extra(f))) < tol
        Paa = f[inplane][np.abs(e[a][inplane]) == 1].sum() + 2 * K + usign * m / 3
        assert abs((e[a] ** 2 * fo).sum() - Paa) < tol
        changed = np.nonzero(fo != f)[0]
        assert set(changed) <= set(np.nonzero(e[axis] == side)[0])


@pytest.mark.parametrize("lat,q,dim,L", LATTICES)
def test_bounceback_and_ntpor_limits(lat, q, dim, L):
    e, w, opp, *_ = tables(lat)
    tau = 0.9
    o = po.Oracle(lat, L, po.NTPOR, tau=tau)
    ob = po.Oracle(lat, L, po.BGK, tau=tau)
    for _ in range(10):
        rho = 1 + 0.05 * RNG.standard_normal()
        u = 0.05 * RNG.standard_normal(dim)
        f = feq_numpy(lat, rho, u) + 1e-3 * RNG.standard_normal(q)
        r = f.sum()
        uu = (e @ f) / r
        # bounceback (code 8): a permutation by opp, rho and u zeroed
        fb, rb, ub, _ = o.collide(8, f, r, uu)
        assert np.array_equal(fb, f[opp]) and rb == 0 and np.all(ub == 0)
        # NTPOR with B = 0 is BGK bit for bit
        f0, *_ = o.collide(2, f, r, uu, B=0.0)
        fz, *_ = ob.collide(0, f, r, uu)
        assert np.array_equal(f0, fz)
        # any B: mass conserved; B -> 1 gives the bounced populations
        B = RNG.uniform(0, 1)
        fp, *_ = o.collide(2, f, r, uu, B=B)
        assert abs(fp.sum() - f.sum()) < 1e-15
        f1, *_ = o.collide(2, f, r, uu, B=1.0)
        assert np.allclose(f1, f[opp], atol=1e-16)
        # the blend is linear in B between its two limits
        assert np.allclose(fp, (1 - B) * fz + B * f1, atol=1e-15)


@pytest.mark.parametrize("lat,q,dim,L", LATTICES)
def test_macro_code_table(lat, q, dim, L):
    """macro_conditions[code-1] (macro_bc.cu:12-17): which of rho, ux, uy, uz
    each code prescribes."""
    sets = {1: "x", 2: "y", 3: "z", 4: "xy", 5: "xz", 6: "yz", 7: "xyz", 8: "r", 9: "rx",
            10: "ry", 11: "rz", 12: "rxy", 13: "rxz", 14: "ryz", 15: "rxyz"}
    o = po.Oracle(lat, L, macro_bc=True)
    for code, s in sets.items():
        r, u = o.macro(code, 1.0, np.full(dim, 0.5), 7.0, np.array([2.0, 3.0, 4.0])[:dim])
        assert r == (7.0 if "r" in s else 1.0)
        for d, ax in enumerate("xyz"[:dim]):
            assert u[d] == ([2.0, 3.0, 4.0][d] if ax in s else 0.5)


@pytest.mark.parametrize("lat,q,dim,L", LATTICES)
def test_sf_copies_entering_directions(lat, q, dim, L):
    e, *_ = tables(lat)
    o = po.Oracle(lat, L, micro_bc=True)
    codes = [5, 6, 7, 8] if lat == po.D2Q9 else [7, 8, 9, 10, 11, 12]
    for k, code in enumerate(codes):
        axis, side = k // 2, (1 if k % 2 == 0 else -1)
        f = RNG.random(q)
        nb = RNG.random(15)
        fo, _, _ = o.micro(code, f, 1.0, np.zeros(dim), nb=nb)
        enter = e[axis] == side
        assert np.array_equal(fo[enter], nb[:q][enter])
        assert np.array_equal(fo[~enter], f[~enter])
