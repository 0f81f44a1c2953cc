"""Host-side closure pipeline (lbm-c_amd/closure.py): filter, strain rate,
subgrid-stress targets and features, against closed forms."""
import numpy as np

import closure


def test_box_filter_is_block_mean():
    a = np.arange(4 * 4 * 8, dtype=np.float64).reshape(4, 4, 8)
    b = closure.box_filter(a, 2)
    assert b.shape == (2, 2, 4)
    assert b[1, 0, 3] == a[2:4, 0:2, 6:8].mean()


def test_strain_rate_of_a_shear_wave():
    L = 32
    k = 2 * np.pi / L
    z, y, x = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    ux = np.sin(k * y)
    u = np.stack([ux, np.zeros_like(ux), np.zeros_like(ux)])
    S = closure.strain_rate(u)
    # central difference of sin(k y): sin(k) cos(k y); S_xy = half of it
    assert np.allclose(S[3], 0.5 * np.sin(k) * np.cos(k * y), atol=1e-12)
    for c in (0, 1, 2, 4, 5):
        assert np.max(np.abs(S[c])) < 1e-12


def test_sgs_targets_vanish_for_blockwise_constant_fields():
    rng = np.random.default_rng(0)
    coarse = rng.standard_normal((3, 4, 4, 4)) * 0.01
    u = np.stack([np.kron(coarse[a], np.ones((2, 2, 2))) for a in range(3)])
    rho = np.ones((8, 8, 8))
    rb, ub, S, tsgs, tau_t = closure.sgs_targets(rho, u, 2, 0.50125)
    assert np.allclose(ub, coarse)
    assert np.max(np.abs(tsgs)) < 1e-15
    assert np.allclose(tau_t, 0.50125)


def test_eddy_viscosity_of_a_prescribed_stress():
    # tau = -2 nu S + c I: the least-squares fit against the deviatoric part
    # is nu (S^dev : S) / (S : S) = nu (1 - tr(S)^2 / (3 S:S)); backscatter
    # (nu < 0) is clipped to zero
    rng = np.random.default_rng(1)
    S = 1e-3 * rng.standard_normal((6, 5, 5, 5))
    c = rng.standard_normal((5, 5, 5))
    for nu in (0.003, -0.002):
        tau = -2.0 * nu * S
        tau[0:3] += c
        got = closure.eddy_viscosity(tau, S)
        ss = S[0] ** 2 + S[1] ** 2 + S[2] ** 2 + 2 * (S[3] ** 2 + S[4] ** 2 + S[5] ** 2)
        tr = S[0] + S[1] + S[2]
        expect = np.maximum(0.0, nu * (1.0 - tr ** 2 / (3.0 * ss)))
        assert np.allclose(got, expect, rtol=1e-9, atol=1e-15)


def test_features_are_scaled_strain():
    S = np.arange(6 * 2 * 2 * 2, dtype=np.float64).reshape(6, 2, 2, 2) * 1e-4
    rho = np.full((2, 2, 2), 1.01)
    tau = np.full((2, 2, 2), 0.52)
    f = closure.features(rho, S, tau)
    assert f.shape == (8, 7)
    assert np.allclose(f[:, 3], -(2.0 / 3.0) * 0.52 * S[3].reshape(-1), rtol=1e-6)
    assert np.allclose(f[:, 6], 1.01)


def test_init_params_layout():
    p = closure.init_params(3)
    assert p.size == closure.NPARAM == 4737
    w1 = p[:512].reshape(64, 8)
    assert np.all(w1[:, 7] == 0) and np.all(p[closure.B2:closure.W3] == 0) and p[closure.B3] == 0


# ---- 2-D pipeline of the ensemble's closure (lbm-c_amd/ensemble_training.py)

def test_ensemble_training_filter_and_strain_2d():
    import ensemble_training as et

    a = np.arange(6 * 8, dtype=np.float64).reshape(6, 8)
    assert et.box2(a, 2)[2, 3] == a[4:6, 6:8].mean()
    L = 32
    k = 2 * np.pi / L
    y, x = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    S = et.strain2(np.stack([np.sin(k * y), np.zeros((L, L))]))
    assert np.allclose(S[2], 0.5 * np.sin(k) * np.cos(k * y), atol=1e-12)
    assert np.max(np.abs(S[0])) < 1e-12 and np.max(np.abs(S[1])) < 1e-12


def test_ensemble_training_eddy_viscosity_2d():
    # traceless 2-D strain (incompressible): tau = -2 nu S + c I gives nu
    # exactly; backscatter clipped
    rng = np.random.default_rng(2)
    sxx = 1e-3 * rng.standard_normal((6, 7))
    S = np.stack([sxx, -sxx, 1e-3 * rng.standard_normal((6, 7))])
    c = rng.standard_normal((6, 7))
    import ensemble_training as et

    for nu in (0.004, -0.001):
        ts = -2.0 * nu * S
        ts[0] += c
        ts[1] += c
        assert np.allclose(et.nu_t2(ts, S), max(nu, 0.0), rtol=1e-9, atol=1e-15)


def test_ensemble_training_refinement_keeps_reynolds_number():
    import ensemble_training as et
    import workloads

    m = workloads.ensemble_d2q9(4, 0)[1]
    mf = et.refine(m, 2)
    assert mf["length"] == (2 * m["length"][0], 2 * m["length"][1])
    assert np.array_equal(mf["geometry"][::2, ::2], m["geometry"])
    nu, nuf = (m["tau"] - 0.5) / 3, (mf["tau"] - 0.5) / 3
    # same physical viscosity in units of the coarse node and step: nu_f = r nu
    assert np.isclose(nuf, 2 * nu) and np.isclose(mf["force"], m["force"] / 2)
    # the start profile: zero in the solids, block means of the fine profile
    # equal the coarse profile's shape (maximum near the centre line)
    ux, uy = et.fine_start(m, 2, 0.0, None)
    assert np.all(ux[mf["geometry"] >= 1] == 0) and np.all(uy == 0)
    prof = et.box2(ux, 2)[:, 0]
    assert abs(int(np.argmax(prof)) - (m["length"][1] - 1) / 2) <= 1


def test_trained_ensemble_closure_file():
    """The committed trained ensemble network loads without pickles, has the
    network's parameter count and finite values, and predicts a correction
    (tau_ref 0) of the size of the eddy viscosities it was trained on."""
    import workloads

    p, mu, isd, out_a, out_b, tau_ref = workloads.ensemble_closure()
    assert p.shape == (4737,) and mu.shape == (7,) and isd.shape == (7,)
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(mu)) and np.all(isd > 0)
    assert tau_ref == 0.0 and 0.0 < out_a < 1e-2 and 0.0 <= out_b < 1e-2
    p0, _, _, a0, b0, r0 = workloads.ensemble_closure(trained=False)
    assert p0.shape == (4737,) and (a0, b0, r0) == (1e-3, 0.55, 0.55)
