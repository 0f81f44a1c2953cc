"""Learned relaxation-time closure on the GPU against the FP64 oracles
(oracle/pymlp.py for the network, oracle/pyclosure.py for the step):
forward values, loss and gradient of the training kernels (bf16 rounding
points emulated in the oracle, FP32 accumulation on the device), the Adam
step, training progress and determinism, and the fused closure step."""
import numpy as np
import pytest

import lbmc
import pyclosure
import pymlp
import pyoracle
import workloads

pytestmark = pytest.mark.gpu


def random_params(seed, bias=0.1):
    rng = np.random.default_rng(seed)
    p = pymlp.init_params(seed)
    p["w1"][:, 7] = bias * rng.standard_normal(64)
    p["b2"] = bias * rng.standard_normal(64)
    p["b3"] = 0.05
    return p


def test_forward_matches_oracle():
    print("[closure] forward S=1000", flush=True)
    rng = np.random.default_rng(10)
    p = random_params(10)
    S = 1000  # not a multiple of 64: the last wave is ragged
    mu = rng.standard_normal(7).astype(np.float32)
    isd = (0.5 + rng.random(7)).astype(np.float32)
    feat = (mu + rng.standard_normal((S, 7)) / isd).astype(np.float32)
    m = lbmc.Mlp(pymlp.flatten(p), mu=mu, isd=isd)
    y = m.forward(feat).astype(np.float64)
    xn = ((feat - mu) * isd).astype(np.float64)
    yq, _ = pymlp.forward(p, xn, quant=True)
    y0, _ = pymlp.forward(p, xn, quant=False)
    scale = np.max(np.abs(yq))
    err = np.abs(y - yq)
    # FP32 accumulation vs FP64 with the same bf16 operands: a bf16 rounding of
    # an activation may land on the other side, one ulp (2^-8) of that value
    assert np.median(err) < 1e-5 * scale
    assert np.max(err) < 3e-3 * scale
    # and the quantised network stays close to the exact one
    assert np.max(np.abs(y - y0)) < 0.05 * np.max(np.abs(y0))
    m.close()


def test_loss_and_gradient_match_oracle():
    print("[closure] loss/grad S=300", flush=True)
    rng = np.random.default_rng(11)
    p = random_params(11)
    S = 300
    x = rng.standard_normal((S, 7)).astype(np.float32)
    t = rng.standard_normal(S).astype(np.float32)
    m = lbmc.Mlp(pymlp.flatten(p))
    loss, g = m.loss_grad(x, t)
    lq, gq = pymlp.loss_and_grad(p, x.astype(np.float64), t.astype(np.float64), quant=True)
    assert abs(loss - lq) < 1e-4 * lq
    gv = pymlp.flatten(gq)
    for name, sl in (("w1", slice(0, 512)), ("w2", slice(512, 4608)), ("b2", slice(4608, 4672)),
                     ("w3", slice(4672, 4736)), ("b3", slice(4736, 4737))):
        d = np.linalg.norm(g[sl] - gv[sl]) / np.linalg.norm(gv[sl])
        assert d < 1e-2, (name, d)
    # bias column of w1 is the sum of the first-layer error signal
    assert np.allclose(g[0:512].reshape(64, 8)[:, 7], gv[0:512].reshape(64, 8)[:, 7],
                       atol=2e-2 * np.max(np.abs(gv[0:512])))
    m.close()


def test_adam_first_step():
    print("[closure] adam step", flush=True)
    rng = np.random.default_rng(12)
    p = random_params(12)
    S = 256
    x = rng.standard_normal((S, 7)).astype(np.float32)
    t = rng.standard_normal(S).astype(np.float32)
    m = lbmc.Mlp(pymlp.flatten(p))
    _, g = m.loss_grad(x, t)
    p0 = m.get_params()
    m.set_data(x, t)
    lr = 1e-3
    m.train(1, batch=S, lr=lr, seed=0)
    dp = m.get_params().astype(np.float64) - p0
    # first Adam step: m_hat = g, v_hat = g^2 -> -lr g / (|g| + eps)
    big = np.abs(g) > 1e-5
    expect = -lr * g / (np.abs(g) + 1e-8)
    assert np.max(np.abs(dp[big] - expect[big])) < 1e-3 * lr
    m.close()


def test_training_fits_and_is_deterministic():
    print("[closure] training 20 epochs S=8192", flush=True)
    rng = np.random.default_rng(13)
    S = 8192
    x = rng.standard_normal((S, 7)).astype(np.float32)
    t = (np.tanh(x[:, 0] - 0.5 * x[:, 3]) + 0.3 * x[:, 6] ** 2).astype(np.float32)
    p = pymlp.flatten(random_params(13))
    runs = []
    for _ in range(2):
        m = lbmc.Mlp(p)
        m.set_data(x, t)
        hist = m.train(20, batch=512, lr=3e-3, seed=5)
        runs.append((hist, m.get_params()))
        m.close()
    hist, params = runs[0]
    assert hist[-1] < 0.1 * hist[0], hist
    assert np.array_equal(runs[0][1], runs[1][1])
    # the device's trained network evaluated by the oracle fits as well
    y, _ = pymlp.forward(pymlp.unflatten(params.astype(np.float64)), x.astype(np.float64), quant=True)
    assert np.mean((y - t) ** 2) < 0.1 * np.var(t)


def _closure_case(lattice, L, tau, c_smag, storage, force=None, solid=None, seed=20, steps=12, every=1):
    """Random network whose tau' falls inside the band at most nodes and
    outside it at some; returns (device macros, oracle macros, fallback share)."""
    dim = 2 if lattice == lbmc.D2Q9 else 3
    n = int(np.prod(L))
    rng = np.random.default_rng(seed)
    if dim == 3:
        ux, uy, uz = workloads.decaying_turbulence(L[0], k0=2.0, u_rms=0.04, seed=seed)
        u = np.stack([ux.reshape(-1), uy.reshape(-1), uz.reshape(-1)])
    else:
        u = 0.03 * rng.standard_normal((2, n))
    rho = 1.0 + 0.002 * rng.standard_normal(n)
    p = random_params(seed)
    # normalisation from the initial non-equilibrium scale
    mu = np.zeros(7, dtype=np.float32)
    mu[6] = 1.0
    isd = np.full(7, 300.0, dtype=np.float32)
    isd[6] = 100.0
    out_a, out_b, lo, hi = 0.02, tau + 0.01, tau, tau + 0.05
    lat = lbmc.Lattice(lattice, L, lbmc.BGK, forcing=force is not None, tau=tau, c_smag=c_smag,
                       storage=storage)
    fields = {}
    if solid is not None:
        fields["geometry"] = solid.astype(np.float64)
    if force is not None:
        fields["force"] = [np.full(n, force[d]) for d in range(dim)]
    lat.set_fields(**fields)
    lat.init_equilibrium(rho, u)
    m = lbmc.Mlp(pymlp.flatten(p), mu=mu, isd=isd, out_a=out_a, out_b=out_b)
    lat.set_closure(m, lo, hi)
    if every > 1:
        lat.set_closure_interval(every)
    lat.run(0, steps, 1)
    rd, ud = lat.get_macros()
    co = pyclosure.ClosureOracle(lattice, L, tau, c_smag, p, mu, isd, out_a, out_b, lo, hi,
                                 force=force, solid=solid, every=every)
    co.init_equilibrium(rho, u)
    ro, uo = co.run(steps)
    lat.close()
    m.close()
    return (rd, ud), (ro, uo), co.fallback.mean()


@pytest.mark.parametrize("storage,tol", [(lbmc.FP32, 2e-6), (lbmc.FP16S, 2e-4), (lbmc.BF16S, 2e-3)])
def test_closure_step_d3q15_matches_oracle(storage, tol):
    print(f"[closure] step D3Q15 16^3 storage={storage}", flush=True)
    (rd, ud), (ro, uo), fb = _closure_case(lbmc.D3Q15, (16, 16, 16), 0.51, 0.17, storage)
    assert 0.02 < fb < 0.9, fb  # both branches taken
    assert np.max(np.abs(ud - uo)) < tol
    assert np.max(np.abs(rd - ro)) < tol


def test_closure_step_d2q9_guo_solids_matches_oracle():
    print("[closure] step D2Q9 Guo + solids", flush=True)
    Lx, Ly = 40, 24
    solid = np.zeros((Ly, Lx))
    solid[0, :] = 1
    solid[-1, :] = 1
    solid[10:14, 18:22] = 1
    (rd, ud), (ro, uo), fb = _closure_case(lbmc.D2Q9, (Lx, Ly), 0.6, 0.1, lbmc.FP32,
                                           force=[2e-5, 0.0], solid=solid.reshape(-1), steps=15)
    assert 0.02 < fb < 0.98, fb
    assert np.max(np.abs(ud - uo)) < 2e-6
    assert np.max(np.abs(rd - ro)) < 2e-6


@pytest.mark.parametrize("every", [2, 3])
def test_closure_interval_matches_oracle(every):
    # 16^3: after the first (evaluating) call, the remaining 99 calls cover
    # the captured period lcm(16, every) at least twice, so graph replays run
    print(f"[closure] interval {every}", flush=True)
    (rd, ud), (ro, uo), fb = _closure_case(lbmc.D3Q15, (16, 16, 16), 0.51, 0.17, lbmc.FP32, steps=100,
                                           every=every)
    assert np.max(np.abs(ud - uo)) < 1e-5
    assert np.max(np.abs(rd - ro)) < 1e-5


def test_closure_off_restores_smagorinsky_bitwise():
    print("[closure] on/off", flush=True)
    L = (16, 16, 16)
    ux, uy, uz = workloads.decaying_turbulence(16, k0=2.0, u_rms=0.04, seed=3)
    u = [ux.reshape(-1), uy.reshape(-1), uz.reshape(-1)]
    rho = np.ones(16 ** 3)
    out = []
    for with_model in (False, True):
        lat = lbmc.Lattice(lbmc.D3Q15, L, lbmc.BGK, tau=0.51, c_smag=0.17, storage=lbmc.FP32)
        lat.set_fields()
        lat.init_equilibrium(rho, u)
        if with_model:
            m = lbmc.Mlp(pymlp.flatten(random_params(4)))
            lat.set_closure(m)
            lat.set_closure(None)
        lat.run(0, 10, 1)
        out.append(lat.get_f())
        lat.close()
    assert np.array_equal(out[0], out[1])


def test_closure_rejects_unsupported():
    m = lbmc.Mlp(pymlp.flatten(random_params(5)))
    lat = lbmc.Lattice(lbmc.D2Q9, (8, 8), lbmc.BGK, tau=0.8, storage=lbmc.FP64)
    with pytest.raises(lbmc.LbmcError) as e:
        lat.set_closure(m)
    assert e.value.status == 5
    lat.close()
    lat = lbmc.Lattice(lbmc.D2Q9, (8, 8), lbmc.MRT, tau=0.8, storage=lbmc.FP32)
    with pytest.raises(lbmc.LbmcError):
        lat.set_closure(m)
    lat.close()
    lat = lbmc.Lattice(lbmc.D2Q9, (8, 8), lbmc.BGK, tau=0.8, storage=lbmc.FP32)
    with pytest.raises(lbmc.LbmcError):
        lat.set_closure(m, 0.4, 2.0)
    lat.close()
    m.close()
