"""CPU pins of the closure MLP oracle (oracle/pymlp.py): the GELU
approximation against erf, the backward pass against central finite
differences, and the bf16 rounding against its definition."""
import numpy as np
from scipy.special import erf

import pymlp


def test_gelu_matches_normal_cdf():
    x = np.linspace(-12, 12, 200001)
    phi_true = 0.5 * (1 + erf(x / np.sqrt(2)))
    assert np.max(np.abs(pymlp.phi(x) - phi_true)) < 8.6e-4
    assert np.max(np.abs(pymlp.gelu(x) - x * phi_true)) < 2.9e-3
    # symmetric: Phi(-x) = 1 - Phi(x)
    assert np.allclose(pymlp.phi(-x), 1 - pymlp.phi(x), atol=1e-15)


def test_dgelu_matches_finite_difference():
    x = np.concatenate([np.linspace(-5, 5, 1001), [-3.49, 3.49, 0.0]])
    h = 1e-6
    fd = (pymlp.gelu(x + h) - pymlp.gelu(x - h)) / (2 * h)
    ok = np.abs(np.abs(x) - pymlp.GELU_R) > 1e-3
    assert np.max(np.abs(fd[ok] - pymlp.dgelu(x)[ok])) < 1e-6


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = pymlp.init_params(1)
    p["b2"] = 0.1 * rng.standard_normal(64)
    p["b3"] = 0.05
    x = rng.standard_normal((32, 7))
    t = rng.standard_normal(32)
    loss, g = pymlp.loss_and_grad(p, x, t)
    v = pymlp.flatten(p)
    gv = pymlp.flatten(g)
    idx = rng.choice(v.size, 120, replace=False)
    h = 1e-6
    for k in idx:
        vp, vm = v.copy(), v.copy()
        vp[k] += h
        vm[k] -= h
        lp, _ = pymlp.loss_and_grad(pymlp.unflatten(vp), x, t)
        lm, _ = pymlp.loss_and_grad(pymlp.unflatten(vm), x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gv[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, gv[k])


def test_bf16_rounding():
    # exactly representable values stay; ties go to even; NaN stays NaN
    assert pymlp.bf16(1.0) == 1.0
    assert pymlp.bf16(1.0 + 2 ** -8) == 1.0          # tie -> even mantissa
    assert pymlp.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert pymlp.bf16(1.0 + 2 ** -8 + 2 ** -20) == 1.0 + 2 ** -7
    assert np.isnan(pymlp.bf16(np.nan))
    a = np.random.default_rng(0).standard_normal(1000)
    assert np.max(np.abs(pymlp.bf16(a) - a) / np.abs(a)) <= 2 ** -8   # half an ulp of 8 bits


def test_quantised_forward_close_to_exact():
    rng = np.random.default_rng(4)
    p = pymlp.init_params(2)
    x = rng.standard_normal((256, 7))
    y0, _ = pymlp.forward(p, x)
    y1, _ = pymlp.forward(p, x, quant=True)
    assert np.max(np.abs(y1 - y0)) < 0.05 * np.max(np.abs(y0))
