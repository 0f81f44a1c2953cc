"""CPU pins of the closure MLP oracle (oracle/pymlp.py): the GELU
approximation against erf (exact-arithmetic form, and the device's fp16
evaluation over every finite fp16 input), the backward pass against central
finite differences, and the bf16 / fp16 roundings against their
definitions."""
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


def _all_fp16():
    h = np.arange(0x10000, dtype=np.uint32).astype(np.uint16).view(np.float16)
    return h[np.isfinite(h)].astype(np.float64)


def test_gelu16_matches_normal_cdf_on_every_fp16_input():
    x = _all_fp16()
    phi_true = 0.5 * (1 + erf(x / np.sqrt(2)))
    s = pymlp.phi16(x)
    assert not np.any(np.isnan(s))
    assert np.max(np.abs(s - phi_true)) < 1.8e-3
    # beyond the fit interval the unclamped polynomial saturates exactly
    assert np.all(s[x > 3.5] == 1.0) and np.all(s[x < -3.5] == 0.0)
    g = pymlp.gelu16(x)
    small = np.abs(x) <= 8
    assert np.max(np.abs(g[small] - x[small] * phi_true[small])) < 5.2e-3
    assert np.array_equal(g[x > 3.5], x[x > 3.5]) and np.all(g[x < -3.5] == 0.0)
    # Phi(-x) = 1 - Phi(x) up to one fp16 rounding of the last sum
    assert np.max(np.abs(pymlp.phi16(-x) - (1 - s))) <= 2.0 ** -11


def test_gelu16h_matches_normal_cdf_on_every_fp16_input():
    """The evaluation's GELU: an fp16 value within the fp16 Phi's error plus
    half an fp16 ulp of x Phi_true, exact tails, and the fp16 rounding of
    x Phi (not of z Phi in FP32, not bf16)."""
    x = _all_fp16()
    phi_true = 0.5 * (1 + erf(x / np.sqrt(2)))
    g = pymlp.gelu16h(x)
    assert not np.any(np.isnan(g))
    assert np.array_equal(g, pymlp._fp16(g))
    small = np.abs(x) <= 8
    ref = x[small] * phi_true[small]
    bound = 1.8e-3 * np.abs(x[small]) + 2.0 ** -11 * np.abs(ref) + 2.0 ** -25
    assert np.all(np.abs(g[small] - ref) <= bound)
    assert np.array_equal(g[x > 3.5], x[x > 3.5]) and np.all(g[x < -3.5] == 0.0)
    # fp16 results agree with the training form's FP32 product to half an fp16 ulp
    g32 = pymlp.gelu16(x)
    assert np.all(np.abs(g - g32) <= 2.0 ** -11 * np.abs(g32) + 2.0 ** -25)
    # a value rounded twice on its way: FP32 pre-activation, then fp16
    assert pymlp.fp16(1.0 + 2 ** -11 + 2 ** -30) == 1.0     # 2^-30 lost in float32: a tie, to even
    assert pymlp._fp16(1.0 + 2 ** -11 + 2 ** -30) == 1.0 + 2 ** -10


def test_fp16_rounding():
    f = pymlp._fp16
    assert f(1.0 + 2 ** -11) == 1.0                  # tie -> even
    assert f(1.0 + 3 * 2 ** -11) == 1.0 + 2 ** -9    # tie -> even (up)
    assert f(1.0 + 2 ** -11 + 2 ** -30) == 1.0 + 2 ** -10
    assert f(65519.0) == 65504.0 and np.isinf(f(65520.0))
    assert f(2.0 ** -24) == 2.0 ** -24 and f(2.0 ** -26) == 0.0   # subnormals kept, then to zero
    # the device's coefficients are the fp16 roundings of the fit's
    assert np.all(np.abs(pymlp.GELU_C16 / pymlp.GELU_C - 1) < 2.5e-3)


def test_dgelu16_close_to_exact_derivative():
    # The fp16 Phi reaches its clamp values 0 / 1 a little inside +-3.5
    # (from |x| ~ 2.9 on, where 1 - Phi < 2e-3), and there the derivative
    # loses its z Phi' term (<= 0.01 for |x| < 3.3); beyond 3.7 both forms
    # are exactly 0 or 1.
    x = np.linspace(-6, 6, 24001)
    d = np.abs(pymlp.dgelu16(x) - pymlp.dgelu(x))
    assert np.max(d[np.abs(x) < 2.8]) < 1e-3
    assert np.max(d[np.abs(x) < 3.3]) < 1.2e-2
    assert np.all(d[np.abs(x) > 3.7] == 0.0)


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
    # the evaluation's fp16 operands carry 3 more mantissa bits than bf16:
    # its rounding error is the smaller, the GELU approximation common to both
    y2, _ = pymlp.forward(p, x, quant="fp16")
    assert np.max(np.abs(y2 - y0)) < 0.05 * np.max(np.abs(y0))
    # the same network with the device's GELU and no operand rounding
    X = np.concatenate([x, np.ones((x.shape[0], 1))], axis=1)
    a1 = pymlp.gelu16(X @ p["w1"].T)
    a2 = pymlp.gelu16(a1 @ p["w2"].T + p["b2"])
    yr = a2 @ p["w3"] + p["b3"]
    e1, e2 = np.abs(y1 - yr), np.abs(y2 - yr)
    assert np.median(e2) < 0.25 * np.median(e1), (np.median(e1), np.median(e2))
