"""FP64 reference of the learned relaxation-time closure's MLP.

TEST INFRASTRUCTURE ONLY: the checker for the HIP training and inference
kernels (tests/ and bench-side experiments); the product never imports it.

Network: 7 -> 64 -> 64 -> 1, activation GELU(x) = x Phi(x), Phi an odd
degree-9 polynomial (fixed FP32 coefficients, least-squares fit of the
normal CDF on [-3.5, 3.5]) added to 1/2 and clamped to [0, 1].
  * phi / gelu / dgelu: the polynomial in exact (float64) arithmetic on x
    clamped to [-3.5, 3.5]; |error| <= 8.6e-4 in Phi, <= 2.9e-3 in GELU.
  * phi16 / gelu16 / dgelu16: the device's evaluation (lbm-c_amd/csrc/mlp.hpp
    gelu_pair, gelu_grad): x = fp16(z) rounded to nearest even, t = x*x and
    four fused multiply-adds each rounded to fp16 once, 1/2 + x p as a fused
    multiply-add clamped to [0, 1], no input clamp; GELU = fp32(z * Phi).
    |error| <= 1.8e-3 in Phi over every finite fp16 input, exactly 0 / 1
    beyond -3.5 / 3.5.
  * gelu16h: the GELU of the evaluation with fp16 operands (mlp.hpp
    gelu_oct<true>): the same Phi, then fp16(x * Phi) with x = fp16(z).
tests/test_closure_math.py checks all of these against erf.

Parameters (one dict of float64 arrays):
    w1 (64, 8)  hidden x (7 normalised inputs + a constant-1 bias column)
    w2 (64, 64), b2 (64,), w3 (64,), b3 ()

`quant=True` is the device's bf16 arithmetic, the training kernels'
(lbm-c_amd/csrc/mlp.hip mlp_train_kernel) and the default evaluation's
(mlp.hpp mlp_wave_eval: the closure steps and lbmc_mlp_forward):
activations and their derivatives as the device evaluates them (gelu16,
dgelu16), and the inputs, the weights, both hidden activations and both
hidden-layer error signals rounded to bfloat16 (round to nearest even,
through float32) before they enter a product.  `quant="fp16"` is the
evaluation with fp16 operands (lbmc_mlp_set_eval_fp16): inputs and weights
rounded to fp16 through float32, activations gelu16h (already fp16),
layer-2 bias in float32.  Sums stay in float64.  The device accumulates in float32, so quantised results
agree with the device to float32 summation error (and the rare fp16 / bf16
rounding that lands on the other side of a tie because of it).
"""
from __future__ import annotations

import numpy as np

GELU_R = 3.5
GELU_C = np.array([0.3964735269546509, -0.061638616025447845, 0.007179733831435442,
                   -0.0004616279620677233, 1.2130119102948811e-05])
NPARAM = 64 * 8 + 64 * 64 + 64 + 64 + 1


def bf16(a):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (r.astype(np.uint32) << 16).view(np.float32)
    out = np.where(np.isnan(a32), a32, out)
    return out.astype(np.float64)


def phi(x):
    xc = np.clip(x, -GELU_R, GELU_R)
    t = xc * xc
    c = GELU_C
    p = xc * (c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * c[4]))))
    return np.clip(0.5 + p, 0.0, 1.0)


def gelu(x):
    return x * phi(x)


def dgelu(x):
    xc = np.clip(x, -GELU_R, GELU_R)
    t = xc * xc
    c = GELU_C
    p = xc * (c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * c[4]))))
    dp = c[0] + t * (3 * c[1] + t * (5 * c[2] + t * (7 * c[3] + t * 9 * c[4])))
    inside = (np.abs(x) < GELU_R) & (0.5 + p > 0.0) & (0.5 + p < 1.0)
    return phi(x) + x * np.where(inside, dp, 0.0)


def _fp16(a):
    """Round float64 values to fp16 (nearest even) and back: one fp16
    rounding of an exactly computed value (products of fp16 values and their
    sums with an fp16 term are exact in float64 in the ranges used here)."""
    with np.errstate(over="ignore", invalid="ignore"):
        return np.asarray(a, dtype=np.float64).astype(np.float16).astype(np.float64)


GELU_C16 = _fp16(GELU_C)


def _phi16_raw(x16):
    """1/2 + x P(x^2) of fp16 inputs, each operation rounded to fp16 once."""
    c = GELU_C16
    with np.errstate(over="ignore", invalid="ignore"):
        t = _fp16(x16 * x16)
        p = _fp16(t * c[4] + c[3])
        p = _fp16(t * p + c[2])
        p = _fp16(t * p + c[1])
        p = _fp16(t * p + c[0])
        return _fp16(x16 * p + 0.5)


def phi16(z):
    """The device's Phi of pre-activation z (fp32 or float64 values)."""
    z32 = np.asarray(z, dtype=np.float32).astype(np.float64)
    return np.clip(_phi16_raw(_fp16(z32)), 0.0, 1.0)


def gelu16(z):
    """The device's GELU before the bf16 rounding: fp32(z * Phi)."""
    z32 = np.asarray(z, dtype=np.float32).astype(np.float64)
    return (z32 * phi16(z32)).astype(np.float32).astype(np.float64)


def fp16(a):
    """Round to fp16 (nearest even) through float32, as the device converts
    FP32 values; returned as float64."""
    return _fp16(np.asarray(a, dtype=np.float32).astype(np.float64))


def gelu16h(z):
    """The evaluation's GELU: x = fp16(z), the device's Phi of x, and the
    product x * Phi rounded to fp16 once (exact in float64 before it)."""
    x = fp16(z)
    with np.errstate(invalid="ignore"):
        return _fp16(x * np.clip(_phi16_raw(x), 0.0, 1.0))


def dgelu16(z):
    """The device's GELU derivative: Phi(x) + z Phi'(x), x = fp16(z), with
    Phi and its clamp test from the fp16 evaluation and Phi' exact."""
    z = np.asarray(z, dtype=np.float64)
    x = _fp16(np.asarray(z, dtype=np.float32).astype(np.float64))
    raw = _phi16_raw(x)
    t = x * x
    c = GELU_C
    dp = c[0] + t * (3 * c[1] + t * (5 * c[2] + t * (7 * c[3] + t * 9 * c[4])))
    inside = (raw > 0.0) & (raw < 1.0)
    return np.clip(raw, 0.0, 1.0) + np.where(inside, z * dp, 0.0)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((64, 8)) * np.sqrt(2.0 / 7.0)
    w1[:, 7] = 0.0
    return {"w1": w1, "w2": rng.standard_normal((64, 64)) * np.sqrt(2.0 / 64.0),
            "b2": np.zeros(64), "w3": rng.standard_normal(64) * np.sqrt(1.0 / 64.0), "b3": 0.0}


def flatten(p):
    return np.concatenate([p["w1"].reshape(-1), p["w2"].reshape(-1), p["b2"], p["w3"],
                           np.atleast_1d(p["b3"])])


def unflatten(v):
    v = np.asarray(v, dtype=np.float64)
    o = 0
    w1 = v[o:o + 512].reshape(64, 8); o += 512
    w2 = v[o:o + 4096].reshape(64, 64); o += 4096
    b2 = v[o:o + 64]; o += 64
    w3 = v[o:o + 64]; o += 64
    return {"w1": w1, "w2": w2, "b2": b2, "w3": w3, "b3": float(v[o])}


def forward(p, x, quant=False):
    """x: (S, 7) normalised features -> y (S,), cache.  quant: False (exact),
    True (bf16 arithmetic) or "fp16" (the evaluation with fp16 operands)."""
    if quant == "fp16":
        X = np.concatenate([x, np.ones((x.shape[0], 1))], axis=1)
        Xq, W1, W2, w3 = fp16(X), fp16(p["w1"]), fp16(p["w2"]), fp16(p["w3"])
        z1 = Xq @ W1.T
        a1 = gelu16h(z1)
        z2 = a1 @ W2.T + np.asarray(p["b2"], dtype=np.float32).astype(np.float64)
        a2 = gelu16h(z2)
        y = a2 @ w3 + p["b3"]
        return y, dict(Xq=Xq, W2=W2, w3=w3, z1=z1, a1=a1, z2=z2, a2=a2)
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    act = gelu16 if quant else gelu
    X = np.concatenate([x, np.ones((x.shape[0], 1))], axis=1)
    Xq, W1, W2, w3 = q(X), q(p["w1"]), q(p["w2"]), q(p["w3"])
    z1 = Xq @ W1.T
    a1 = q(act(z1))
    z2 = a1 @ W2.T + p["b2"]
    a2 = q(act(z2))
    y = a2 @ w3 + p["b3"]
    return y, dict(Xq=Xq, W2=W2, w3=w3, z1=z1, a1=a1, z2=z2, a2=a2)


def loss_and_grad(p, x, t, quant=False):
    """Mean squared error over the S samples and its gradient (same dict form)."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    dact = dgelu16 if quant else dgelu
    y, c = forward(p, x, quant)
    S = x.shape[0]
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / S
    gw3 = c["a2"].T @ dy
    gb3 = float(dy.sum())
    dz2 = np.outer(dy, c["w3"]) * dact(c["z2"])
    gb2 = dz2.sum(axis=0)
    dz2q = q(dz2)
    gw2 = dz2q.T @ c["a1"]
    da1 = dz2q @ c["W2"]
    dz1q = q(da1 * dact(c["z1"]))
    gw1 = dz1q.T @ c["Xq"]
    return loss, {"w1": gw1, "w2": gw2, "b2": gb2, "w3": gw3, "b3": gb3}
