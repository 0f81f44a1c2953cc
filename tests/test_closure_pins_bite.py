"""The closure-oracle pins catch plausible mistakes: each mutation below is
planted in a copy of the oracle's function (on the CPU) and the
corresponding check of tests/test_closure_math.py must fail on it."""
import numpy as np
import pytest

import pymlp
import test_closure_math as pins


def _fails(check):
    with pytest.raises(AssertionError):
        check()


def test_gelu_pins_catch_wrong_coefficient(monkeypatch):
    c = pymlp.GELU_C.copy()
    c[1] = -c[1]  # sign of the cubic term
    monkeypatch.setattr(pymlp, "GELU_C", c)
    _fails(pins.test_gelu_matches_normal_cdf)


def test_gelu16_pin_catches_wrong_coefficient(monkeypatch):
    c = pymlp.GELU_C16.copy()
    c[2] = pymlp._fp16(1.01 * c[2])  # a 1 % slip in the quintic coefficient
    monkeypatch.setattr(pymlp, "GELU_C16", c)
    _fails(pins.test_gelu16_matches_normal_cdf_on_every_fp16_input)


def test_gelu16_pin_catches_missing_output_clamp(monkeypatch):
    monkeypatch.setattr(pymlp, "phi16", lambda z: pymlp._phi16_raw(pymlp._fp16(np.asarray(z, np.float64))))
    _fails(pins.test_gelu16_matches_normal_cdf_on_every_fp16_input)


def test_derivative_pin_catches_dropped_term(monkeypatch):
    monkeypatch.setattr(pymlp, "dgelu", lambda x: pymlp.phi(x))  # x Phi'(x) dropped
    _fails(pins.test_dgelu_matches_finite_difference)


def test_backward_pin_catches_transposed_weights(monkeypatch):
    orig = pymlp.loss_and_grad

    def mutated(p, x, t, quant=False):
        q = dict(p)
        q["w2"] = np.ascontiguousarray(p["w2"])
        loss, g = orig(p, x, t, quant)
        # da1 = dz2 @ W2 computed as dz2 @ W2^T: recompute gw1 with the transpose
        y, c = pymlp.forward(p, x, quant)
        dy = 2.0 * (y - t) / x.shape[0]
        dz2 = np.outer(dy, c["w3"]) * pymlp.dgelu(c["z2"])
        da1 = dz2 @ c["W2"].T
        g["w1"] = (da1 * pymlp.dgelu(c["z1"])).T @ c["Xq"]
        return loss, g

    monkeypatch.setattr(pymlp, "loss_and_grad", mutated)
    _fails(pins.test_backward_matches_finite_differences)


def test_bf16_pin_catches_truncation(monkeypatch):
    def truncate(a):
        a32 = np.asarray(a, dtype=np.float32)
        u = a32.view(np.uint32) & np.uint32(0xFFFF0000)
        return u.view(np.float32).astype(np.float64)

    monkeypatch.setattr(pymlp, "bf16", truncate)
    _fails(pins.test_bf16_rounding)


def test_gelu16h_pin_catches_unrounded_product(monkeypatch):
    def unrounded(z):
        x = pymlp.fp16(z)
        return x * np.clip(pymlp._phi16_raw(x), 0.0, 1.0)

    monkeypatch.setattr(pymlp, "gelu16h", unrounded)
    _fails(pins.test_gelu16h_matches_normal_cdf_on_every_fp16_input)


def test_gelu16h_pin_catches_bf16_product(monkeypatch):
    def to_bf16(z):
        x = pymlp.fp16(z)
        return pymlp.bf16(x * np.clip(pymlp._phi16_raw(x), 0.0, 1.0))

    monkeypatch.setattr(pymlp, "gelu16h", to_bf16)
    _fails(pins.test_gelu16h_matches_normal_cdf_on_every_fp16_input)


def test_eval_forward_pin_catches_bf16_operands(monkeypatch):
    # the evaluation form with the training form's bf16 operand rounding
    monkeypatch.setattr(pymlp, "fp16", pymlp.bf16)
    _fails(pins.test_quantised_forward_close_to_exact)
