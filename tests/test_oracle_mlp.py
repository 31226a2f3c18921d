"""Pins the MLP reference (oracle/mlp_oracle.py) used to check the HIP
forward/backward kernels of the warm start.

* The analytic FP64 gradients match central finite differences of the FP64
  loss (relative 1e-6): a dropped GELU' factor or a transposed product fails.
* The bf16-rounded gradients (the kernels' arithmetic) stay within bf16
  resolution of the exact ones.
* bf16 rounding is round-to-nearest-even on known bit patterns.
* GELU and its derivative against the closed form and a finite difference.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _params(rng, scale=0.3):
    return {"W1": rng.normal(0, scale, (8, 128)), "b1": rng.normal(0, 0.1, 128),
            "W2": rng.normal(0, scale / 4, (128, 128)), "b2": rng.normal(0, 0.1, 128),
            "W3": rng.normal(0, scale / 4, (128, 7)), "b3": rng.normal(0, 0.1, 7)}


def test_bf16_round_to_nearest_even():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert mo.bf16(np.array([1.0 + 2 ** -8]))[0] == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6 -> 1 + 2^-6 (even)
    assert mo.bf16(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6
    assert mo.bf16(np.array([1.0 + 2 ** -8 + 2 ** -12]))[0] == 1.0 + 2 ** -7
    assert mo.bf16(np.array([-3.140625]))[0] == -3.140625


def test_gelu_derivative():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.allclose(mo.gelu_grad(z), fd, rtol=1e-7, atol=1e-9)
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mo.gelu(np.array([10.0]))[0] - 10.0) < 1e-12


def test_exact_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(rng)
    X = rng.normal(0, 1, (32, 8))
    T = rng.normal(0, 1, (32, 7))
    g = mo.grads_exact(p, X, T)
    h = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = mo.loss_exact(p, X, T)
            flat[idx] = old - h
            lm = mo.loss_exact(p, X, T)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            an = g[k].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-3), (k, idx, fd, an)


def test_rounded_gradients_close_to_exact():
    rng = np.random.default_rng(2)
    p = _params(rng)
    X = rng.normal(0, 1, (256, 8))
    T = rng.normal(0, 1, (256, 7))
    _, g = mo.loss_and_grads(p, X, T)
    ge = mo.grads_exact(p, X, T)
    for k in p:
        err = np.linalg.norm(g[k] - ge[k]) / np.linalg.norm(ge[k])
        assert err < 3e-2, (k, err)


def test_transposed_product_is_detected():
    """A planted mistake (W2 used untransposed in the backward) changes the
    gradient far beyond bf16 resolution."""
    rng = np.random.default_rng(3)
    p = _params(rng)
    X = rng.normal(0, 1, (64, 8))
    T = rng.normal(0, 1, (64, 7))
    ge = mo.grads_exact(p, X, T)
    Z1 = X @ p["W1"] + p["b1"]
    H1 = mo.gelu(Z1)
    Z2 = H1 @ p["W2"] + p["b2"]
    H2 = mo.gelu(Z2)
    Y = H2 @ p["W3"] + p["b3"]
    dY = 2.0 * (Y - T) / Y.size
    dZ2 = (dY @ p["W3"].T) * mo.gelu_grad(Z2)
    wrong = (dZ2 @ p["W2"]) * mo.gelu_grad(Z1)        # should be W2.T
    gW1_wrong = X.T @ wrong
    assert np.linalg.norm(gW1_wrong - ge["W1"]) > 0.1 * np.linalg.norm(ge["W1"])
