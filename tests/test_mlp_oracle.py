"""Pins of the FP64 MLP oracle (CPU)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_oracle as mo  # noqa: E402


def test_gelu_closed_form():
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mo.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    z = np.linspace(-4, 4, 101)
    fd = (mo.gelu(z + 1e-6) - mo.gelu(z - 1e-6)) / 2e-6
    assert np.allclose(mo.gelu_grad(z), fd, atol=1e-8)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    n_in, h, n_out, b = 7, 5, 3, 4
    n = n_in * h + h + h * h + h + h * n_out + n_out
    p = rng.normal(size=n) * 0.5
    x = rng.normal(size=(b, n_in))
    t = rng.normal(size=(b, n_out))
    loss, g = mo.loss_and_grads(p, x, t, n_in, h, n_out)
    eps = 1e-6
    for k in range(0, n, 3):
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (mo.loss_and_grads(pp, x, t, n_in, h, n_out)[0] -
              mo.loss_and_grads(pm, x, t, n_in, h, n_out)[0]) / (2 * eps)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(fd)), k


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, 3.0e-5, -2.5])
    r = mo.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert r[4] == -2.5


def test_adam_first_step_is_signed_lr():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    p1, m, v = mo.adam(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-7)
