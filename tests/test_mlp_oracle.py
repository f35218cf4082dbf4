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


def test_restatements_from_a_trace_match_the_pinned_fp64_pipeline():
    """With inputs and weights exact in bfloat16, the FP64 pipeline's own
    intermediates fed to stages_from_trace / grads_from_trace give back its
    stages and its (finite-difference pinned) gradients: the formulas the
    GPU check uses multiply the right operands in the right orientation."""
    rng = np.random.default_rng(5)
    n_in, h, n_out, b = 9, 6, 4, 5
    n = n_in * h + h + h * h + h + h * n_out + n_out
    p = mo.bf16(rng.normal(size=n) * 0.5)
    x = mo.bf16(rng.normal(size=(b, n_in)))
    t = rng.normal(size=(b, n_out))
    loss, g, tr = mo.loss_and_grads(p, x, t, n_in, h, n_out, trace=True)
    st = mo.stages_from_trace(p, x, tr, n_in, h, n_out)
    for k in ("z1", "z2", "y", "dz2", "dz1"):
        assert np.allclose(st[k], tr[k], rtol=1e-13, atol=1e-13), k
    gt = mo.grads_from_trace(x, tr, n_in, h, n_out)
    assert np.allclose(gt, g, rtol=1e-13, atol=1e-13)
    # a swapped operand is seen: the W1 gradient from dz2 instead of dz1 differs
    bad = dict(tr, dz1b=tr["dz2b"])
    assert not np.allclose(mo.grads_from_trace(x, bad, n_in, h, n_out), g)


def test_adam_second_step_closed_form():
    # two steps on constant gradients: the bias corrections make the moment
    # estimates exact, so each step moves every parameter by lr g/(|g| + eps')
    p = np.array([0.5, -1.0])
    g = np.array([0.2, -3.0])
    p1, m, v = mo.adam(p, g, np.zeros(2), np.zeros(2), 1, 0.01)
    p2, _, _ = mo.adam(p1, g, m, v, 2, 0.01)
    step = 0.01 * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1, p - step, rtol=0, atol=1e-15)
    assert np.allclose(p2, p - 2 * step, rtol=0, atol=1e-14)
