"""Surrogate kernels (BF16 MFMA, FP32 accumulation) against the FP64 oracle."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_oracle as mo  # noqa: E402
from bd_aging_amd.surrogate import MLPEnsemble  # noqa: E402

pytestmark = pytest.mark.gpu
N_IN, H, N_OUT = 387, 256, 32


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def test_forward_matches_oracle():
    rng = np.random.default_rng(1)
    e = MLPEnsemble(3, N_IN, H, N_OUT, seed=5)
    x = rng.normal(size=(70, N_IN)).astype(np.float32)  # not a multiple of 32
    y = e.forward(x)
    p = e.params()
    for m in range(3):
        yo, _ = mo.forward(p[m].astype(np.float64), x, N_IN, H, N_OUT, bf=True)
        assert rel(y[m], yo) < 2e-3, m


def test_gradients_match_oracle():
    rng = np.random.default_rng(2)
    e = MLPEnsemble(2, N_IN, H, N_OUT, seed=7)
    x = rng.normal(size=(64, N_IN)).astype(np.float32)
    t = rng.uniform(0, 1, size=(64, N_OUT)).astype(np.float32)
    g, loss = e.gradients(x, t)
    p = e.params()
    sizes = [n for _, n in [(a, int(np.prod(s))) for a, s in mo.shapes(N_IN, H, N_OUT)]]
    for m in range(2):
        lo, go = mo.loss_and_grads(p[m].astype(np.float64), x, t, N_IN, H, N_OUT, bf=True)
        assert abs(loss[m] - lo) < 1e-3 * lo
        k = 0
        for n in sizes:  # every tensor separately
            assert rel(g[m, k:k + n], go[k:k + n]) < 3e-2, (m, k)
            k += n


def test_gradients_from_the_gpu_intermediates():
    """Each stage of forward and backward against FP64 from the GPU's own
    inputs to that stage (bd_mlp_trace), so only the accumulation order
    differs: pre-activations and the FP32 pre-activation gradients within
    1e-5 (a slip of 1e-3 in GELU' fails), the BF16 operands the correctly
    rounded copies of the FP32 values, and every parameter gradient within
    1e-4 of the oracle's products of the GPU's operands."""
    rng = np.random.default_rng(4)
    B = 64
    e = MLPEnsemble(2, N_IN, H, N_OUT, seed=9)
    x = rng.normal(size=(B, N_IN)).astype(np.float32)
    t = rng.uniform(0, 1, size=(B, N_OUT)).astype(np.float32)
    g, loss = e.gradients(x, t)
    p = e.params()
    for m in range(2):
        tr = e.trace(m, B)
        pm = p[m].astype(np.float64)
        st = mo.stages_from_trace(pm, x, tr, N_IN, H, N_OUT)
        for k in ("z1", "z2", "y", "dz2", "dz1"):
            assert rel(tr[k], st[k]) < 1e-5, (m, k)
        # the activations are BF16 roundings of the GPU's GELU (one BF16 ulp at
        # most from rounding the FP64 GELU of the same z)
        for hk, zk in (("h1", "z1"), ("h2", "z2")):
            ref = mo.gelu(tr[zk])
            # (plus FP32 cancellation in 1 + erf for very negative z, ~1e-7 absolute)
            assert np.all(np.abs(tr[hk] - ref) <= 2.0 ** -7 * np.abs(ref) + 1e-6), (m, hk)
        # dy = 2 (y - t) / n in FP32, and the BF16 operands round the FP32 values
        y32, t32 = tr["y"].astype(np.float32), t.astype(np.float32)
        dy32 = np.float32(2.0 / (B * N_OUT)) * (y32 - t32)
        assert np.array_equal(tr["dy"], dy32.astype(np.float64))
        for fk, bk in (("dy", "dyb"), ("dz2", "dz2b"), ("dz1", "dz1b")):
            assert np.array_equal(tr[bk], mo.bf16(tr[fk])), (m, bk)
        go = mo.grads_from_trace(x, tr, N_IN, H, N_OUT)
        k = 0
        for name, shp in mo.shapes(N_IN, H, N_OUT):
            n = int(np.prod(shp))
            assert rel(g[m, k:k + n], go[k:k + n]) < 1e-4, (m, name)
            k += n
        assert abs(loss[m] - np.mean((tr["y"] - t) ** 2)) < 1e-5 * loss[m]


def test_adam_steps_match_oracle():
    """Two bd_mlp_train_step calls against the oracle's Adam on the GPU's own
    gradients (FP64 moments): every parameter within 1e-6 of the step size
    plus the FP32 spacing of the parameters."""
    rng = np.random.default_rng(6)
    e = MLPEnsemble(2, N_IN, H, N_OUT, seed=13)
    x = rng.normal(size=(32, N_IN)).astype(np.float32)
    t = rng.uniform(0, 1, size=(32, N_OUT)).astype(np.float32)
    lr = 1e-3
    p0 = e.params().astype(np.float64)
    g1, _ = e.gradients(x, t)
    e.train_step(x, t, lr)
    p1 = e.params().astype(np.float64)
    g2, _ = e.gradients(x, t)
    e.train_step(x, t, lr)
    p2 = e.params().astype(np.float64)
    for m in range(2):
        zero = np.zeros_like(p0[m])
        q1, m1, v1 = mo.adam(p0[m], g1[m].astype(np.float64), zero, zero, 1, lr)
        q2, _, _ = mo.adam(p1[m], g2[m].astype(np.float64), m1, v1, 2, lr)
        tol1 = 1e-6 * lr + float(np.spacing(np.float32(np.abs(q1).max())))
        tol2 = 1e-6 * lr + float(np.spacing(np.float32(np.abs(q2).max())))
        assert np.abs(p1[m] - q1).max() <= tol1
        assert np.abs(p2[m] - q2).max() <= tol2
        assert np.abs(p2[m] - p1[m]).max() > 0.5 * lr  # the second step moved


def test_fallback_gate_sends_out_of_distribution_points_to_bd():
    """Points far outside the training inputs make the 5 members disagree by
    more than the 0.03 gate, and evaluate() answers them with BD itself;
    points like the training ones pass the gate."""
    from bd_aging_amd import surrogate as sg
    rng = np.random.default_rng(7)

    def points(n, scale=1.0, shift=0.0):
        # a g(r)-like block that depends smoothly on the state, with a little
        # noise; out of distribution: the block stretched and shifted far away
        X = np.empty((n, sg.N_FEATURES))
        T = rng.uniform(0.05, 0.5, n)
        phi = rng.uniform(0.52, 0.60, n)
        r = np.arange(3 * sg.N_BINS) % sg.N_BINS * (sg.G_RANGE / sg.N_BINS)
        g = 1.0 + (1.5 + 10.0 * (phi[:, None] - 0.56)) * np.exp(-3.0 * (r - 1.2) ** 2) * \
            np.cos(4.0 * r + T[:, None])
        X[:, :3 * sg.N_BINS] = shift + scale * (g + 0.01 * rng.normal(size=g.shape))
        X[:, -3] = T
        X[:, -2] = phi
        X[:, -1] = rng.uniform(1.0, 4.0, n)    # log10 t_w
        return X

    def curves(X):
        rate = 2.0 * X[:, -3] / (X[:, -2] * 10.0 ** (0.2 * X[:, -1]))
        return np.exp(-np.outer(rate, sg.LAGS)).astype(np.float32)

    Xtr = points(8192)
    sur = sg.Surrogate(seed=5)
    sur.fit(Xtr, curves(Xtr), epochs=300)
    Xin = points(256)
    _, std_in, fb_in = sur.predict(Xin)
    Xood = points(256, scale=6.0, shift=3.0)
    rep = sg.evaluate(sur, Xood, curves(Xood), 1.0)
    _, std_ood, fb_ood = sur.predict(Xood)
    print(f"[gate] in: {fb_in.mean():.3f} fall back (max std {std_in.max():.4f}); "
          f"out: {fb_ood.mean():.3f} (median std {np.median(std_ood.max(1)):.3f})", flush=True)
    assert fb_in.mean() < 0.1
    assert fb_ood.mean() > 0.9
    assert rep["fallback_fraction"] == fb_ood.mean()
    assert rep["rms_with_fallback"] < rep["rms_all"]
    sur.net.close()


def test_training_reduces_loss_and_members_differ():
    rng = np.random.default_rng(3)
    n = 2048
    x = rng.normal(size=(n, N_IN)).astype(np.float32)
    w = rng.normal(size=(N_IN,)) / np.sqrt(N_IN)
    lags = np.linspace(0, 1, N_OUT)
    t = np.exp(-np.outer(np.exp(x @ w), lags)).astype(np.float32)  # decaying curves
    e = MLPEnsemble(5, N_IN, H, N_OUT, seed=11)
    losses = e.train(x, t, epochs=15, batch=128, lr=1e-3, seed=4)
    assert losses[-1].max() < 0.2 * losses[0].min()
    mean, std = e.predict(x[:256])
    assert np.sqrt(np.mean((mean - t[:256]) ** 2)) < 0.08
    assert std.mean() > 0  # independent initialisations


def test_dataset_generation_small():
    from bd_aging_amd import surrogate as sg
    X, Y, cost, wall = sg.generate(R=4, n_range=(400, 600), t_w_steps=(10, 20, 40, 80, 160, 320,
                                                                      640, 1280),
                                   seed=3, equil_steps=200)
    assert X.shape == (32, sg.N_FEATURES) and Y.shape == (32, sg.N_LAGS)
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(Y))
    # F_s starts near 1 at the shortest lag and does not grow on average
    assert Y[:, 0].min() > 0.8
    assert np.mean(Y[:, -1]) < np.mean(Y[:, 0])
    g = X[:, :3 * sg.N_BINS].reshape(-1, 3, sg.N_BINS)
    r = (np.arange(sg.N_BINS) + 0.5) * sg.G_RANGE / sg.N_BINS
    assert np.all(g[:, :, r < 0.7] == 0)          # no overlaps
    assert g[:, 2, :].max() > 1.5                 # a first peak
    assert cost > 0


def test_surrogate_gain_over_simulation():
    """What the trained network gains over simulating: trained only on batched-BD
    data, it predicts held-out F_s(k=7, t_w, t) curves of unseen state points
    (N, phi, T, seed) much closer than the training curves' mean at the same t_w
    (the best guess without the network), at >= 100x lower cost than the BD
    that measured them; points the ensemble disagrees on go back to BD."""
    from bd_aging_amd import surrogate as sg
    runs = [sg.generate(R=64, n_range=(3000, 6000), seed=500 + k) for k in range(6)]
    Xtr = np.concatenate([r[0] for r in runs[:5]])
    Ytr = np.concatenate([r[1] for r in runs[:5]])
    Xte, Yte, _, wall = runs[5]
    sur = sg.Surrogate(seed=3)
    sur.fit(Xtr, Ytr, epochs=300)
    rep = sg.evaluate(sur, Xte, Yte, wall / len(Xte))
    # without the network: the mean training curve at the same waiting time
    tw = Xtr[:, -1]
    naive = np.stack([Ytr[tw == v].mean(0) for v in Xte[:, -1]])
    rms_naive = float(np.sqrt(np.mean((naive - Yte) ** 2)))
    print(f"[surrogate gain] rms {rep['rms_all']:.4f} vs naive {rms_naive:.4f}, "
          f"fallback {rep['fallback_fraction']:.3f}, cost ratio "
          f"{rep['cost_ratio_bd_over_surrogate']:.0f}x", flush=True)
    assert rep["rms_all"] < 0.6 * rms_naive
    assert rep["rms_all"] < 0.04
    assert rep["cost_ratio_bd_over_surrogate"] >= 100
    assert rep["rms_with_fallback"] <= rep["rms_all"] + 1e-12
    sur.net.close()
