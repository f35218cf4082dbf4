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
