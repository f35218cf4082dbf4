"""Learned surrogate of the self-intermediate scattering function F_s(k, t_w, t).

Input (387 features): the three partial g(r) (big-big, big-small, small-small;
128 bins to r = 5 sigma_sml, the reference's normalisation, main.cu:849-877)
at the waiting time t_w after the quench, the quench temperature T, the
packing fraction phi and log10(t_w).  Output: F_s(k = 7.0) of the big species
at 32 log-spaced lags.  The training data come only from the batched BD
engine (Engine with one replica per state point, F_s against up to 8
waiting-time origins per replica in one launch).

Lag grid: 32 log-spaced lags in [1e-2, 2] time units (100 to 2e4 steps of
dt = 1e-4).  The requested [1e-2, 1e3] would need runs 500x longer than the
batched ensemble's 2e5 steps, so the grid is capped at what the data cover.

The network (bd_mlp_*: BF16 matrix-core kernels, FP32 accumulation) is an
ensemble of 5 independently initialised members; a state point whose
ensemble standard deviation exceeds `std_gate` at any lag falls back to
plain BD, which stays the yardstick of the surrogate's error.
"""
import ctypes as C
import math
import time

import numpy as np

from . import abi
from .abi import check, load
from .engine import Engine, make_params, box_side

N_BINS = 128
G_RANGE = 5.0
K_FS = 7.0
N_LAGS = 32
LAGS = np.logspace(-2, np.log10(2.0), N_LAGS)  # time units
N_FEATURES = 3 * N_BINS + 3


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLPEnsemble:
    """members x (n_in -> hidden -> hidden -> n_out) MLPs on the GPU."""

    def __init__(self, members=5, n_in=N_FEATURES, hidden=256, n_out=N_LAGS, seed=1):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.bd_mlp_create(members, n_in, hidden, n_out, int(seed), C.byref(h)))
        self.h = h
        self.members, self.n_in, self.hidden, self.n_out = members, n_in, hidden, n_out
        n = C.c_int64()
        check(self.lib.bd_mlp_num_params(self.h, C.byref(n)))
        self.nparam = n.value

    def close(self):
        if self.h:
            self.lib.bd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros((self.members, self.nparam), dtype=np.float32)
        check(self.lib.bd_mlp_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(self.members, self.nparam)
        check(self.lib.bd_mlp_set_params(self.h, _fp(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((self.members, len(x), self.n_out), dtype=np.float32)
        check(self.lib.bd_mlp_forward(self.h, _fp(x), len(x), _fp(y)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros((self.members, self.nparam), dtype=np.float32)
        loss = np.zeros(self.members, dtype=np.float32)
        check(self.lib.bd_mlp_gradients(self.h, _fp(x), _fp(t), len(x), _fp(g), _fp(loss)))
        return g, loss

    def trace(self, member, batch):
        """Intermediates of the last gradients/train_step call (test hook):
        dict of z1, h1, z2, h2 (batch x hidden), y, dy, dyb (batch x n_out),
        dz2, dz2b, dz1, dz1b (batch x hidden)."""
        bh, bo = batch * self.hidden, batch * self.n_out
        out = np.zeros(8 * bh + 3 * bo, dtype=np.float32)
        check(self.lib.bd_mlp_trace(self.h, int(member), int(batch), _fp(out)))
        names = [("z1", bh), ("h1", bh), ("z2", bh), ("h2", bh), ("y", bo), ("dy", bo),
                 ("dyb", bo), ("dz2", bh), ("dz2b", bh), ("dz1", bh), ("dz1b", bh)]
        res, k = {}, 0
        for name, n in names:
            cols = self.n_out if n == bo else self.hidden
            res[name] = out[k:k + n].reshape(batch, cols).astype(np.float64)
            k += n
        return res

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(self.members, dtype=np.float32)
        check(self.lib.bd_mlp_train_step(self.h, _fp(x), _fp(t), len(x), float(lr), _fp(loss)))
        return loss

    def train(self, x, t, epochs=50, batch=256, lr=1e-3, seed=1):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        check(self.lib.bd_mlp_set_data(self.h, _fp(x), _fp(t), len(x)))
        losses = np.zeros((epochs, self.members), dtype=np.float32)
        check(self.lib.bd_mlp_train(self.h, int(epochs), int(batch), float(lr), int(seed),
                                    _fp(losses)))
        return losses

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        mean = np.zeros((len(x), self.n_out), dtype=np.float32)
        std = np.zeros((len(x), self.n_out), dtype=np.float32)
        check(self.lib.bd_mlp_predict(self.h, _fp(x), len(x), _fp(mean), _fp(std)))
        return mean, std


# ---------------------------------------------------------------- data
def normalized_gr(counts, p, nbins=N_BINS, rng=G_RANGE):
    """main.cu:849-877 normalisation of counts (R, 4, nbins) for params list p."""
    bs = rng / nbins
    nb = np.arange(nbins, dtype=np.float64)
    shell = 4.0 * 3.1415927410125732 * bs ** 3 * (1.0 / 3.0 + nb * nb + nb)
    out = np.zeros((len(p), 3, nbins))
    for r, q in enumerate(p):
        vol = q.side ** 3
        xb, xs = float(q.n_big), float(q.n_sml)
        sbs = 0.5 * (q.sigma_big + q.sigma_sml)
        vf = lambda s: vol - 4.0 / 3.0 * 3.1415927410125732 * s ** 3  # noqa: E731
        out[r, 0] = counts[r, 0] / shell / (xb * (xb - 1.0) / vf(q.sigma_big))
        out[r, 1] = 0.5 * (counts[r, 1] + counts[r, 2]) / shell / (xb * xs / vf(sbs))
        out[r, 2] = counts[r, 3] / shell / (xs * (xs - 1.0) / vf(q.sigma_sml))
    return out


def generate(R=256, n_range=(1000, 16000), phi_range=(0.52, 0.60), T_range=(0.05, 0.5),
             t_w_steps=(10, 30, 100, 300, 1000, 3000, 10000, 20000), dt=1e-4, seed=1,
             equil_steps=2000, relaxed=None, log=None):
    """One batched BD run: R replicas of uneven N, phi and quench T (the
    batched-ensemble configuration), F_s(k=7) of every replica against 8
    waiting-time origins at every lag, g(r) at every waiting time.
    Returns X (R*8, 387), Y (R*8, 32), and the run cost in particle-steps."""
    from .init import relax_batch
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), R))).astype(int)
    phi = rng.uniform(*phi_range, R)
    T = rng.uniform(*T_range, R)
    specs, params = [], []
    for r in range(R):
        nb = int(n[r]) // 2
        ns = int(n[r]) - nb
        side = box_side(nb, ns, 1.4, 1.0, float(phi[r]))
        specs.append((nb, ns, side, 1.4, 1.0, int(seed) * 100003 + r))
        params.append(make_params(nb, ns, side=side, dt=dt, T0=1.0, Tf=float(T[r]), t0=0.0,
                                  tf=0.0, first_step=-equil_steps,
                                  seed=int(seed) * 7919 + r, ntags_big=12, ntags_sml=12))
    xs = relax_batch(specs) if relaxed is None else relaxed
    lag_steps = np.unique(np.rint(LAGS / dt).astype(int))
    assert len(lag_steps) == N_LAGS
    t_w = np.array(t_w_steps)
    events = {}
    for a, tw in enumerate(t_w):
        events.setdefault(int(tw), []).append(("origin", a))
        for j, L in enumerate(lag_steps):
            events.setdefault(int(tw + L), []).append(("fs", a, j))
    fs = np.zeros((R, len(t_w), N_LAGS))
    gr = np.zeros((R, len(t_w), 3, N_BINS))
    with Engine(params) as eng:
        for r, x in enumerate(xs):
            eng.set_positions(x, r)
        eng.step(equil_steps)  # at T = T0 = 1 (transient)
        eng.mark_origin()
        t0 = time.perf_counter()
        done = 0  # steps after the quench
        for at in sorted(events):
            eng.step(at - done)
            done = at
            evs = events[at]
            if any(e[0] == "origin" for e in evs):
                for e in evs:
                    if e[0] == "origin":
                        eng.set_origin(e[1])
                        c = eng.gder_counts_all(N_BINS, G_RANGE).astype(np.float64)
                        gr[:, e[1]] = normalized_gr(c, params)
            f = [e for e in evs if e[0] == "fs"]
            if f:
                res = eng.fself_origins(K_FS, len(t_w))
                for _, a, j in f:
                    fs[:, a, j] = res[:, a, 0]
        wall = time.perf_counter() - t0
        if log:
            log(f"[surrogate] run seed={seed}: {done} steps, sum N={int(n.sum())}, {wall:.1f} s")
    X = np.zeros((R, len(t_w), N_FEATURES))
    X[:, :, :3 * N_BINS] = gr.reshape(R, len(t_w), -1)
    X[:, :, 3 * N_BINS] = T[:, None]
    X[:, :, 3 * N_BINS + 1] = phi[:, None]
    X[:, :, 3 * N_BINS + 2] = np.log10(t_w * dt)[None, :]
    cost = float(n.sum()) * done
    return X.reshape(-1, N_FEATURES), fs.reshape(-1, N_LAGS), cost, wall


class Surrogate:
    """Standardised inputs -> 5-member ensemble; std gate for the BD fallback."""

    def __init__(self, members=5, hidden=256, seed=1, std_gate=0.03):
        self.net = MLPEnsemble(members, N_FEATURES, hidden, N_LAGS, seed)
        self.std_gate = std_gate
        self.mu = None
        self.sd = None

    def fit(self, X, Y, epochs=150, batch=256, lr=1e-3, seed=1):
        """Adam in four phases of epochs/4 with the rate stepped down
        (lr, 0.3 lr, 0.1 lr, 0.03 lr); returns the per-epoch losses."""
        self.mu = X.mean(0)
        # The g(r) block is scaled as a whole (one scale for all 3 x 128 bins),
        # so a bin that barely varies (below contact, or noise at small N)
        # stays a small input instead of being blown up to unit variance; T,
        # phi and log10(t_w) are standardised one by one.
        ng = 3 * N_BINS
        self.sd = np.empty(X.shape[1])
        self.sd[:ng] = max(float((X[:, :ng] - self.mu[:ng]).std()), 1e-6)
        self.sd[ng:] = np.maximum(X[:, ng:].std(0), 1e-6)
        Xn = self.standardise(X)
        phases = [max(1, epochs // 4)] * 3
        phases.append(max(1, epochs - sum(phases)))
        out = []
        for k, (ep, f) in enumerate(zip(phases, (1.0, 0.3, 0.1, 0.03))):
            out.append(self.net.train(Xn, Y, ep, batch, lr * f, seed + 7919 * k))
        return np.concatenate(out)

    def standardise(self, X):
        return np.clip((X - self.mu) / self.sd, -30.0, 30.0)

    def predict(self, X):
        """Ensemble mean, std and the fallback mask (True: run plain BD)."""
        mean, std = self.net.predict(self.standardise(X))
        return mean, std, std.max(1) > self.std_gate


def evaluate(sur, X, Y, bd_seconds_per_sample):
    """RMS of the ensemble mean against BD-measured F_s, with and without the
    std gate, and the cost ratio of simulating versus predicting."""
    t0 = time.perf_counter()
    mean, std, fallback = sur.predict(X)
    pred_s = (time.perf_counter() - t0) / len(X)
    err = mean - Y
    rms_all = float(np.sqrt(np.mean(err ** 2)))
    keep = ~fallback
    rms_kept = float(np.sqrt(np.mean(err[keep] ** 2))) if keep.any() else float("nan")
    # the gated answer: surrogate where trusted, BD itself where it falls back
    gated = np.where(fallback[:, None], Y, mean)
    rms_gated = float(np.sqrt(np.mean((gated - Y) ** 2)))
    return {"rms_all": rms_all, "rms_accepted": rms_kept, "rms_with_fallback": rms_gated,
            "fallback_fraction": float(fallback.mean()), "samples": int(len(X)),
            "predict_seconds_per_sample": pred_s,
            "bd_seconds_per_sample": bd_seconds_per_sample,
            "cost_ratio_bd_over_surrogate": bd_seconds_per_sample / pred_s,
            "cost_ratio_with_fallback": bd_seconds_per_sample /
            (pred_s + float(fallback.mean()) * bd_seconds_per_sample)}


def baselines(Xtr, Ytr, Xte, Yte):
    """RMS on the held-out points of two predictors that use the state
    variables only: the mean training curve at the same waiting time, and a
    least-squares cubic polynomial in (T, phi, log10 t_w) (all 20 monomials of
    degree <= 3, one fit per lag)."""
    ng = 3 * N_BINS
    tw_tr, tw_te = np.round(Xtr[:, ng + 2], 9), np.round(Xte[:, ng + 2], 9)
    mean_pred = np.empty_like(Yte)
    for tw in np.unique(tw_te):
        mean_pred[tw_te == tw] = Ytr[tw_tr == tw].mean(0)
    mu, sd = Xtr[:, ng:].mean(0), Xtr[:, ng:].std(0)

    def design(X):
        v = (X[:, ng:] - mu) / sd
        cols = [np.ones(len(X))]
        for i in range(3):
            cols.append(v[:, i])
        for i in range(3):
            for j in range(i, 3):
                cols.append(v[:, i] * v[:, j])
        for i in range(3):
            for j in range(i, 3):
                for k in range(j, 3):
                    cols.append(v[:, i] * v[:, j] * v[:, k])
        return np.stack(cols, 1)

    coef, *_ = np.linalg.lstsq(design(Xtr), Ytr, rcond=None)
    cubic_pred = design(Xte) @ coef
    return {"rms_mean_curve_at_tw": float(np.sqrt(np.mean((mean_pred - Yte) ** 2))),
            "rms_cubic_in_T_phi_logtw": float(np.sqrt(np.mean((cubic_pred - Yte) ** 2)))}


def main(argv=None):
    import argparse
    import json
    import os
    import sys
    ap = argparse.ArgumentParser(description="train and evaluate the F_s surrogate")
    ap.add_argument("--runs", type=int, default=20)
    ap.add_argument("--test-runs", type=int, default=2)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--nmin", type=int, default=1000)
    ap.add_argument("--nmax", type=int, default=16000)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--out", default="surrogate_out")
    ap.add_argument("--save-data", action="store_true",
                    help="also write the generated data set (dataset.npz) to --out")
    ap.add_argument("--load-data", default=None,
                    help="train on a dataset.npz written by --save-data instead of simulating")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)

    def log(m):
        print(m, flush=True)

    if a.load_data:
        d = np.load(a.load_data)
        Xtr, Ytr, Xte, Yte = d["Xtr"], d["Ytr"], d["Xte"], d["Yte"]
        walls, costs = list(d["walls"]), list(d["costs"])
        a.runs, a.test_runs = int(d["runs"]), int(d["test_runs"])
    else:
        xs, ys, walls, costs = [], [], [], []
        for k in range(a.runs + a.test_runs):
            X, Y, cost, wall = generate(R=a.replicas, n_range=(a.nmin, a.nmax), seed=1000 + k,
                                        log=log)
            xs.append(X)
            ys.append(Y)
            walls.append(wall)
            costs.append(cost)
        Xtr, Ytr = np.concatenate(xs[:a.runs]), np.concatenate(ys[:a.runs])
        Xte, Yte = np.concatenate(xs[a.runs:]), np.concatenate(ys[a.runs:])
        if a.save_data:
            np.savez_compressed(os.path.join(a.out, "dataset.npz"), Xtr=Xtr.astype(np.float32),
                                Ytr=Ytr.astype(np.float32), Xte=Xte.astype(np.float32),
                                Yte=Yte.astype(np.float32), walls=np.array(walls),
                                costs=np.array(costs), runs=a.runs, test_runs=a.test_runs)
    log(f"[surrogate] train {len(Xtr)} samples, test {len(Xte)}")
    sur = Surrogate()
    t0 = time.perf_counter()
    losses = sur.fit(Xtr, Ytr, epochs=a.epochs)
    train_s = time.perf_counter() - t0
    log(f"[surrogate] trained in {train_s:.1f} s, final losses {losses[-1]}")
    bd_per_sample = sum(walls[a.runs:]) / len(Xte)
    rep = evaluate(sur, Xte, Yte, bd_per_sample)
    rep.update(baselines(Xtr, Ytr, Xte, Yte))
    rep.update({"train_samples": int(len(Xtr)), "epochs": a.epochs, "train_seconds": train_s,
                "final_train_loss": losses[-1].tolist(), "lags": LAGS.tolist(), "k": K_FS,
                "members": 5, "std_gate": sur.std_gate,
                "bd_particle_steps_per_s": sum(costs) / sum(walls)})
    log(json.dumps(rep))
    json.dump(rep, open(os.path.join(a.out, "report.json"), "w"), indent=1)
    np.savez(os.path.join(a.out, "surrogate_model.npz"), params=sur.net.params(), mu=sur.mu,
             sd=sur.sd, lags=LAGS)
    np.savez_compressed(os.path.join(a.out, "test_curves.npz"), Y=Yte.astype(np.float32),
                        P=sur.predict(Xte)[0])
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
