"""Disorder averaging with the learned control variate.

The ensemble mean of an echo curve, E[y], is estimated as

    mu = mean_n(y - beta * g) + beta * mean_M(g)

where y are exact curves (the propagator) on n realisations, g the
surrogate's predictions (libspx's graph network) on the same n realisations,
and the second mean runs over M further geometries on which only the
surrogate is evaluated. For any beta fixed independently of those n + M
draws the estimator is unbiased; its variance is
Var(y - beta g) / n + beta^2 Var(g) / M, so a surrogate that tracks y lets
n shrink by Var(y) / Var(y - beta g) at the same error bar. The reference's
plain Monte-Carlo mean (src/ensemble/Ensemble.cpp, runEnsemble) is the
beta = 0 case and stays the fallback.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import spx


def coupling_matrix(L: int, lo, hi, j) -> np.ndarray:
    """Dense symmetric J[L, L] from a pair list (the graph's edge weights)."""
    J = np.zeros((L, L))
    lo = np.asarray(lo, dtype=np.int64)
    hi = np.asarray(hi, dtype=np.int64)
    if lo.size and (lo.max() >= L or hi.max() >= L):
        raise ValueError("pair index out of range")
    J[lo, hi] = j
    J[hi, lo] = j
    return J


class Model:
    """The device network with the scales a training set fixes: s_h the
    field spread, s_J the median nonzero |J_ij|, and the targets centred per
    curve point and divided by one common spread, so the network fits O(1)
    numbers whatever the observable."""

    def __init__(self, ctx: spx.Context, J, h, Y, seed: int = 0):
        J = np.asarray(J, dtype=float)
        h = np.asarray(h, dtype=float)
        Y = np.asarray(Y, dtype=float)
        nz = np.abs(J[J != 0])
        self.scale_j = float(np.median(nz)) if nz.size else 1.0
        sh = float(h.std())
        self.scale_h = sh if sh > 0 else 1.0
        self.y_mean = Y.mean(axis=0)
        sy = float(np.sqrt(Y.var(axis=0).mean()))
        self.y_scale = sy if sy > 0 else 1.0
        self.net = spx.Surrogate(ctx, Y.shape[1], seed, self.scale_h, self.scale_j)

    def normalise(self, Y) -> np.ndarray:
        return ((np.asarray(Y, dtype=float) - self.y_mean) / self.y_scale).astype(np.float32)

    def train(self, J, h, Y, epochs: int, batch: int, lr: float, seed: int = 0, decay: float = 1.0):
        """Adam on shuffled minibatches; the mean loss (normalised units) per
        epoch. `decay` multiplies the learning rate after every epoch."""
        J = np.asarray(J, dtype=float)
        h = np.asarray(h, dtype=float)
        T = self.normalise(Y)
        n = J.shape[0]
        if n < batch:
            raise ValueError("fewer training graphs than one batch")
        rng = np.random.default_rng(seed)
        self.net.set_training_set(J, h, T)
        history = []
        for _ in range(epochs):
            order = rng.permutation(n)
            for s in range(0, n - batch + 1, batch):
                order[s:s + batch] = np.sort(order[s:s + batch])
            history.append(self.net.train_epoch(order, batch, lr))
            lr *= decay
        return history

    def predict_dipolar(self, seed: int, first: int, count: int, nspins: int, **draw) -> np.ndarray:
        """Predicted curves on realisations first .. first+count-1 of the
        dipolar ensemble, drawn inside libspx (spx_dipolar_graphs)."""
        raw = self.net.predict_dipolar(seed, first, count, nspins, **draw).astype(np.float64)
        return raw * self.y_scale + self.y_mean

    def predict(self, J, h, chunk: int = 1024) -> np.ndarray:
        """Predicted curves for any number of graphs, in chunks of `chunk`."""
        J = np.asarray(J, dtype=float)
        h = np.asarray(h, dtype=float)
        raw = np.concatenate([self.net.predict(J[s:s + chunk], h[s:s + chunk]).astype(np.float64)
                              for s in range(0, J.shape[0], chunk)])
        return raw * self.y_scale + self.y_mean


def fit_beta(y, g) -> np.ndarray:
    """Per-point optimal coefficient Cov(y, g) / Var(g), to be fitted on
    realisations that the estimate does not use."""
    y = np.asarray(y, dtype=float)
    g = np.asarray(g, dtype=float)
    gc = g - g.mean(axis=0)
    var = (gc * gc).mean(axis=0)
    cov = ((y - y.mean(axis=0)) * gc).mean(axis=0)
    return np.where(var > 0, cov / np.where(var > 0, var, 1.0), 0.0)


@dataclass
class Estimate:
    mean: np.ndarray
    stderr: np.ndarray
    n_exact: int
    n_surrogate: int


def plain_estimate(y) -> Estimate:
    """Plain Monte-Carlo mean and standard error (sample variance, n - 1)."""
    y = np.asarray(y, dtype=float)
    n = y.shape[0]
    se = y.std(axis=0, ddof=1) / np.sqrt(n) if n > 1 else np.zeros(y.shape[1:])
    return Estimate(y.mean(axis=0), se, n, 0)


def cv_estimate(y, g, g_extra, beta) -> Estimate:
    """Control-variate mean: exact y with predictions g on the same n
    realisations, predictions g_extra on M independent ones."""
    y = np.asarray(y, dtype=float)
    g = np.asarray(g, dtype=float)
    g_extra = np.asarray(g_extra, dtype=float)
    beta = np.asarray(beta, dtype=float)
    n, M = y.shape[0], g_extra.shape[0]
    if n < 2 or M < 2:
        raise ValueError("need at least two exact and two surrogate-only realisations")
    r = y - beta * g
    mean = r.mean(axis=0) + beta * g_extra.mean(axis=0)
    var = r.var(axis=0, ddof=1) / n + beta * beta * g_extra.var(axis=0, ddof=1) / M
    return Estimate(mean, np.sqrt(var), n, M)


def variance_reduction(y, g, beta) -> np.ndarray:
    """Var(y) / Var(y - beta g) per point: the factor by which the exact
    sample shrinks at a fixed error bar when surrogate draws are free."""
    y = np.asarray(y, dtype=float)
    g = np.asarray(g, dtype=float)
    vy = y.var(axis=0, ddof=1)
    vr = (y - beta * g).var(axis=0, ddof=1)
    return np.where(vr > 0, vy / np.where(vr > 0, vr, 1.0), np.inf)


def timing(ctx: spx.Context, L: int = 12, graphs: int = 16384, epochs: int = 40, points: int = 4,
           draws: int = 1_000_000, seed: int = 20240726, eager: bool = False) -> dict:
    """The surrogate's kernels alone, at the control variate's shapes:
    training on `graphs` dipolar geometries of L spins (drawn inside libspx;
    the targets are a fixed smooth function of each geometry, since the
    kernels' cost does not depend on the values), then predictions on
    `draws` further geometries. No exact runs, so the kernel trace and the
    counters cover the network only. eager: train with one
    spx_gnn_train_step call per minibatch (its batch uploaded from the host,
    every launch issued by the host) instead of the epochs' replayed steps."""
    import time
    J, h = spx.dipolar_graphs(seed, 0, graphs, L)
    t = np.linspace(0.25, 1.0, points)
    Y = np.cos(np.abs(J).sum(axis=(1, 2))[:, None] * t[None, :] / L) + 0.1 * h.std(axis=1)[:, None]
    model = Model(ctx, J, h, Y, seed=11)
    ctx.synchronize()
    t0 = time.perf_counter()
    if eager:
        T = model.normalise(Y)
        rng = np.random.default_rng(1)
        hist, lr = [], 2e-3
        for _ in range(epochs):
            order = rng.permutation(graphs)
            losses = [model.net.train_step(J[order[s:s + 128]], h[order[s:s + 128]], T[order[s:s + 128]], lr)
                      for s in range(0, graphs - 127, 128)]
            hist.append(float(np.mean(losses)))
            lr *= 0.95
    else:
        hist = model.train(J, h, Y, epochs=epochs, batch=128, lr=2e-3, seed=1, decay=0.95)
    ctx.synchronize()
    t_train = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = model.predict_dipolar(seed, 10_000_000, draws, L)
    ctx.synchronize()
    t_pred = time.perf_counter() - t0
    return {"workload": f"surrogate alone: training on {graphs} L={L} geometries ({epochs} epochs, batches of 128, "
                        f"{points} curve points{', one host-issued step per batch' if eager else ''}), then {draws} "
                        f"predictions on further geometries",
            "train_s": round(t_train, 3), "steps": epochs * (graphs // 128), "loss_first": round(hist[0], 4),
            "loss_last": round(hist[-1], 4), "predict_s": round(t_pred, 3),
            "predictions_per_s": round(draws / t_pred, 1), "prediction_mean": [round(float(v), 5) for v in g.mean(axis=0)]}


if __name__ == "__main__":
    import json
    import sys
    args = sys.argv[1:]
    if not args or args[0] != "timing":
        raise SystemExit("usage: python -m spinsim_amd.surrogate timing [graphs] [epochs] [draws] [eager]")
    kw = {"eager": "eager" in args[1:]}
    for key, value in zip(("graphs", "epochs", "draws"), [a for a in args[1:] if a != "eager"]):
        kw[key] = int(value)
    print(json.dumps(timing(spx.Context(0), **kw)), flush=True)
