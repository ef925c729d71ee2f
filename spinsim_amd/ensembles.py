"""The disorder-averaged echo experiments the surrogate is built for, on the
device, end to end: pulse programs, exact curves from the batch runner,
surrogate training and validation, and the control-variate estimate against
plain Monte-Carlo.

Every realisation is the reference's runRealization (src/io/Experiment.cpp:10-55)
for a run with zero base fields, dipolar geometry, Gaussian Sz disorder and a
random initial state, drawn inside libspx from Rng::forRealization(seed,
index): 2D positions at unit density with the both-coordinate minimum spacing
0.5, secular couplings (1 - 3cos^2 theta)/r^3 on Sz Sz - (Sx Sx + Sy Sy)/2,
Gaussian hz of spread 0.5. The echo is corr:x (<ref| sum_i sigma_x(i) |psi>,
src/core/Observable.cpp:57-74, with ref = sum_i sigma_x(i) psi0) divided by
L, so that it starts at 1.

Two experiments:
  * the surrogate's training set: L = 14, the CPMG sequence (pi/2_x, then
    64 x [10 steps of 0.05, ideal pi_y, 10 steps of 0.05, measure]), 8192
    curves for training and 1024 held out;
  * the control-variate ensemble: L = 16, XY-16 (X Y X Y Y X Y X and its
    phase-inverted copy) with finite pulses of hx or hy = +-50 lasting
    pi / 50, one Chebyshev step of 0.05 of free evolution before each pulse,
    the echo measured after every four pulses; the surrogate is evaluated on
    1e6 random geometries, plus an exact correction on as few realisations
    as reach a +-1e-3 error bar on every point, against plain Monte-Carlo
    reaching the same error bar.

Usage: python -m spinsim_amd.ensembles training-set   (the full L = 14 set)
       python -m spinsim_amd.ensembles control-variate (the L = 16 estimate)
Each prints one JSON line.
"""
from __future__ import annotations

import json
import math
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import spx
from . import surrogate as S

SEED = 20240726
OMEGA = 50.0          # pulse field during a finite pulse
DRAW = dict(scale=1.0, min_separation=0.5, angular=True, disorder=0.5)


def cpmg_program(L: int, points: int = 64, steps: int = 10, dt: float = 0.05) -> spx.Program:
    """pi/2_x on every spin, then `points` x [steps x dt, ideal pi_y,
    steps x dt, measure]."""
    P = spx.Program()
    every = list(range(L))
    P.pulse((1, 0, 0), math.pi / 2, every)
    P.repeat(points)
    P.evolve(dt, steps)
    P.pulse((0, 1, 0), math.pi, every)
    P.evolve(dt, steps)
    P.measure()
    P.end()
    return P


XY16 = [("x", 1), ("y", 1), ("x", 1), ("y", 1), ("y", 1), ("x", 1), ("y", 1), ("x", 1),
        ("x", -1), ("y", -1), ("x", -1), ("y", -1), ("y", -1), ("x", -1), ("y", -1), ("x", -1)]


def xy16_program(blocks: int = 1, free: float = 0.05, omega: float = OMEGA, per_measure: int = 4) -> spx.Program:
    """`blocks` x XY-16: before each pulse one free Chebyshev step of `free`,
    each pulse an evolve of pi / omega under (+-omega, 0, 0) or (0, +-omega,
    0) on every spin (an evolve step's hx / hy keys; the unnamed hz is the
    base value 0, as the reference's run files give it); a measure after
    every `per_measure` pulses."""
    P = spx.Program()
    tp = math.pi / omega
    P.repeat(blocks)
    for k, (axis, sign) in enumerate(XY16):
        P.evolve(free, 1)
        hx, hy = (sign * omega, 0.0) if axis == "x" else (0.0, sign * omega)
        P.evolve_uniform(tp, 1, hx=hx, hy=hy, hz=0.0)
        if (k + 1) % per_measure == 0:
            P.measure()
    P.end()
    return P


def exact_curves(ctx: spx.Context, program: spx.Program, L: int, first: int, count: int,
                 seed: int = SEED, epsilon: float = 1e-7):
    """Echo curves corr_x / L [count, points] of realisations first ..
    first+count-1 by the batch runner (the LDS-resident launch up to L = 12,
    the streaming groups above), and the run's wall time. Couplings and
    disorder are drawn when the realisations are added; each initial state
    is drawn inside the run, on host threads, and timed with it."""
    b = spx.Batch(ctx, program, [(spx.OBS_CORR_X, 0)], epsilon=epsilon, reference_axis=0)
    b.add_dipolar(seed, first, count, L, **DRAW)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, v = b.run()
    dt = time.perf_counter() - t0
    if not np.isfinite(v).all():
        raise RuntimeError("non-finite echo from the batch runner")
    return v[:, :, 0] / L, dt


def graphs(L: int, first: int, count: int, seed: int = SEED):
    """The surrogate's inputs (J [count, L, L], h [count, L]) of the same realisations."""
    return spx.dipolar_graphs(seed, first, count, L, DRAW["scale"], DRAW["min_separation"], DRAW["angular"],
                              DRAW["disorder"])


# ------------------------------------------------------------ training set

def exact_curves_chunked(ctx, program, L, first, count, seed=SEED, chunk=2048, log=lambda m: None):
    """exact_curves in runs of `chunk` realisations (bounded host memory,
    a progress line per run); the summed run time."""
    parts, total = [], 0.0
    for s in range(0, count, chunk):
        y, dt = exact_curves(ctx, program, L, first + s, min(chunk, count - s), seed)
        parts.append(y)
        total += dt
        log(f"  {s + y.shape[0]} / {count} L={L} curves, {total:.1f} s")
    return np.concatenate(parts), total


def training_set(ctx: spx.Context, L: int = 14, n_train: int = 8192, n_val: int = 1024, points: int = 64,
                 epochs: int = 60, batch: int = 128, lr: float = 2e-3, seed: int = SEED,
                 log=lambda m: None) -> dict:
    """Exact CPMG curves of n_train + n_val realisations, the surrogate
    trained on the first n_train on the device, validated on the other
    n_val: per-point correlation and variance reduction on the held-out
    curves (Var(y) / Var(y - beta g), beta fitted on the training curves)."""
    P = cpmg_program(L, points)
    Ytr, t_tr = exact_curves_chunked(ctx, P, L, 0, n_train, seed, log=log)
    Yva, t_va = exact_curves_chunked(ctx, P, L, n_train, n_val, seed, log=log)
    Jtr, htr = graphs(L, 0, n_train, seed)
    Jva, hva = graphs(L, n_train, n_val, seed)
    model = S.Model(ctx, Jtr, htr, Ytr, seed=11)
    log(f"training on {n_train} curves, {epochs} epochs")
    t0 = time.perf_counter()
    hist = model.train(Jtr, htr, Ytr, epochs=epochs, batch=batch, lr=lr, seed=1, decay=0.97)
    t_train = time.perf_counter() - t0
    gtr = model.predict(Jtr, htr)
    gva = model.predict(Jva, hva)
    beta = S.fit_beta(Ytr, gtr)
    vr = S.variance_reduction(Yva, gva, beta)
    corr = np.array([np.corrcoef(Yva[:, p], gva[:, p])[0, 1] if Yva[:, p].std() > 0 else 0.0
                     for p in range(points)])
    mse = float(np.mean((gva - Yva) ** 2))
    worst = int(np.argmax(Yva.var(axis=0)))
    return {
        "workload": f"L={L} CPMG (pi/2_x, {points} x [10 x 0.05, pi_y, 10 x 0.05, measure]), corr_x/L, "
                    f"{n_train} training + {n_val} held-out exact realisations (seed {seed})",
        "exact_s": round(t_tr + t_va, 3), "exact_realisations_per_s": round((n_train + n_val) / (t_tr + t_va), 1),
        "train_s": round(t_train, 3), "epochs": epochs, "loss_first": round(hist[0], 4), "loss_last": round(hist[-1], 4),
        "held_out_mse": mse, "held_out_mean_sd": float(Yva.std(axis=0).mean()),
        "held_out_corr": [round(float(c), 4) for c in corr],
        "held_out_variance_reduction": [round(float(v), 3) for v in vr],
        "worst_point": worst, "worst_point_variance_reduction": round(float(vr[worst]), 3),
        "curve_mean": [round(float(v), 5) for v in Yva.mean(axis=0)],
        "curve_sd": [round(float(v), 5) for v in Yva.std(axis=0)],
    }


# ------------------------------------------------------------ control variate

@dataclass
class CVResult:
    plain: S.Estimate
    cv: S.Estimate
    line: dict


def control_variate(ctx: spx.Context, L: int = 16, target: float = 1e-3, draws: int = 1_000_000,
                    L_train: int = 10, n_train: int = 65536, n_val: int = 256, epochs: int = 60,
                    decay: float | None = None, seed: int = SEED, log=lambda m: None) -> CVResult:
    """Seconds to a +-target error bar on every point of the mean XY-16 echo
    at L: plain batched Monte-Carlo against the control variate (the
    surrogate on `draws` random geometries plus the exact correction), both
    arms actually run to the bar.

    The surrogate is trained on exact curves of the same sequence at the
    smaller L_train (the LDS-resident runner: about 2.4 s for 65536 curves
    at L = 10); the network is size-agnostic and the short-time echo is set
    by local geometry. The learning rate falls by `decay` after every epoch,
    by default to an eighth of its start over the epochs. The training set
    trades against the exact runs, which scale as the inverse of the worst
    point's variance reduction. On 1024 held-out L = 16 curves (one
    training run each): L = 12, 32768 curves x 100 epochs 29.7 (curves
    3.1 s, training 1.95 s); L = 11, 65536 x 60 33.8; L = 10, 65536 x 60
    35.2 (2.4 s, 2.3 s), x 80 36.2, 98304 x 40 34.9, 131072 x 30 36.1;
    L = 9, 131072 x 30 32.7. With fewer curves the smaller sizes fall
    behind: L = 10, 32768 x 100 reached 17-19 on 512 held-out curves,
    where L = 11 reached 30 and L = 12 31-37. beta is fitted on n_val held-out exact curves at L (its error
    adds a fraction of about 1 / n_val to the residual variance), whose
    variances also give both arms their first sample sizes; each arm then
    runs further realisations until its own standard error is at most
    `target` on every point. Whatever the surrogate's quality the estimate is
    unbiased, and plain Monte-Carlo runs beside it as the judge."""
    P = xy16_program()
    # disjoint index ranges: training, validation, plain arm, exact side of
    # the control variate, surrogate-only geometries
    base_val, base_plain, base_cv, base_draws = 4_000_000, 5_000_000, 6_000_000, 10_000_000
    log(f"training set: {n_train} exact L={L_train} curves")
    Ytr, t_ytr = exact_curves(ctx, P, L_train, 0, n_train, seed)
    Jtr, htr = graphs(L_train, 0, n_train, seed)
    model = S.Model(ctx, Jtr, htr, Ytr, seed=11)
    t0 = time.perf_counter()
    model.train(Jtr, htr, Ytr, epochs=epochs, batch=128, lr=2e-3, seed=1,
                decay=decay if decay is not None else 0.125 ** (1.0 / epochs))
    t_train = time.perf_counter() - t0
    log(f"validation: {n_val} exact L={L} curves")
    Yva, t_va = exact_curves(ctx, P, L, base_val, n_val, seed)
    Jva, hva = graphs(L, base_val, n_val, seed)
    gva = model.predict(Jva, hva)
    beta = S.fit_beta(Yva, gva)
    var_y = Yva.var(axis=0, ddof=1)
    var_r = (Yva - beta * gva).var(axis=0, ddof=1)
    se2 = target * target

    def more(var, share, n):
        """Runs still needed by a per-point variance at a per-point share of se^2."""
        return max(64, int(math.ceil(float((var / share).max()) * 1.01)) - n)

    log(f"plain: from {int(math.ceil(var_y.max() / se2))} exact L={L} runs")
    parts, t_plain, n_plain = [], 0.0, 0
    todo = int(math.ceil(var_y.max() / se2))
    for _ in range(4):
        y, dt = exact_curves_chunked(ctx, P, L, base_plain + n_plain, todo, seed, chunk=4096, log=log)
        parts.append(y)
        t_plain += dt
        n_plain += todo
        Yp = np.concatenate(parts)
        plain = S.plain_estimate(Yp)
        if plain.stderr.max() <= target:
            break
        todo = more(Yp.var(axis=0, ddof=1), se2, n_plain)
        log(f"  plain stderr {plain.stderr.max():.6f}: {todo} more")

    ctx.synchronize()
    t0 = time.perf_counter()
    gx = model.predict_dipolar(seed, base_draws, draws, L, **DRAW)
    t_sur = time.perf_counter() - t0
    var_g = beta * beta * gx.var(axis=0, ddof=1)
    if (var_g / draws).max() > 0.1 * se2:
        raise RuntimeError("the surrogate-only draws cannot carry their share of the error bar")
    share = se2 - var_g / draws
    log(f"control variate: from {int(math.ceil((var_r / share).max()))} exact runs + {draws} surrogate draws")
    ys, gs, t_exact, n_cv = [], [], 0.0, 0
    todo = int(math.ceil(float((var_r / share).max())))
    for _ in range(4):
        y, dt = exact_curves(ctx, P, L, base_cv + n_cv, todo, seed)
        t_exact += dt
        ctx.synchronize()
        t0 = time.perf_counter()
        J, h = graphs(L, base_cv + n_cv, todo, seed)
        gs.append(model.predict(J, h))
        t_sur += time.perf_counter() - t0
        ys.append(y)
        n_cv += todo
        Yc, gc = np.concatenate(ys), np.concatenate(gs)
        cv = S.cv_estimate(Yc, gc, gx, beta)
        if cv.stderr.max() <= target:
            break
        todo = more((Yc - beta * gc).var(axis=0, ddof=1), share, n_cv)
        log(f"  control variate stderr {cv.stderr.max():.6f}: {todo} more")
    t_cv = t_exact + t_sur
    t_oneshot = t_cv + t_ytr + t_train + t_va
    z = float(np.max(np.abs(cv.mean - plain.mean) / np.sqrt(cv.stderr ** 2 + plain.stderr ** 2)))
    line = {
        "workload": f"L={L} dipolar XXZ (min spacing 0.5, Hz sigma 0.5), XY-16 with finite pulses (Omega={OMEGA:g}, "
                    f"pi/Omega long, one 0.05 step of free evolution before each), corr_x/L after every 4 pulses; "
                    f"surrogate on {draws} random geometries + exact correction; +-{target:g} on every point",
        "plain": {"s": round(t_plain, 3), "exact_runs": n_plain, "max_stderr": float(plain.stderr.max())},
        "control_variate": {"s": round(t_cv, 3), "exact_runs": n_cv, "surrogate_draws": draws,
                            "exact_s": round(t_exact, 3), "surrogate_s": round(t_sur, 3),
                            "max_stderr": float(cv.stderr.max())},
        "speedup": round(t_plain / t_cv, 2),
        "one_shot": {"s": round(t_oneshot, 3), "speedup": round(t_plain / t_oneshot, 2),
                     "note": "training curves, training and the held-out beta fit counted once, as for a single estimate"},
        "worst_point_variance_reduction": round(float(var_y.max() / var_r.max()), 2),
        "held_out_corr": [round(float(np.corrcoef(Yva[:, p], gva[:, p])[0, 1]), 4) for p in range(Yva.shape[1])],
        "max_abs_z_cv_vs_plain": round(z, 2),
        "mean_plain": [round(float(v), 5) for v in plain.mean],
        "mean_cv": [round(float(v), 5) for v in cv.mean],
        "training": {"nspins": L_train, "exact_runs": n_train, "exact_s": round(t_ytr, 3), "train_s": round(t_train, 3),
                     "validation_runs": n_val, "validation_s": round(t_va, 3)},
    }
    return CVResult(plain, cv, line)


def main(argv):
    which = argv[1] if len(argv) > 1 else ""
    ctx = spx.Context(0)
    log = lambda m: print(f"[ensembles] {m}", file=sys.stderr, flush=True)  # noqa: E731
    if which == "training-set":
        print(json.dumps(training_set(ctx, log=log)), flush=True)
    elif which == "control-variate":
        print(json.dumps(control_variate(ctx, log=log).line), flush=True)
    elif which == "xy16-curves":
        # exact XY-16 curves alone (the control variate's exact runs):
        # L and count from the command line, after a warm-up of 16
        L = int(argv[2]) if len(argv) > 2 else 16
        count = int(argv[3]) if len(argv) > 3 else 256
        P = xy16_program()
        exact_curves(ctx, P, L, 0, 16)
        _, dt = exact_curves(ctx, P, L, 1000, count)
        print(json.dumps({"workload": f"{count} exact XY-16 curves at L={L}", "s": round(dt, 3),
                          "realisations_per_s": round(count / dt, 1)}), flush=True)
    else:
        raise SystemExit("usage: python -m spinsim_amd.ensembles training-set | control-variate | "
                         "xy16-curves [L] [count]")


if __name__ == "__main__":
    main(sys.argv)
