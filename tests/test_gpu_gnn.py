"""The surrogate graph network on the device (BF16 MFMA, FP32 accumulate)
against the FP64 numpy oracle, and its use as a control variate for the
disorder average of an echo curve against plain Monte-Carlo."""
import math
import time

import numpy as np
import pytest

from oracle import gnn_oracle as N
from oracle import oracle as O
from spinsim_amd import spx
from spinsim_amd import surrogate as S

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def random_graphs(G, L, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(G, L, L)) / L
    J = np.triu(A, 1)
    J = J + np.swapaxes(J, 1, 2)
    h = rng.normal(scale=0.5, size=(G, L))
    return J, h


def block_errors(got, want, P):
    out = {}
    o = 0
    for name, shape in N.layout(P):
        n = int(np.prod(shape))
        a, b = got[o:o + n], want[o:o + n]
        out[name] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
        o += n
    return out


def test_gnn_forward_and_gradient_match_oracle(ctx):
    """Predictions and every parameter block's gradient: against the oracle
    with BF16 operands rounded where the device rounds them (FP32 vs FP64
    accumulation only), and against the exact FP64 network."""
    banner("gnn forward/gradient vs oracle G=8 L=10 P=16")
    P, G, L = 16, 8, 10
    sur = spx.Surrogate(ctx, P, seed=3, scale_h=0.5, scale_j=1.0)
    p = sur.get_params().astype(np.float64)
    assert p.size == N.num_params(P)
    J, h = random_graphs(G, L, 1)
    t = np.random.default_rng(2).normal(scale=0.3, size=(G, P))
    y = sur.predict(J, h).astype(np.float64)
    yb, _ = N.forward(p, J, h, P, 0.5, 1.0, use_bf16=True)
    ye, _ = N.forward(p, J, h, P, 0.5, 1.0, use_bf16=False)
    eb = np.linalg.norm(y - yb) / np.linalg.norm(yb)
    ee = np.linalg.norm(y - ye) / np.linalg.norm(ye)
    banner(f"forward rel err: vs bf16 oracle {eb:.2e}, vs fp64 {ee:.2e}")
    assert eb < 2e-3 and ee < 3e-2
    loss, g = sur.gradient(J, h, t)
    lb, gb = N.loss_and_grad(p, J, h, t, P, 0.5, 1.0, use_bf16=True)
    le, ge = N.loss_and_grad(p, J, h, t, P, 0.5, 1.0, use_bf16=False)
    assert abs(loss - lb) <= 2e-3 * lb and abs(loss - le) <= 3e-2 * le
    errb = block_errors(g.astype(np.float64), gb, P)
    erre = block_errors(g.astype(np.float64), ge, P)
    banner("grad rel err vs bf16 oracle: " + " ".join(f"{k}={v:.1e}" for k, v in errb.items()))
    banner("grad rel err vs fp64 oracle: " + " ".join(f"{k}={v:.1e}" for k, v in erre.items()))
    for k in errb:
        assert errb[k] < 1e-2, (k, errb[k])
        assert erre[k] < 5e-2, (k, erre[k])
    cos = float(g @ ge / (np.linalg.norm(g) * np.linalg.norm(ge)))
    assert cos > 0.999
    # deterministic: a second call is bitwise the same
    loss2, g2 = sur.gradient(J, h, t)
    assert loss2 == loss and np.array_equal(g, g2)


def test_gnn_adam_step_matches_oracle(ctx):
    banner("gnn adam step")
    P, G, L = 8, 4, 7
    sur = spx.Surrogate(ctx, P, seed=5)
    J, h = random_graphs(G, L, 4)
    t = np.random.default_rng(5).normal(size=(G, P))
    p0 = sur.get_params().astype(np.float64)
    _, g = sur.gradient(J, h, t)
    sur.train_step(J, h, t, 1e-3)
    p1 = sur.get_params().astype(np.float64)
    want, _, _ = N.adam_step(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    assert np.max(np.abs(p1 - want)) < 1e-6
    # set/get round trip
    sur.set_params(p0.astype(np.float32))
    assert np.array_equal(sur.get_params(), p0.astype(np.float32))


def test_resident_epoch_equals_step_by_step_training(ctx):
    """An epoch over the device-resident training set applies exactly the
    updates of train_step on the same minibatches (bitwise parameters)."""
    banner("gnn resident epoch vs train_step")
    P, L, n, batch = 8, 9, 64, 16
    J, h = random_graphs(n, L, 8)
    t = np.random.default_rng(9).normal(size=(n, P)).astype(np.float32)
    a = spx.Surrogate(ctx, P, seed=2)
    b = spx.Surrogate(ctx, P, seed=2)
    a.set_training_set(J, h, t)
    rng = np.random.default_rng(1)
    for epoch in range(2):
        order = rng.permutation(n).astype(np.uint32)
        loss_a = a.train_epoch(order, batch, 1e-3)
        losses = [b.train_step(J[order[s:s + batch]], h[order[s:s + batch]], t[order[s:s + batch]], 1e-3)
                  for s in range(0, n, batch)]
        assert np.array_equal(a.get_params(), b.get_params()), epoch
        assert abs(loss_a - np.mean(losses)) <= 1e-5 * max(1.0, abs(loss_a))
    with pytest.raises(spx.OutOfRange):
        a.train_epoch(np.full(n, n, np.uint32), batch, 1e-3)
    with pytest.raises(spx.InvalidArgument):
        a.train_epoch(np.arange(n, dtype=np.uint32), n + 1, 1e-3)


@pytest.mark.parametrize("G,L,P", [(3, 1, 5), (9, 30, 3), (700, 2, 40), (2100, 12, 4)])
def test_gnn_gradient_block_shapes(ctx, G, L, P):
    """The training kernel's edge shapes against the BF16-rounded oracle:
    one spin per graph (16 graphs a block), 30 spins (2 a block), outputs
    not a multiple of 8 and more than one 16-wide tile, and a batch of 420
    blocks (four launches adding into the same rows of partial sums).
    Bitwise repeatable."""
    banner(f"gnn gradient G={G} L={L} P={P}")
    sur = spx.Surrogate(ctx, P, seed=G + L, scale_h=0.5, scale_j=1.0)
    p = sur.get_params().astype(np.float64)
    J, h = random_graphs(G, L, G)
    t = np.random.default_rng(L).normal(scale=0.3, size=(G, P))
    loss, g = sur.gradient(J, h, t)
    lb, gb = N.loss_and_grad(p, J, h, t, P, 0.5, 1.0, use_bf16=True)
    assert abs(loss - lb) <= 2e-3 * lb
    errb = block_errors(g.astype(np.float64), gb, P)
    banner("grad rel err vs bf16 oracle: " + " ".join(f"{k}={v:.1e}" for k, v in errb.items()))
    for k, e in errb.items():
        assert e < 1e-2, (k, e)
    loss2, g2 = sur.gradient(J, h, t)
    assert loss2 == loss and np.array_equal(g, g2)


def test_gnn_argument_errors(ctx):
    banner("gnn argument errors")
    with pytest.raises(spx.InvalidArgument):
        spx.Surrogate(ctx, 0)
    with pytest.raises(spx.InvalidArgument):
        spx.Surrogate(ctx, 129)
    sur = spx.Surrogate(ctx, 4)
    J, h = random_graphs(2, 5, 0)
    J[1, 0, 1] += 1.0  # not symmetric
    with pytest.raises(spx.InvalidArgument):
        sur.predict(J, h)
    with pytest.raises(ValueError):
        sur.predict(J[:, :, :4], h)
    with pytest.raises(ValueError):
        sur.set_params(np.zeros(3, np.float32))


# ------------------------------------------------------- control variate

SEED = 20240726


def cpmg_y(P, L, cycles, steps):
    """pi/2_x on every spin, then `cycles` x [tau, pi_y, tau, measure]."""
    allspins = list(range(L))
    P.pulse((1, 0, 0), math.pi / 2, allspins)
    P.repeat(cycles)
    P.evolve(0.05, steps)
    P.pulse((0, 1, 0), math.pi, allspins)
    P.evolve(0.05, steps)
    P.measure()
    P.end()
    return P


def exact_curves(ctx, L, first, count, cycles, steps):
    """Mean transverse polarisation <sigma_y> over spins after each cycle,
    from the all-down state, for realisations first .. first+count-1 of the
    dipolar ensemble (min spacing 0.5, secular angle, Sz sigma 0.5), by the
    resident batch runner; also the run's wall time."""
    obs = [(O.OBS_MAG_Y, s) for s in range(L)]
    b = spx.Batch(ctx, cpmg_y(spx.Program(), L, cycles, steps), obs, epsilon=1e-7, reference_axis=0)
    psi0 = np.zeros(2 << L)
    psi0[0] = 1.0
    f = np.zeros((L, 3))
    for i in range(first, first + count):
        d = spx.dipolar_realisation(SEED, i, L)
        f[:, 2] = d.hz
        b.add(L, f, d.lo, d.hi, d.jx, d.jy, d.jz, psi0)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, v = b.run()
    return v.mean(axis=2), time.perf_counter() - t0


def test_control_variate_beats_plain_monte_carlo(ctx):
    """Trains the surrogate on exact CPMG curves (L=10, 32 echoes) from the
    device's own kernels. On held-out realisations: (1) it removes most of
    the disorder variance at the curve's worst point (the one that sets the
    number of exact runs for a common error bar); (2) the control-variate
    mean from 1024 exact + 65536 surrogate-only realisations agrees with the
    plain mean of 3072 exact ones within their error bars; (3) at a +-1e-3
    error bar on every point it costs several times less device time than
    plain Monte-Carlo, counting the surrogate's own draws."""
    L, cycles, steps = 10, 32, 5
    n_train, n_val, n_est, M = 12288, 2048, 1024, 65536
    banner(f"control variate: exact curves L={L} for {n_train + n_val + n_est} realisations")
    Ytr, t_tr = exact_curves(ctx, L, 0, n_train, cycles, steps)
    Yva, _ = exact_curves(ctx, L, n_train, n_val, cycles, steps)
    Yes, _ = exact_curves(ctx, L, n_train + n_val, n_est, cycles, steps)
    t_exact = t_tr / n_train
    banner(f"exact: {t_exact * 1e6:.1f} us per realisation (batched)")
    Jtr, htr = spx.dipolar_graphs(SEED, 0, n_train, L)
    Jva, hva = spx.dipolar_graphs(SEED, n_train, n_val, L)
    Jes, hes = spx.dipolar_graphs(SEED, n_train + n_val, n_est, L)
    banner(f"curves: mean {np.round(Ytr.mean(axis=0)[[0, 3, 15, 31]], 4)} sd {np.round(Ytr.std(axis=0)[[0, 3, 15, 31]], 4)}")
    sur = S.Model(ctx, Jtr, htr, Ytr, seed=11)
    t0 = time.perf_counter()
    hist = sur.train(Jtr, htr, Ytr, epochs=60, batch=128, lr=2e-3, seed=1, decay=0.97)
    t_train = time.perf_counter() - t0
    banner(f"training {t_train:.1f} s: loss first {hist[0]:.3f} last {hist[-1]:.3f}")
    assert hist[-1] < 0.6 * hist[0]
    gva = sur.predict(Jva, hva)
    beta = S.fit_beta(Yva, gva)
    vrf = S.variance_reduction(Yva, gva, beta)
    var_y = Yva.var(axis=0, ddof=1)
    var_r = (Yva - beta * gva).var(axis=0, ddof=1)
    gain = var_y.max() / var_r.max()
    banner(f"held-out VRF first points {np.round(vrf[:6], 2)}; worst-point gain {gain:.2f}")
    assert gain > 4.0
    # surrogate-only geometries drawn and evaluated inside libspx, timed
    ctx.synchronize()
    t0 = time.perf_counter()
    gex = sur.predict_dipolar(SEED, 10_000_000, M, L)
    t_sur = (time.perf_counter() - t0) / M
    # the bulk path sees the same graphs as the explicit one
    Jx, hx = spx.dipolar_graphs(SEED, 10_000_000, 64, L)
    assert np.allclose(sur.predict(Jx, hx), gex[:64], rtol=0, atol=1e-6)
    ges = sur.predict(Jes, hes)
    cv = S.cv_estimate(Yes, ges, gex, beta)
    plain = S.plain_estimate(np.concatenate([Yva, Yes]))
    z = np.abs(cv.mean - plain.mean) / np.sqrt(cv.stderr ** 2 + plain.stderr ** 2)
    banner(f"cv vs plain: max |z| {z.max():.2f}; worst se cv(1024+{M}) {cv.stderr.max():.2e} "
           f"plain(3072) {plain.stderr.max():.2e}; surrogate {t_sur * 1e6:.2f} us per geometry")
    assert z.max() < 4.5
    assert cv.stderr.max() < plain.stderr.max()
    # Device time to a +-1e-3 error bar on every point: plain needs
    # max Var(y) / se^2 exact runs; the control variate max Var(y - beta g)
    # / (0.9 se^2) exact runs plus enough surrogate draws for the other 10%.
    se2 = 1e-6
    cost_plain = var_y.max() / se2 * t_exact
    n_c = var_r.max() / (0.9 * se2)
    m_c = (beta * beta * gva.var(axis=0, ddof=1)).max() / (0.1 * se2)
    cost_cv = n_c * t_exact + m_c * t_sur
    banner(f"to +-1e-3: plain {cost_plain:.2f} s ({var_y.max() / se2:.0f} runs), control variate "
           f"{cost_cv:.2f} s ({n_c:.0f} runs + {m_c:.0f} draws) = x{cost_plain / cost_cv:.1f}; "
           f"one-off training {t_train:.1f} s on {n_train} runs")
    assert cost_plain / cost_cv > 3.0
