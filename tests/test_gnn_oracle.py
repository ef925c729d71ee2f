"""Pins the FP64 surrogate-network oracle on the CPU: finite differences,
permutation invariance, the J = 0 reduction and Adam's first step."""
import math

import numpy as np

from oracle import gnn_oracle as N


def sample(G=3, L=6, P=5, seed=1):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(G, L, L))
    J = np.triu(A, 1)
    J = J + np.swapaxes(J, 1, 2)
    h = rng.normal(size=(G, L))
    p = np.zeros(N.num_params(P))
    o = 0
    for name, shape in N.layout(P):
        n = int(np.prod(shape))
        scale = 0.0 if name.startswith("b") else np.sqrt(2.0 / shape[-1])
        blk = rng.normal(scale=scale, size=shape) if scale else rng.normal(scale=0.1, size=shape)
        if name == "W0":
            blk[:, 4:] = 0.0
        p[o:o + n] = blk.reshape(-1)
        o += n
    t = rng.normal(size=(G, P))
    return J, h, p, t


def test_backward_matches_central_differences_per_block():
    J, h, p, t = sample()
    P = t.shape[1]
    _, g = N.loss_and_grad(p, J, h, t, P, sh=1.3, sj=0.7)
    rng = np.random.default_rng(7)
    o = 0
    for name, shape in N.layout(P):
        n = int(np.prod(shape))
        idx = rng.choice(n, size=min(n, 6), replace=False) + o
        if name == "W0":  # live input columns only: features 0..3
            rows = rng.choice(shape[0], size=6)
            idx = o + rows * shape[1] + rng.integers(0, 4, size=6)
        for i in idx:
            e = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += e
            pm[i] -= e
            fd = (N.loss_and_grad(pp, J, h, t, P, 1.3, 0.7)[0] - N.loss_and_grad(pm, J, h, t, P, 1.3, 0.7)[0]) / (2 * e)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (name, i, fd, g[i])
        o += n


def test_backward_matches_directional_differences():
    J, h, p, t = sample(G=4, L=7, P=3, seed=3)
    P = 3
    _, g = N.loss_and_grad(p, J, h, t, P)
    rng = np.random.default_rng(11)
    for _ in range(4):
        d = rng.normal(size=p.size)
        d /= np.linalg.norm(d)
        e = 1e-6
        fd = (N.loss_and_grad(p + e * d, J, h, t, P)[0] - N.loss_and_grad(p - e * d, J, h, t, P)[0]) / (2 * e)
        assert abs(fd - g @ d) <= 1e-6 * abs(fd) + 1e-8


def test_forward_is_invariant_under_spin_relabelling():
    J, h, p, _ = sample(G=2, L=8, P=4, seed=5)
    y, _ = N.forward(p, J, h, 4)
    perm = np.random.default_rng(2).permutation(8)
    y2, _ = N.forward(p, J[:, perm][:, :, perm], h[:, perm], 4)
    assert np.allclose(y, y2, rtol=1e-12, atol=1e-12)
    # a non-symmetric relabelling of J alone changes the prediction
    y3, _ = N.forward(p, J[:, perm], h, 4)
    assert not np.allclose(y, y3)


def test_forward_without_couplings_is_a_per_spin_perceptron():
    _, h, p, _ = sample(G=2, L=5, P=3, seed=9)
    J = np.zeros((2, 5, 5))
    P = 3
    w = N.unpack(p, P)
    want = np.zeros((2, P))
    for g in range(2):
        acc = np.zeros(N.H)
        for i in range(5):
            f = np.zeros(N.F)
            f[0], f[1], f[2], f[3] = h[g, i] / 2.0, 0.0, 0.0, 1.0
            x = np.maximum(w["W0"] @ f + w["b0"], 0.0)
            for l in range(1, N.NLAYER + 1):
                x = np.maximum(w[f"W{l}"][:, :N.H] @ x + w[f"b{l}"], 0.0)
            acc += x
        want[g] = w["Wout"] @ (acc / 5) + w["bout"]
    y, _ = N.forward(p, J, h, P, sh=2.0)
    assert np.allclose(y, want, rtol=1e-12, atol=1e-12)


def test_graph_inputs_match_plain_loops():
    J, h, _, _ = sample(G=2, L=5, P=2, seed=4)
    A, f = N.graph_inputs(J, h, 0.5, 1.7)
    for g in range(2):
        w = [[J[g, i, j] / 1.7 for j in range(5)] for i in range(5)]
        d = [sum(abs(v) for v in row) for row in w]
        for i in range(5):
            assert math.isclose(f[g, i, 0], h[g, i] / 0.5, rel_tol=1e-15)
            assert math.isclose(f[g, i, 1], d[i], rel_tol=1e-14)
            s2 = sum((J[g, i, j] / 1.7) ** 2 for j in range(5))
            assert math.isclose(f[g, i, 2], math.log(1 + s2), rel_tol=1e-14)
            assert f[g, i, 3] == 1.0 and not f[g, i, 4:].any()
            for j in range(5):
                want = w[i][j] / math.sqrt((1 + d[i]) * (1 + d[j]))
                assert math.isclose(A[g, i, j], want, rel_tol=1e-14, abs_tol=1e-300)
    assert np.allclose(A, np.swapaxes(A, 1, 2), rtol=0, atol=0)


def test_adam_first_step_moves_each_live_coordinate_by_lr():
    rng = np.random.default_rng(0)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    g[::7] = 0.0
    p1, m, v = N.adam_step(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    step = np.abs(p1 - p)
    assert np.allclose(step[g != 0], 1e-3, rtol=1e-6)
    assert np.all(step[g == 0] == 0)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.14159, 0.0])
    r = N.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie, even mantissa stays
    assert r[2] == 1.0 + 4 * 2 ** -8  # tie, rounds up to even
    assert r[3] == 1.0
    assert abs(r[4] + 3.140625) < 1e-12
    assert r[5] == 0.0
