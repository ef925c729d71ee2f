"""FP64 numpy restatement of the surrogate graph network of include/spx_gnn.h:
forward, loss, hand-derived backward and one Adam step.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
network, never by the product path.

The network has no counterpart in the reference, so nothing there pins it.
What pins this restatement (tests/test_gnn_oracle.py):
  * the backward against central finite differences of the forward in FP64,
    per parameter block and along random directions (a dropped term, a
    transposed operand or a wrong ReLU mask in the backward fails it);
  * the forward's invariance under relabelling the spins of a graph, and its
    reduction at J = 0 to a per-spin perceptron evaluated by plain loops
    (a transposed J or a wrong pooling axis fails one of them), and the
    edge weights and features against a plain-loop evaluation;
  * Adam against its textbook first step (|dp| = lr for every coordinate
    with a nonzero gradient).
With `bf16=True` every matrix-product operand is rounded to BF16 (nearest
even) at the points where the device rounds it, so the device can be held
to a tight tolerance; `bf16=False` is the exact FP64 network.
"""
from __future__ import annotations

import numpy as np

H = 128
F = 16
NLAYER = 3


def layout(P: int):
    """(name, shape) of each parameter block, in the device's order."""
    blocks = [("W0", (H, F)), ("b0", (H,))]
    for l in range(1, NLAYER + 1):
        blocks += [(f"W{l}", (H, 2 * H)), (f"b{l}", (H,))]
    blocks += [("Wout", (P, H)), ("bout", (P,))]
    return blocks


def num_params(P: int) -> int:
    return sum(int(np.prod(s)) for _, s in layout(P))


def unpack(p: np.ndarray, P: int) -> dict:
    out, o = {}, 0
    for name, shape in layout(P):
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d: dict, P: int) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in layout(P)])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16, nearest even, through FP32 (as the device does)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def graph_inputs(J: np.ndarray, h: np.ndarray, sh: float, sj: float):
    """Edge weights A [G, L, L] and node features [G, L, F]."""
    G, L = h.shape
    x = J / sj
    d = np.abs(x).sum(axis=2)
    A = x / np.sqrt((1.0 + d)[:, :, None] * (1.0 + d)[:, None, :])
    f = np.zeros((G, L, F))
    f[:, :, 0] = h / sh
    f[:, :, 1] = d
    f[:, :, 2] = np.log1p((x * x).sum(axis=2))
    f[:, :, 3] = 1.0
    return A, f


def forward(p, J, h, P, sh=1.0, sj=1.0, use_bf16=False):
    """Predictions [G, P] and the cache the backward needs."""
    r = bf16 if use_bf16 else (lambda x: x)
    q = _f32 if use_bf16 else (lambda x: x)
    w = unpack(p, P)
    A, f = graph_inputs(J, h, sh, sj)
    Aq = q(A)
    feat = r(q(f))
    cache = {"A": Aq, "feat": feat, "z": [], "cat": [None]}
    z = feat @ r(w["W0"]).T + w["b0"]
    cache["z"].append(z)
    x = np.maximum(z, 0.0)
    for l in range(1, NLAYER + 1):
        m = Aq @ x
        cat = r(np.concatenate([x, m], axis=2))
        z = cat @ r(w[f"W{l}"]).T + w[f"b{l}"]
        cache["cat"].append(cat)
        cache["z"].append(z)
        x = np.maximum(z, 0.0)
    pool = r(x.mean(axis=1))
    cache["pool"] = pool
    y = pool @ r(w["Wout"]).T + w["bout"]
    return y, cache


def loss_and_grad(p, J, h, target, P, sh=1.0, sj=1.0, use_bf16=False):
    """MSE over G x P and its gradient in the parameter order."""
    r = bf16 if use_bf16 else (lambda x: x)
    w = unpack(p, P)
    y, c = forward(p, J, h, P, sh, sj, use_bf16)
    G, L = h.shape
    d = y - target
    loss = float(np.mean(d * d))
    dy = 2.0 * d / d.size
    g = {}
    dyr = r(dy)
    g["Wout"] = dyr.T @ c["pool"]
    g["bout"] = dy.sum(axis=0)
    dpool = dyr @ r(w["Wout"])
    dx = np.repeat(dpool[:, None, :] / L, L, axis=1)
    for l in range(NLAYER, -1, -1):
        dz = dx * (c["z"][l] > 0)
        dzr = r(dz)
        g[f"b{l}"] = dz.sum(axis=(0, 1))
        if l == 0:
            g["W0"] = dzr.reshape(-1, H).T @ c["feat"].reshape(-1, F)
            break
        cat = c["cat"][l]
        g[f"W{l}"] = dzr.reshape(-1, H).T @ cat.reshape(-1, 2 * H)
        dcat = dzr @ r(w[f"W{l}"])
        dm = dcat[:, :, H:]
        dx = dcat[:, :, :H] + np.swapaxes(c["A"], 1, 2) @ dm
    return loss, pack(g, P)


def adam_step(p, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update; `step` counts from 1. Returns (p, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
