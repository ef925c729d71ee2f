"""The algebra behind the Hadamard path (spinsim_amd/csrc/spx_hadamard.hip),
checked on the CPU with small numpy restatements: the flip-flop pair sum
the reference applies term by term (axpyFlipFlop,
src/core/SpinOperators.cpp:254-276) equals W E W + P W E W P, E = D / 2^(L+1); the column pass's
factored E equals the direct sum; and the row/column split of W is W."""
import itertools

import numpy as np
import pytest


def popcount(k):
    return bin(int(k)).count("1")


def flip_flop_direct(L, pairs, w, psi):
    """sum_q w_q (s+_i s-_j + s-_i s+_j) psi: partner k ^ m where bits i, j differ."""
    k = np.arange(1 << L)
    out = np.zeros_like(psi)
    for (i, j), c in zip(pairs, w):
        m = (1 << i) | (1 << j)
        differ = ((k >> i) & 1) != ((k >> j) & 1)
        out += np.where(differ, c * psi[k ^ m], 0.0)
    return out


def wht(x, bits):
    """Unnormalised Walsh-Hadamard butterflies over the given index bits."""
    x = x.copy()
    for b in bits:
        y = x.reshape(-1, 2, 1 << b)
        a, c = y[:, 0, :].copy(), y[:, 1, :].copy()
        y[:, 0, :], y[:, 1, :] = a + c, a - c
        x = y.reshape(-1)
    return x


def energies(L, pairs, w):
    """E(s) = sum_q w_q z_i(s) z_j(s) / 2^(L+1), z = +1 for bit 0."""
    k = np.arange(1 << L)
    e = np.zeros(1 << L)
    for (i, j), c in zip(pairs, w):
        e += c * np.where(((k >> i) & 1) == ((k >> j) & 1), 1.0, -1.0)
    return np.ldexp(e, -(L + 1))


@pytest.mark.parametrize("L", [2, 3, 6, 9])
def test_flip_flop_is_a_hadamard_sandwich(L):
    rng = np.random.default_rng(L)
    pairs = list(itertools.combinations(range(L), 2))
    w = rng.standard_normal(len(pairs))
    psi = rng.standard_normal(1 << L) + 1j * rng.standard_normal(1 << L)
    want = flip_flop_direct(L, pairs, w, psi)
    P = np.array([-1.0 if popcount(k) & 2 else 1.0 for k in range(1 << L)])
    E = energies(L, pairs, w)
    allbits = range(L)
    got = wht(E * wht(psi, allbits), allbits) + P * wht(E * wht(P * psi, allbits), allbits)
    assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
    # dropping the P term leaves only (X X) / 2: the check above is sensitive to it
    half = 2 * wht(E * wht(psi, allbits), allbits)
    assert np.max(np.abs(half - want)) > 1e-3


def test_single_pair_matches_the_textbook_matrix():
    """L = 2, one pair of weight c: the 4 x 4 operator is c (|01><10| + |10><01|)."""
    L, c = 2, 0.7
    M = np.zeros((4, 4))
    for col in range(4):
        e = np.zeros(4)
        e[col] = 1.0
        M[:, col] = flip_flop_direct(L, [(0, 1)], [c], e)
    want = np.zeros((4, 4))
    want[1, 2] = want[2, 1] = c
    assert np.array_equal(M, want)
    E = energies(L, [(0, 1)], [c])
    P = np.array([1.0, 1.0, 1.0, -1.0])
    for col in range(4):
        e = np.zeros(4)
        e[col] = 1.0
        got = wht(E * wht(e, [0, 1]), [0, 1]) + P * wht(E * wht(P * e, [0, 1]), [0, 1])
        assert np.allclose(got, want[:, col], atol=1e-15)


@pytest.mark.parametrize("L,TB", [(8, 5), (10, 6), (9, 3)])
def test_factored_diagonal_and_row_column_split(L, TB):
    """E(x, c) = E_lo(c) + E_hi(x) + sum_j z_j(x) A_j(c) for k = x 2^TB + c
    (the tables the column pass reads), and W = W_high W_low."""
    rng = np.random.default_rng(100 + L)
    pairs = list(itertools.combinations(range(L), 2))
    w = rng.standard_normal(len(pairs))
    weight = dict(zip(pairs, w))
    G = L - TB
    E = energies(L, pairs, w)
    z = lambda k, b: -1.0 if (k >> b) & 1 else 1.0
    elo = [sum(weight[(i, j)] * z(c, i) * z(c, j) for i, j in itertools.combinations(range(TB), 2))
           for c in range(1 << TB)]
    ehi = [sum(weight[(TB + i, TB + j)] * z(x, i) * z(x, j) for i, j in itertools.combinations(range(G), 2))
           for x in range(1 << G)]
    A = [[sum(weight[(i, TB + j)] * z(c, i) for i in range(TB)) for c in range(1 << TB)] for j in range(G)]
    for k in range(1 << L):
        x, c = k >> TB, k & ((1 << TB) - 1)
        f = elo[c] + ehi[x] + sum(z(x, j) * A[j][c] for j in range(G))
        assert abs(np.ldexp(f, -(L + 1)) - E[k]) < 1e-14
    v = rng.standard_normal(1 << L)
    assert np.allclose(wht(wht(v, range(TB)), range(TB, L)), wht(v, range(L)), atol=1e-12)
