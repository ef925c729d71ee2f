"""Independent dense reference for small L, after tests/DenseReference.cpp:
Pauli operators built by explicit Kronecker products (never by bit
arithmetic), the dense Hamiltonian, and exp(c H) psi through numpy's
Hermitian eigendecomposition (LAPACK, as the reference's zheev_)."""
from __future__ import annotations

import numpy as np

_P = {
    # basis (index 0 = down, index 1 = up), as tests/DenseReference.cpp:43-66
    "x": np.array([[0, 1], [1, 0]], dtype=complex),
    "y": np.array([[0, 1j], [-1j, 0]], dtype=complex),
    "z": np.array([[-1, 0], [0, 1]], dtype=complex),
}


def pauli(L: int, spin: int, axis: str) -> np.ndarray:
    """Spin s at Kronecker position L-1-s from the most significant factor."""
    out = np.eye(1, dtype=complex)
    for pos in range(L - 1, -1, -1):
        out = np.kron(out, _P[axis] if pos == spin else np.eye(2, dtype=complex))
    return out


def hamiltonian(L: int, fields, lo=(), hi=(), jx=(), jy=(), jz=()) -> np.ndarray:
    f = np.asarray(fields, dtype=float).reshape(L, 3)
    H = np.zeros((1 << L, 1 << L), dtype=complex)
    for s in range(L):
        for a, ax in enumerate("xyz"):
            if f[s, a] != 0.0:
                H += 0.5 * f[s, a] * pauli(L, s, ax)
    for p in range(len(lo)):
        for j, ax in ((jx[p], "x"), (jy[p], "y"), (jz[p], "z")):
            if j != 0.0:
                H += 0.25 * j * (pauli(L, int(lo[p]), ax) @ pauli(L, int(hi[p]), ax))
    return H


def to_complex(psi: np.ndarray) -> np.ndarray:
    return psi[0::2] + 1j * psi[1::2]


def to_interleaved(c: np.ndarray) -> np.ndarray:
    out = np.empty(2 * c.size)
    out[0::2] = c.real
    out[1::2] = c.imag
    return out


def expm_apply(H: np.ndarray, factor: complex, psi: np.ndarray) -> np.ndarray:
    w, V = np.linalg.eigh(H)
    c = V.conj().T @ to_complex(psi)
    return to_interleaved(V @ (np.exp(factor * w) * c))


def max_abs_diff(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.max(np.abs(to_complex(a) - to_complex(b))))


def sample_hamiltonian(L: int):
    """tests/test_Propagator.cpp:19-32: every field component populated,
    anisotropic all-to-all couplings."""
    fields = np.zeros((L, 3))
    for s in range(L):
        k = s + 1.0
        fields[s] = (0.4 * k, -0.3 / k, 0.9)
    lo, hi, jx, jy, jz = [], [], [], [], []
    for a in range(L):
        for b in range(a + 1, L):
            r = float(b - a)
            lo.append(a); hi.append(b)
            jx.append(0.5 / r); jy.append(-0.2 / r); jz.append(0.8 / (r * r))
    return fields, np.array(lo, np.uint32), np.array(hi, np.uint32), np.array(jx), np.array(jy), np.array(jz)
