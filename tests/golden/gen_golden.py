"""Writes tests/golden/ref_vectors.npz from the reference's own compiled code.

The reference's RNG, Bessel, StateVector, SpinOperators and Hamiltonian
sources compile here unmodified; oracle/Makefile (target `ref`) builds them
with a C-ABI shim into oracle/_ref/libspinsim_ref.so. This script records
their outputs on fixed inputs, so that the oracle stays pinned to the
reference where the reference sources are not present.

    make -C oracle ref && python tests/golden/gen_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def sample_system(L: int, index: int, anisotropic: bool, transverse: bool):
    """Dipolar geometry + disorder drawn by the oracle (its RNG is pinned
    separately), then optionally made anisotropic / given transverse fields so
    every kernel class is live."""
    c, hz, _ = O.dipolar_lattice(L, 20240726, index, disorder=1.5)
    fields = np.zeros((L, 3))
    fields[:, 2] = hz
    if transverse:
        fields[:, 0] = 0.3 + 0.1 * np.arange(L)
        fields[:, 1] = -0.2
    jx = c.jx * (1.3 if anisotropic else 1.0)
    return fields, c.lo.copy(), c.hi.copy(), jx, c.jy.copy(), c.jz.copy()


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref not built (make -C oracle ref)")
    out = {}
    # RNG: uniform and gaussian streams of forRealization(seed, index)
    for seed, index in ((20240726, 0), (20240726, 1), (20240726, 179), (12345, 7), (31337, 0)):
        for g in (0, 1):
            v = np.zeros(64)
            R.ref_rng_draws(seed, index, 64, g, O._d(v))
            out[f"rng_{seed}_{index}_{'g' if g else 'u'}"] = v
    v = np.zeros(64)
    R.ref_rng_seeded(2024, 64, 1, O._d(v))
    out["rng_seeded_2024_g"] = v
    # Bessel sequences at arguments beyond the Fortran fixture's grid
    for x in (0.3, 1.7, 3.05, 12.5, 88.0):
        for k, kind in ((0, "J"), (1, "I")):
            v = np.zeros(160)
            nz = O.C.c_size_t()
            R.ref_bessel(k, x, 160, O._d(v), O.C.byref(nz))
            out[f"bessel_{kind}_{x}"] = v
    # Hamiltonian applications (applyPhysical / applyTwiceNormalized /
    # accumulateTwiceNormalized) on fixed inputs
    for L, aniso, trans in ((4, True, True), (6, False, True), (7, True, False), (9, False, False)):
        fields, lo, hi, jx, jy, jz = sample_system(L, L, aniso, trans)
        psi = O.random_state(L, 100 + L)
        acc0 = O.random_state(L, 200 + L)
        bound = R.ref_ham_bound(L, O._d(np.ascontiguousarray(fields)), lo.size, O._u32(lo),
                                O._u32(hi), O._d(jx), O._d(jy), O._d(jz))
        key = f"ham_L{L}"
        out[key + "_fields"] = fields
        out[key + "_pairs"] = np.stack([lo.astype(float), hi.astype(float), jx, jy, jz])
        out[key + "_psi"] = psi
        out[key + "_acc0"] = acc0
        out[key + "_bound"] = np.array([bound])
        for mode in (0, 1, 2):
            o = acc0.copy() if mode == 2 else np.zeros_like(psi)
            R.ref_ham_apply(mode, L, O._d(np.ascontiguousarray(fields)), lo.size, O._u32(lo),
                            O._u32(hi), O._d(jx), O._d(jy), O._d(jz), bound, O._d(psi), O._d(o))
            out[f"{key}_out{mode}"] = o
        # a few Chebyshev steps through the reference's own state operations.
        # The expansion is built by the oracle (the reference's Chebyshev.cpp
        # needs <format>, absent here) and recorded as the steps' INPUT, not
        # as reference output; the step results are the reference's.
        for mode, tau in ((0, 0.05), (0, -0.3), (1, 0.1)):
            rc, coeffs, eff = O.cheb_build(mode, bound, tau, 1e-10)
            p = psi.copy()
            R.ref_one_step(mode, L, O._d(np.ascontiguousarray(fields)), lo.size, O._u32(lo),
                           O._u32(hi), O._d(jx), O._d(jy), O._d(jz), O._d(coeffs), coeffs.size - 1,
                           eff, 3, O._d(p))
            out[f"{key}_step_m{mode}_t{tau}"] = p
            out[f"{key}_stepinput_oracle_coeffs_m{mode}_t{tau}"] = coeffs
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({len(out)} arrays)")


if __name__ == "__main__":
    main()
