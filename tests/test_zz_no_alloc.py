"""spx_prop_evolve_noalloc (the allocation-free evolve, through the C ABI)
against spx_prop_evolve on the same inputs: two calls in a row, the second
at a smaller size, real and imaginary time, with transverse fields and
flip-both weights present; device memory in use does not change across the
calls; sizes beyond the static scratch are rejected."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from spinsim_amd import spx

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def device_free_bytes():
    """hipMemGetInfo's free bytes on the current device (libamdhip64)."""
    hip = None
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so", "libamdhip64.so.7"):
        try:
            hip = C.CDLL(name)
            break
        except OSError:
            continue
    assert hip is not None
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def system(L, index, transverse, aniso):
    c, hz, _ = O.dipolar_lattice(L, 20240726, index, min_sep=0.5, angular=True, disorder=0.5)
    f = np.zeros((L, 3))
    f[:, 2] = hz
    if transverse:
        f[:, 0] = 0.6
        f[:, 1] = -0.25
    if aniso:
        c.jx = c.jx * 1.3
    return f, c


def ham(ctx, L, f, c):
    h = spx.Hamiltonian(ctx, L)
    h.set_fields(f)
    h.set_couplings(c.lo, c.hi, c.jx, c.jy, c.jz)
    return h


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("mode", [spx.REAL, spx.IMAGINARY])
def test_noalloc_evolve_matches_the_propagator(ctx, mode):
    cases = [(14, 3, True, True, 0.05, 3), (9, 5, False, False, 0.2, 2)]   # second call smaller
    # objects first, so that their allocations are done before the reading
    made = []
    for L, index, tr, an, tau, steps in cases:
        f, c = system(L, index, tr, an)
        psi0 = O.random_state(L, 40 + index)
        h = ham(ctx, L, f, c)
        want = spx.StateVector(ctx, L, psi0)
        got = spx.StateVector(ctx, L, psi0)
        prop = spx.Propagator(ctx, L, 1e-10)
        prop.evolve(h, mode, want, tau, steps)
        made.append((L, h, got, want, prop.last_order(), tau, steps))
    # a first call on the device (the scratch's address is looked up there),
    # on a state of its own
    warm = spx.StateVector(ctx, made[1][0], O.random_state(made[1][0], 1))
    spx.evolve_noalloc(made[1][1], mode, warm, 0.05, 1, 1e-10)
    ctx.synchronize()
    # the two calls back to back, device memory read around them only
    before = device_free_bytes()
    orders = []
    for L, h, got, want, order, tau, steps in made:
        orders.append(spx.evolve_noalloc(h, mode, got, tau, steps, 1e-10))
        ctx.synchronize()
    after = device_free_bytes()
    banner(f"device free bytes before {before}, after {after}")
    assert after == before
    for (L, h, got, want, order, tau, steps), k in zip(made, orders):
        err = rel(got.download(), want.download())
        banner(f"no-alloc evolve L={L} mode={mode} order {k}: rel err {err:.2e}")
        assert k == order
        assert err < 1e-12
    with pytest.raises(spx.OutOfRange):
        L = 17
        spx.evolve_noalloc(spx.Hamiltonian(ctx, L), mode, spx.StateVector(ctx, L), 0.05, 1, 1e-10)
