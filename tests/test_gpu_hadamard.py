"""The Hadamard-sandwich path for flip-flop-only couplings (schedule dense=2):
H psi through Walsh-Hadamard transforms instead of one gather per coupled
pair. Every application kind against the oracle, evolution against the
table and all-pairs kernels up to L=20, batched states bitwise equal to
single ones, conservation at L=18, transverse fields and flip-both
weights, and the fallbacks outside its sizes. FP64; agreement to rounding."""
import numpy as np
import pytest

import dense as D
from oracle import oracle as O
from spinsim_amd import spx

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def dipolar_system(L, index, transverse=False, anisotropic=False):
    c, hz, _ = O.dipolar_lattice(L, 20240726, index, min_sep=0.5, angular=True, disorder=0.5)
    fields = np.zeros((L, 3))
    fields[:, 2] = hz
    if transverse:
        fields[:, 0] = 0.7
    if anisotropic:
        c.jx = c.jx * 1.25
    return fields, c


def device_ham(ctx, L, fields, c, dense):
    h = spx.Hamiltonian(ctx, L)
    h.set_fields(fields)
    h.set_couplings(c.lo, c.hi, c.jx, c.jy, c.jz)
    h.set_schedule(dense=dense)
    return h


@pytest.mark.parametrize("L", [14, 16])
def test_apply_kinds_match_the_oracle(ctx, L):
    banner(f"hadamard apply L={L}")
    fields, c = dipolar_system(L, L)
    ho = O.Ham(L, fields, c)
    h = device_ham(ctx, L, fields, c, 2)
    psi = O.random_state(L, 3 + L)
    acc0 = O.random_state(L, 4 + L)
    sin, sout = spx.StateVector(ctx, L, psi), spx.StateVector(ctx, L)
    bound = ho.bound()
    h.apply_physical(sin, sout)
    assert h.kernel_name().startswith("k_wht_"), h.kernel_name()
    assert rel_err(sout.download(), ho.apply_physical(psi)) < 1e-13
    want = np.zeros_like(psi)
    O.lib().so_ham_apply_twice_normalized(ho.h, bound, O._d(psi), O._d(want))
    h.apply_twice_normalized(bound, sin, sout)
    assert rel_err(sout.download(), want) < 1e-13
    want = acc0.copy()
    O.lib().so_ham_accumulate_twice_normalized(ho.h, bound, O._d(psi), O._d(want))
    sout.upload(acc0)
    h.accumulate_twice_normalized(bound, sin, sout)
    assert rel_err(sout.download(), want) < 1e-13


def test_evolution_matches_the_oracle_at_l14(ctx):
    L = 14
    banner("hadamard evolve vs oracle L=14")
    fields, c = dipolar_system(L, 2)
    ho = O.Ham(L, fields, c)
    psi0 = O.random_state(L, 41)
    want = psi0.copy()
    rc, _ = ho.evolve(want, 0.05, 3, 0, False, 1e-12)
    assert rc == 0
    h = device_ham(ctx, L, fields, c, 2)
    s = spx.StateVector(ctx, L, psi0)
    spx.Propagator(ctx, L, 1e-12).evolve(h, spx.REAL, s, 0.05, 3)
    assert float(np.max(np.abs(s.download() - want))) < 1e-12 * max(1.0, float(np.max(np.abs(want))))


@pytest.mark.parametrize("L", [18, 20])
def test_evolution_matches_the_gather_kernels(ctx, L):
    banner(f"hadamard evolve vs gather kernels L={L}")
    fields, c = dipolar_system(L, 3)
    psi0 = O.random_state(L, 5)
    prop = spx.Propagator(ctx, L, 1e-12)
    got = {}
    for dense in ((2, 1, 0) if L <= 18 else (2, 0)):
        h = device_ham(ctx, L, fields, c, dense)
        s = spx.StateVector(ctx, L, psi0)
        prop.evolve(h, spx.REAL, s, 0.05, 2)
        got[dense] = (s.download(), h.kernel_name())
    assert got[2][1].startswith("k_wht_") and not got[0][1].startswith("k_wht")
    scale = max(1.0, float(np.max(np.abs(got[0][0]))))
    for dense in got:
        assert float(np.max(np.abs(got[dense][0] - got[2][0]))) < 1e-12 * scale, dense


def test_batched_states_equal_single_ones_bitwise(ctx):
    L = 16
    banner("hadamard batched states L=16")
    fields, c = dipolar_system(L, 4)
    h = device_ham(ctx, L, fields, c, 2)
    prop = spx.Propagator(ctx, L, 1e-12)
    psis = [O.random_state(L, 60 + i) for i in range(3)]
    many = [spx.StateVector(ctx, L, p) for p in psis]
    prop.evolve_many(h, spx.REAL, many, 0.05, 2)
    for p, m in zip(psis, many):
        one = spx.StateVector(ctx, L, p)
        prop.evolve(h, spx.REAL, one, 0.05, 2)
        assert np.array_equal(one.download(), m.download())


def test_l18_conservation_and_time_reversal(ctx):
    """Norm and energy over 200 steps at eps 1e-12 with autonormalisation
    off, then -tau returns to the initial state."""
    L = 18
    banner("hadamard L=18 conservation")
    fields, c = dipolar_system(L, 1)
    h = device_ham(ctx, L, fields, c, 2)
    psi0 = O.random_state(L, 181)
    s = spx.StateVector(ctx, L, psi0)
    e0, n0 = spx.measure_energy(h, s), s.norm()
    prop = spx.Propagator(ctx, L, 1e-12)
    prop.evolve(h, spx.REAL, s, 0.05, 200)
    assert h.kernel_name().startswith("k_wht_")
    assert abs(s.norm() - n0) < 1e-10
    assert abs(spx.measure_energy(h, s) - e0) < 1e-10 * max(1.0, abs(e0))
    prop.evolve(h, spx.REAL, s, -0.05, 200)
    assert D.max_abs_diff(s.download(), psi0) < 1e-10


@pytest.mark.parametrize("L", [12, 21])
def test_falls_back_outside_its_sizes(ctx, L):
    banner(f"hadamard fallback L={L}")
    fields, c = dipolar_system(L, 7)
    h = device_ham(ctx, L, fields, c, 2)
    s, o = spx.StateVector(ctx, L, O.random_state(L, 9)), spx.StateVector(ctx, L)
    h.apply_physical(s, o)
    assert not h.kernel_name().startswith("k_wht"), h.kernel_name()


@pytest.mark.parametrize("L,transverse,aniso", [(13, True, True), (14, True, False), (14, False, True), (16, True, True)])
def test_transverse_fields_and_flip_both_weights(ctx, L, transverse, aniso):
    """Pulses (hx, hy) and jx != jy: the Y-string sandwich S W B W S^+ and
    the fields in A and B, against the oracle's apply and evolution."""
    banner(f"hadamard general H L={L} transverse={transverse} aniso={aniso}")
    fields, c = dipolar_system(L, 11, transverse=transverse, anisotropic=aniso)
    if transverse:
        fields[:, 1] = np.linspace(-0.6, 0.9, L)   # hy as well as hx
    ho = O.Ham(L, fields, c)
    h = device_ham(ctx, L, fields, c, 2)
    psi = O.random_state(L, 12)
    sin, sout = spx.StateVector(ctx, L, psi), spx.StateVector(ctx, L)
    h.apply_physical(sin, sout)
    assert h.kernel_name().startswith("k_wht_"), h.kernel_name()
    assert rel_err(sout.download(), ho.apply_physical(psi)) < 1e-13
    want = psi.copy()
    rc, _ = ho.evolve(want, 0.05, 2, 0, False, 1e-12)
    assert rc == 0
    s = spx.StateVector(ctx, L, psi)
    spx.Propagator(ctx, L, 1e-12).evolve(h, spx.REAL, s, 0.05, 2)
    assert float(np.max(np.abs(s.download() - want))) < 1e-12 * max(1.0, float(np.max(np.abs(want))))


def test_imaginary_time_matches_the_table_kernels(ctx):
    L = 15
    banner("hadamard imaginary time vs table kernels L=15")
    fields, c = dipolar_system(L, 6)
    psi0 = O.random_state(L, 8)
    prop = spx.Propagator(ctx, L, 1e-12)
    got = {}
    for dense in (2, 0):
        h = device_ham(ctx, L, fields, c, dense)
        s = spx.StateVector(ctx, L, psi0)
        prop.evolve(h, spx.IMAGINARY, s, 0.05, 3, autonormalize=True)
        got[dense] = (s.download(), h.kernel_name())
    assert got[2][1].startswith("k_wht_") and not got[0][1].startswith("k_wht")
    assert float(np.max(np.abs(got[2][0] - got[0][0]))) < 1e-12 * max(1.0, float(np.max(np.abs(got[0][0]))))
