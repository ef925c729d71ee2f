"""GPU parity of libspx (through the C ABI) against the pinned CPU oracle and
the dense oracle. FP64 throughout; the device path uses FMA contraction and a
one-pass gather order, so agreement is to rounding (stated per test), except
where the arithmetic is identical by construction (diagonal, Bessel,
Chebyshev coefficients, identity steps), which is checked bitwise."""
import math
import time

import numpy as np
import pytest

import dense as D
from oracle import oracle as O
from spinsim_amd import spx

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def dipolar_system(L, index, disorder=0.5, transverse=False, anisotropic=False, angular=True):
    c, hz, _ = O.dipolar_lattice(L, 20240726, index, min_sep=0.5, angular=angular, disorder=disorder)
    fields = np.zeros((L, 3))
    fields[:, 2] = hz
    if transverse:
        fields[:, 0] = 0.7
        fields[:, 1] = -0.3
    if anisotropic:
        c.jx = c.jx * 1.25
    return fields, c


def device_ham(ctx, L, fields, c):
    h = spx.Hamiltonian(ctx, L)
    h.set_fields(fields)
    h.set_couplings(c.lo, c.hi, c.jx, c.jy, c.jz)
    return h


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ------------------------------------------------------------ Hamiltonian

@pytest.mark.parametrize("L,transverse,aniso", [(1, True, False), (3, True, True), (6, False, True),
                                                (9, True, False), (10, False, False),
                                                (11, True, True), (13, False, False),
                                                (16, True, False)])
def test_hamiltonian_apply_matches_oracle(ctx, L, transverse, aniso):
    banner(f"ham apply L={L}")
    fields, c = dipolar_system(L, L, transverse=transverse, anisotropic=aniso)
    ho = O.Ham(L, fields, c)
    h = device_ham(ctx, L, fields, c)
    assert h.spectral_bound() == ho.bound()
    psi = O.random_state(L, 3 + L)
    acc0 = O.random_state(L, 4 + L)
    sin, sout = spx.StateVector(ctx, L, psi), spx.StateVector(ctx, L)
    bound = ho.bound()
    h.apply_physical(sin, sout)
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
    # the sigma_z diagonal is the reference's, bit for bit
    assert np.array_equal(h.diagonal(), ho.diag())


def test_hamiltonian_errors(ctx):
    h = spx.Hamiltonian(ctx, 3)
    with pytest.raises(spx.OutOfRange):
        h.set_field(3, 1.0)
    with pytest.raises(spx.OutOfRange):
        h.set_coupling(0, 3, 1, 1, 1)
    with pytest.raises(spx.InvalidArgument):
        h.set_coupling(1, 1, 1, 1, 1)
    a, b = spx.StateVector(ctx, 3), spx.StateVector(ctx, 3)
    with pytest.raises(spx.DomainError):
        h.apply_twice_normalized(0.0, a, b)
    with pytest.raises(spx.InvalidArgument):
        h.apply_physical(a, spx.StateVector(ctx, 4))
    h.set_coupling(2, 0, 0.1, 0.2, 0.3)
    h.set_coupling(0, 2, 0.4, 0.5, 0.6)   # repeated pair overwrites in place
    assert h.num_couplings() == 1


# ------------------------------------------------------------ StateVector

def test_state_vector_ops(ctx):
    L = 12
    a0, b0 = O.random_state(L, 1), O.random_state(L, 2)
    a, b = spx.StateVector(ctx, L, a0), spx.StateVector(ctx, L, b0)
    assert np.array_equal(a.download(), a0)
    assert abs(a.norm() - O.lib().so_state_norm(O._d(a0), 1 << L)) < 1e-14
    re, im = a.inner(b)
    wr, wi = O.C.c_double(), O.C.c_double()
    O.lib().so_state_inner(O._d(a0), O._d(b0), 1 << L, O.C.byref(wr), O.C.byref(wi))
    assert abs(re - wr.value) < 1e-14 and abs(im - wi.value) < 1e-14
    a.add_scaled(0.3, -0.7, b)
    want = a0.copy()
    O.lib().so_state_add_scaled(O._d(want), 1 << L, 0.3, -0.7, O._d(b0))
    assert np.array_equal(a.download(), want)
    a.negate()
    a.scale(2.0)
    assert np.array_equal(a.download(), -2.0 * want)
    a.set_basis(5)
    d = a.download()
    assert d[10] == 1.0 and np.count_nonzero(d) == 1
    with pytest.raises(spx.OutOfRange):
        a.set_basis(1 << L)
    a.set_zero()
    with pytest.raises(spx.DomainError):
        a.normalize()


def test_large_transfers_round_trip_in_pieces(ctx):
    """Host copies go through a pinned bounce buffer of 4 MiB, in pieces: an
    L = 20 state (16 MiB, four pieces) uploads and downloads bitwise, a
    basis state lands in the last piece, and the L = 20 diagonal (8 MiB, two
    pieces) comes back as the oracle builds it (the same terms in the same
    order, so bitwise)."""
    L = 20
    banner(f"bounce-buffer transfers L={L}")
    psi = O.random_state(L, 9)
    s = spx.StateVector(ctx, L, psi)
    assert np.array_equal(s.download(), psi)
    s.set_basis((1 << L) - 1)
    d = s.download()
    assert d[-2] == 1.0 and np.count_nonzero(d) == 1
    fields, c = dipolar_system(L, 3)
    h = device_ham(ctx, L, fields, c)
    assert np.array_equal(h.diagonal(), O.Ham(L, fields, c).diag())


# ------------------------------------------------------ kernels/observables

@pytest.mark.parametrize("L", [1, 4, 9, 14])
def test_sigma_rotation_and_observables_match_oracle(ctx, L):
    banner(f"sigma/rotation/observables L={L}")
    psi, ref = O.random_state(L, 10 + L), O.random_state(L, 20 + L)
    sp, sr, so = spx.StateVector(ctx, L, psi), spx.StateVector(ctx, L, ref), spx.StateVector(ctx, L)
    for axis in range(3):
        for spin in sorted({0, L // 2, L - 1}):
            so.upload(ref)
            spx.apply_sigma(axis, 0.75, sp, spin, so, True)
            want = ref.copy()
            O.lib().so_axpy_sigma(axis, 0.75, O._d(psi), L, spin, O._d(want))
            assert np.array_equal(so.download(), want)
            spx.apply_sigma(axis, 1.0, sp, spin, so, False)
            O.lib().so_apply_sigma(axis, O._d(psi), L, spin, O._d(want))
            assert np.array_equal(so.download(), want)
            scratch = np.zeros_like(psi)
            m = spx.measure_mag(axis, spin, sp)
            assert abs(m - O.lib().so_mag(axis, spin, O._d(psi), L, O._d(scratch))) < 1e-13
        spx.apply_total_spin(axis, sp, so)
        want = np.zeros_like(psi)
        O.lib().so_apply_total_spin(axis, O._d(psi), L, O._d(want))
        assert np.array_equal(so.download(), want)
        re, im = spx.measure_corr(axis, sp, sr)
        wr, wi = O.C.c_double(), O.C.c_double()
        O.lib().so_corr(axis, O._d(psi), O._d(ref), L, O._d(np.zeros_like(psi)), O.C.byref(wr),
                        O.C.byref(wi))
        assert abs(re - wr.value) < 1e-12 and abs(im - wi.value) < 1e-12
    for axis, angle in (((1, 0, 0), math.pi / 2), ((0, 1, 0), math.pi), ((0.3, -0.5, 0.8), 1.7)):
        targets = list(range(L)) if L < 9 else [0, L - 1, L // 2]
        sp.upload(psi)
        spx.rotate(sp, axis, angle, targets)
        want = psi.copy()
        t = np.array(targets, np.uint32)
        O.lib().so_rotate(O._d(want), L, *axis, angle, O._u32(t), t.size)
        assert rel_err(sp.download(), want) < 1e-14
    fields, c = dipolar_system(L, 7)
    ho, h = O.Ham(L, fields, c), device_ham(ctx, L, fields, c)
    sp.upload(psi)
    e = spx.measure_energy(h, sp)
    assert abs(e - O.lib().so_energy(ho.h, O._d(psi), O._d(np.zeros_like(psi)))) < 1e-12 * max(1, abs(e))


@pytest.mark.parametrize("L", [5, 10, 13, 16, 18])
def test_rotation_of_many_spins_matches_oracle(ctx, L):
    """Rotations over ascending target lists (every spin, as a pulse step
    gives them, and a sorted subset) run as two LDS passes instead of one
    launch per target; each pair sees the per-target arithmetic and order,
    so the result is the oracle's sequential rotation to rounding."""
    banner(f"rotation of many spins L={L}")
    psi = O.random_state(L, 60 + L)
    sp = spx.StateVector(ctx, L, psi)
    for axis, angle, targets in (((0, 1, 0), math.pi, list(range(L))),
                                 ((1, 0, 0), math.pi / 2, list(range(L))),
                                 ((0.3, -0.5, 0.8), 1.7, sorted({0, 2, L // 2, L - 2, L - 1}))):
        sp.upload(psi)
        spx.rotate(sp, axis, angle, targets)
        want = psi.copy()
        t = np.array(targets, np.uint32)
        O.lib().so_rotate(O._d(want), L, *axis, angle, O._u32(t), t.size)
        assert rel_err(sp.download(), want) < 1e-14, (axis, targets)


def test_two_spin_gate_matches_oracle(ctx):
    L = 6
    g = np.random.default_rng(5)
    m = g.standard_normal((4, 4)) + 1j * g.standard_normal((4, 4))
    U, _ = np.linalg.qr(m)
    psi = O.random_state(L, 8)
    for a, b in ((0, 1), (4, 1), (2, 5)):
        s = spx.StateVector(ctx, L, psi)
        spx.two_spin_gate(s, U, a, b)
        want = psi.copy()
        u = np.ascontiguousarray(U.reshape(16)).view(np.float64)
        O.lib().so_two_spin_gate(O._d(want), L, O._d(u), a, b)
        assert rel_err(s.download(), want) < 1e-14


# --------------------------------------------------------------- Propagator

def test_propagator_matches_exact_exponential(ctx):
    """tests/test_Propagator.cpp:36-68 on the device: L<=6, eps 1e-12."""
    banner("propagator vs dense L<=6")
    for L in range(1, 7):
        f, lo, hi, jx, jy, jz = D.sample_hamiltonian(L)
        h = spx.Hamiltonian(ctx, L)
        h.set_fields(f)
        h.set_couplings(lo, hi, jx, jy, jz)
        H = D.hamiltonian(L, f, lo, hi, jx, jy, jz)
        psi0 = O.ref_test_state(L, 0x9001 + L)
        prop = spx.Propagator(ctx, L, 1e-12)
        for tau in (0.05, 0.3, 1.0):
            s = spx.StateVector(ctx, L, psi0)
            prop.evolve(h, spx.REAL, s, tau, 1)
            assert D.max_abs_diff(s.download(), D.expm_apply(H, 1j * tau, psi0)) < 1e-10
        s = spx.StateVector(ctx, L, psi0)
        prop.evolve(h, spx.IMAGINARY, s, 0.4, 1)
        assert D.max_abs_diff(s.download(), D.expm_apply(H, 0.4, psi0)) < 1e-10


@pytest.mark.parametrize("L,eps,tau,steps,mode", [(8, 1e-7, 0.05, 10, 0), (10, 1e-7, 0.05, 10, 0),
                                                  (12, 1e-12, 0.05, 4, 0), (13, 1e-7, -0.3, 2, 0),
                                                  (14, 1e-12, 0.05, 2, 0), (11, 1e-10, 0.1, 3, 1)])
def test_propagator_matches_oracle(ctx, L, eps, tau, steps, mode):
    banner(f"propagator vs oracle L={L} eps={eps} tau={tau} steps={steps} mode={mode}")
    fields, c = dipolar_system(L, 100 + L, transverse=(L == 13))
    ho, h = O.Ham(L, fields, c), device_ham(ctx, L, fields, c)
    psi0 = O.random_state(L, 30 + L)
    want = psi0.copy()
    rc, order = ho.evolve(want, tau, steps, mode, False, eps)
    assert rc == 0
    s = spx.StateVector(ctx, L, psi0)
    prop = spx.Propagator(ctx, L, eps)
    prop.evolve(h, mode, s, tau, steps)
    assert prop.last_order() == order
    got = s.download()
    scale = max(1.0, float(np.max(np.abs(want))))
    assert float(np.max(np.abs(got - want))) < 1e-12 * scale


def test_propagator_edge_cases(ctx):
    L = 5
    fields, c = dipolar_system(L, 1)
    h = device_ham(ctx, L, fields, c)
    psi0 = O.random_state(L, 99)
    prop = spx.Propagator(ctx, L, 1e-7)
    s = spx.StateVector(ctx, L, psi0)
    prop.evolve(h, spx.REAL, s, 1e-12, 1)          # identity shortcut: bitwise untouched
    assert prop.last_order() == 0 and np.array_equal(s.download(), psi0)
    prop.evolve(h, spx.REAL, s, 0.3, 0)            # zero steps: untouched
    assert np.array_equal(s.download(), psi0)
    prop.evolve(h, spx.REAL, s, 0.7, 1)
    prop.evolve(h, spx.REAL, s, -0.7, 1)           # time reversal
    assert D.max_abs_diff(s.download(), psi0) < 1e-6
    with pytest.raises(spx.ConvergenceError):
        spx.Propagator(ctx, L, 1e-7).evolve(h, spx.REAL, s, 1e5, 1)
    with pytest.raises(spx.InvalidArgument):
        prop.evolve(h, spx.REAL, spx.StateVector(ctx, L + 1), 0.1, 1)
    with pytest.raises(spx.InvalidArgument):
        spx.Propagator(ctx, L, 0.0)
    # single-spin Rabi: <sigma_z> = -cos(Hx t) (tests/test_Propagator.cpp:126-152)
    h1 = spx.Hamiltonian(ctx, 1)
    h1.set_field(0, 2.0, 0.0, 0.0)
    s1 = spx.StateVector(ctx, 1)
    s1.set_basis(0)
    p1 = spx.Propagator(ctx, 1, 1e-13)
    for step in range(1, 61):
        p1.evolve(h1, spx.REAL, s1, 0.05, 1)
        assert abs(spx.measure_mag(2, 0, s1) + math.cos(2.0 * 0.05 * step)) < 1e-9


def test_evolve_many_equals_separate_evolves_bitwise(ctx):
    L = 14
    fields, c = dipolar_system(L, 3)
    h = device_ham(ctx, L, fields, c)
    vecs = [O.random_state(L, 40 + i) for i in range(5)]
    prop = spx.Propagator(ctx, L, 1e-7)
    many = [spx.StateVector(ctx, L, v) for v in vecs]
    prop.evolve_many(h, spx.REAL, many, 0.05, 3)
    for v, m in zip(vecs, many):
        one = spx.StateVector(ctx, L, v)
        prop.evolve(h, spx.REAL, one, 0.05, 3)
        assert np.array_equal(one.download(), m.download())


def test_l18_single_step_matches_oracle_and_is_deterministic(ctx):
    """Full-size L=18 (153 pairs): one Chebyshev step against the oracle, and
    the same run twice is bitwise identical."""
    L = 18
    banner("L=18 step vs oracle")
    fields, c = dipolar_system(L, 0)
    ho, h = O.Ham(L, fields, c), device_ham(ctx, L, fields, c)
    psi0 = O.random_state(L, 18)
    want = psi0.copy()
    rc, order = ho.evolve(want, 0.05, 1, 0, False, 1e-12)
    assert rc == 0
    prop = spx.Propagator(ctx, L, 1e-12)
    outs = []
    for _ in range(2):
        s = spx.StateVector(ctx, L, psi0)
        prop.evolve(h, spx.REAL, s, 0.05, 1)
        outs.append(s.download())
    assert prop.last_order() == order
    assert np.array_equal(outs[0], outs[1])
    assert float(np.max(np.abs(outs[0] - want))) < 1e-12


def test_l18_conservation_and_time_reversal(ctx):
    """Size-independent properties at L=18, eps 1e-12, autonormalise off:
    norm and energy drift over 200 steps, and +tau then -tau returns."""
    L = 18
    banner("L=18 conservation")
    fields, c = dipolar_system(L, 1)
    h = device_ham(ctx, L, fields, c)
    psi0 = O.random_state(L, 181)
    s = spx.StateVector(ctx, L, psi0)
    e0, n0 = spx.measure_energy(h, s), s.norm()
    prop = spx.Propagator(ctx, L, 1e-12)
    prop.evolve(h, spx.REAL, s, 0.05, 200)
    # the truncated expansion drifts systematically by ~a_{K+1} per step (the
    # reference arithmetic drifts 1.3e-10 over 1e4 steps at L=5 on its own test
    # state); 200 steps stay well inside 1e-10
    assert abs(s.norm() - n0) < 1e-10
    assert abs(spx.measure_energy(h, s) - e0) < 1e-10 * max(1.0, abs(e0))
    prop.evolve(h, spx.REAL, s, -0.05, 200)
    assert D.max_abs_diff(s.download(), psi0) < 1e-10


# ------------------------------------------------------------------ Ensemble

def realisation_inputs(L, seed, index, disorder=0.5, min_sep=0.5, angular=True):
    """runRealization's draws (src/io/Experiment.cpp:10-55) from the oracle's
    pinned stream: geometry, Hz disorder, random initial state."""
    c, hz, r = O.dipolar_lattice(L, seed, index, min_sep=min_sep, angular=angular, disorder=disorder)
    psi0 = np.zeros(2 << L)
    assert O.lib().so_prepare_initial_state(O._d(psi0), L, 0, O.C.byref(r), 0) == 0
    fields = np.zeros((L, 3))
    fields[:, 2] = hz
    return fields, c, psi0


def cpmg(P, L, cycles=8):
    allspins = list(range(L))
    P.pulse((1, 0, 0), math.pi / 2, allspins)
    P.repeat(cycles)
    P.evolve(0.05, 10)
    P.pulse((0, 1, 0), math.pi, allspins)
    P.evolve(0.05, 10)
    P.measure()
    P.end()
    return P


def test_config0_cpmg_echoes_match_oracle_per_realisation(ctx):
    """Plain CPMG echo, L=10, 20 realisations (20240726, 0..19): per-realisation
    echo curves within 1e-10 of the CPU path, and the ensemble statistics."""
    banner("config0 CPMG ensemble")
    L, n = 10, 20
    obs = ((O.OBS_CORR_X, 0), (O.OBS_NORM, 0), (O.OBS_ENERGY, 0))
    cfg = O.Config(L=L, program=cpmg(O.Program(), L), observables=obs, field_disorder=0.5,
                   min_separation=0.5, angular=True)
    times, per, mean, sd, se = O.run_ensemble(cfg, n, threads=8)
    b = spx.Batch(ctx, cpmg(spx.Program(), L), obs, epsilon=1e-7, reference_axis=0)
    for i in range(n):
        fields, c, psi0 = realisation_inputs(L, 20240726, i)
        b.add(L, fields, c.lo, c.hi, c.jx, c.jy, c.jz, psi0)
    gt, gv = b.run()
    assert np.array_equal(gt, times)
    assert float(np.max(np.abs(gv - per))) < 1e-10
    assert float(np.max(np.abs(gv.mean(axis=0) - mean))) < 1e-10


def xy8(P, L, blocks, omega=50.0, oracle=False):
    """XY-8 (X Y X Y Y X Y X) with finite pulses: each pulse is an evolve of
    tau_p = pi/omega under hx = omega (X) or hy = omega (Y) on every spin,
    preceded by free evolution 10 x 0.05; a measure after each block."""
    tp = math.pi / omega
    P.measure()
    P.repeat(blocks)
    for ax in "XYXYYXYX":
        P.evolve(0.05, 10)
        hx, hy = (omega, 0.0) if ax == "X" else (0.0, omega)
        if oracle:
            f = np.zeros((L, 3))
            f[:, 0], f[:, 1] = hx, hy
            P.evolve(tp, 1, fields=f)
        else:
            P.evolve_uniform(tp, 1, hx=hx, hy=hy, hz=0.0)
    P.measure()
    P.end()
    return P


def test_resident_batch_uneven_sizes_finite_pulses_match_oracle(ctx):
    """The batched resident runner on realisations of L = 8..12 (L from the
    index), finite XY-8 pulses: every realisation's rows within 1e-10 of the
    oracle's, and a second run is bitwise identical."""
    banner("resident batch uneven L XY-8")
    obs = ((O.OBS_CORR_X, 0), (O.OBS_NORM, 0), (O.OBS_MAG_Z, 1), (O.OBS_ENERGY, 0))
    n = 10
    Ls = [8 + (i % 5) for i in range(n)]
    b = spx.Batch(ctx, xy8(spx.Program(), 0, 1), obs, epsilon=1e-7, reference_axis=0)
    want = []
    for i, L in enumerate(Ls):
        fields, c, psi0 = realisation_inputs(L, 20240726, i)
        b.add(L, fields, c.lo, c.hi, c.jx, c.jy, c.jz, psi0)
        cfg = O.Config(L=L, program=xy8(O.Program(), L, 1, oracle=True), observables=obs,
                       field_disorder=0.5, min_separation=0.5, angular=True)
        t, v = O.run_realization(cfg, i)
        want.append(v)
    gt, gv = b.run()
    assert np.allclose(gt, t)
    for i in range(n):
        err = float(np.max(np.abs(gv[i] - want[i])))
        assert err < 1e-10, (i, Ls[i], err)
    _, gv2 = b.run()
    assert np.array_equal(gv, gv2)


def test_resident_batch_autonormalised_steps_match_oracle(ctx):
    """Evolve steps with autonormalisation in the resident runner, in
    imaginary and real time, at every walker (L = 8..12: pair gathers up to
    9, two vectors held as one at 10, half-workgroup vectors at 11, the
    whole workgroup at 12): rows within 1e-10 of the oracle's."""
    banner("resident batch, autonormalised imaginary and real time at L = 8..12")
    obs = ((O.OBS_CORR_X, 0), (O.OBS_NORM, 0), (O.OBS_ENERGY, 0))

    def program(P):
        P.measure()
        P.evolve(0.1, 4, mode=1, autonorm=True)
        P.measure()
        P.evolve(0.05, 3, mode=0, autonorm=True)
        P.measure()
        return P

    b = spx.Batch(ctx, program(spx.Program()), obs, epsilon=1e-7, reference_axis=0)
    want = []
    Ls = list(range(8, 13))
    for i, L in enumerate(Ls):
        fields, c, psi0 = realisation_inputs(L, 20240726, 40 + i)
        b.add(L, fields, c.lo, c.hi, c.jx, c.jy, c.jz, psi0)
        cfg = O.Config(L=L, program=program(O.Program()), observables=obs, field_disorder=0.5,
                       min_separation=0.5, angular=True)
        want.append(O.run_realization(cfg, 40 + i)[1])
    _, gv = b.run()
    for i, L in enumerate(Ls):
        err = float(np.max(np.abs(gv[i] - want[i])))
        assert err < 1e-10, (L, err)
        assert np.allclose(gv[i][1:, 2:4], 1.0, atol=1e-12), L   # both norms after each step


def test_resident_batch_full_xy8_program_matches_oracle(ctx):
    """The uneven-size batched workload's whole pulse program: XY-8 with
    finite pulses repeated 16 times (128 pulses), a measure before and after
    every block, on two realisations of each L = 8..12 drawn inside libspx
    (spx_batch_add_dipolar: lattice, disorder and the random initial state
    from the realisation's stream); every row within 1e-10 of the oracle's
    runRealization."""
    from concurrent.futures import ThreadPoolExecutor
    banner("resident batch, XY-8 x 16 at L = 8..12: oracle (CPU, about 20 s)")
    Ls = list(range(8, 13))
    obs = ((O.OBS_CORR_X, 0), (O.OBS_NORM, 0))

    def oracle_rows(L):
        return O.run_ensemble(O.dipolar_config(L, O.xy8_program(L), observables=obs), 2, threads=2)[1]

    with ThreadPoolExecutor(len(Ls)) as pool:
        want = list(pool.map(oracle_rows, Ls))
    b = spx.Batch(ctx, xy8(spx.Program(), 0, 16), obs, epsilon=1e-7, reference_axis=0)
    for L in Ls:
        b.add_dipolar(20240726, 0, 2, L)
    _, gv = b.run()
    assert gv.shape[0] == 2 * len(Ls) and gv.shape[1] == 17
    for g, L in enumerate(Ls):
        for i in range(2):
            err = float(np.max(np.abs(gv[2 * g + i] - want[g][i])))
            assert err < 1e-10, (L, i, err)
    banner("resident batch, XY-8 x 16: every row within 1e-10")


def test_imaginary_time_relaxes_to_the_ground_state(ctx):
    """Imaginary-time propagation of -H with autonormalisation (the ground-state
    run of the single-large-state workload) reaches the lowest eigenvalue of H
    from a dense eigensolver, L=10, and its eigenvector."""
    L = 10
    banner("imaginary time ground state L=10")
    for index in range(7, 40):  # a realisation with a gap, so the relaxation converges fast
        fields, c = dipolar_system(L, index)
        H = D.hamiltonian(L, fields, c.lo, c.hi, c.jx, c.jy, c.jz)
        w, V = np.linalg.eigh(H)
        if w[1] - w[0] > 0.05:
            break
    assert w[1] - w[0] > 0.05
    hneg = device_ham(ctx, L, -fields, type(c)(c.lo, c.hi, -c.jx, -c.jy, -c.jz))
    hpos = device_ham(ctx, L, fields, c)
    s = spx.StateVector(ctx, L, O.random_state(L, 3))
    prop = spx.Propagator(ctx, L, 1e-12)
    tau = 0.1
    steps = int(math.ceil(40.0 / ((w[1] - w[0]) * tau)))
    prop.evolve(hneg, spx.IMAGINARY, s, tau, steps, autonormalize=True)
    e = spx.measure_energy(hpos, s)
    assert abs(s.norm() - 1.0) < 1e-12
    assert abs(e - w[0]) < 1e-9 * max(1.0, abs(w[0])), (e, w[0])
    overlap = abs(np.vdot(V[:, 0], D.to_complex(s.download())))
    assert overlap > 1.0 - 1e-9


def test_benchmarked_arrangement_matches_the_oracle(ctx):
    """The headline workload exactly as bench.py runs it: realisations
    (20240726, 0..11) at L = 18 from bench.synthetic_system, expansions built
    at dt = 0.05 and eps = 1e-12, the schedule the tuner picks at the set's
    one-stream width and adopted through the tuner's cache, two halves of
    six on two streams, three chained steps through steps_set. Every
    realisation agrees with the oracle's three steps amplitude by amplitude
    to 1e-12 of its largest amplitude, and the kernels are the Hadamard
    path's."""
    import bench
    L = 18
    S = 12
    banner(f"benchmarked arrangement: oracle three steps of {S} L=18 realisations (CPU, about 25 s)")
    systems = [bench.synthetic_system(L, index=i) for i in range(S)]
    hams, states, wants = [], [], []
    for f, lo, hi, jx, jy, jz, p0 in systems:
        x = p0.copy()
        rc, _ = O.Ham(L, f, O.Couplings(lo, hi, jx, jy, jz)).evolve(x, 0.05, 3, eps=1e-12)
        assert rc == 0
        wants.append(x)
        h = spx.Hamiltonian(ctx, L)
        h.set_fields(f)
        h.set_couplings(lo, hi, jx, jy, jz)
        hams.append(h)
        states.append(spx.StateVector(ctx, L, p0))
    built = [spx.cheb_build(spx.REAL, h.spectral_bound(), 0.05, 1e-12) for h in hams]
    banner("benchmarked arrangement: tune, then three steps on the device")
    hams[0].autotune(built[0][1], nvec=S)
    spx.autotune_mode(1)
    try:
        prop = spx.Propagator(ctx, L, 1e-12)
        prop.set_streams(2)
        prop.steps_set(hams, spx.REAL, states, [c for c, _ in built], [e for _, e in built], 3)
    finally:
        spx.autotune_mode(0)
    for i, (s, w) in enumerate(zip(states, wants)):
        err = rel_err(s.download(), w)
        banner(f"realisation {i}: {hams[i].kernel_name()} rel err {err:.2e}")
        assert err < 1e-12, (i, err)
        assert hams[i].kernel_name().startswith("k_wht_"), hams[i].kernel_name()


def test_l18_drift_over_1e4_steps_is_the_expansions_own(ctx):
    """The single-large-state drift check at full length: eight L = 18
    realisations (indices 0..7, two halves of four on two streams), 1e4 real
    steps of 0.05 at eps 1e-12, autonormalisation off. The truncated
    expansion changes the norm and the energy by a fixed amount per step (it
    depends only on the state's spectral weights, which the evolution keeps),
    so after 1e4 steps each realisation's drift must be 5000 times what the
    oracle's own two steps give for it: the device reproduces the method's
    drift to 1e-3, and adds none of its own."""
    L, n = 18, 1 << 18
    steps = 10_000
    banner("L=18 drift: oracle two steps per realisation (CPU, about 10 s)")
    hams, states, want_dn, want_de, n0, e0 = [], [], [], [], [], []
    for i in range(8):
        fields, c = dipolar_system(L, i)
        psi = O.random_state(L, 900 + i)
        oh = O.Ham(L, fields, c)
        energy = lambda v: float(np.dot(v, oh.apply_physical(v)))
        x = psi.copy()
        rc, _ = oh.evolve(x, 0.05, 2, eps=1e-12)
        assert rc == 0
        want_dn.append((np.linalg.norm(x) - np.linalg.norm(psi)) * steps / 2)
        want_de.append((energy(x) - energy(psi)) * steps / 2)
        hams.append(device_ham(ctx, L, fields, c))
        states.append(spx.StateVector(ctx, L, psi))
        n0.append(states[-1].norm())
        e0.append(spx.measure_energy(hams[-1], states[-1]))
    banner("L=18 drift: 1e4 steps of eight realisations on the device")
    prop = spx.Propagator(ctx, L, 1e-12)
    for _ in range(10):
        prop.evolve_set(hams, spx.REAL, states, 0.05, steps // 10)
    for i in range(8):
        dn = states[i].norm() - n0[i]
        de = spx.measure_energy(hams[i], states[i]) - e0[i]
        assert abs(dn - want_dn[i]) <= 1e-3 * abs(want_dn[i]) + 1e-12, (i, dn, want_dn[i])
        assert abs(de - want_de[i]) <= 1e-3 * abs(want_de[i]) + 1e-12, (i, de, want_de[i])


def test_imaginary_time_ground_state_l18(ctx):
    """The ground-state run of the single-large-state workload at its own size:
    L = 18, realisation 7 of the dipolar ensemble (153 pairs), a state and a
    second one relaxed together (two states per launch: the Hadamard path) in
    imaginary time under -H with autonormalisation. Both reach the lowest
    eigenvalue that Lanczos (scipy's eigsh, 233 products on the CPU) finds
    with the oracle's H·ψ, and the same state. The gap to the next level is
    0.258, so 60 time units leave exp(-31) of any excited component."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    L, n = 18, 1 << 18
    fields, c = dipolar_system(L, 7)
    oh = O.Ham(L, fields, c)

    def matvec(v):
        x = np.zeros(2 * n)
        x[0::2] = np.asarray(v).ravel()
        return oh.apply_physical(x)[0::2].copy()

    banner("L=18 ground energy by Lanczos over the oracle's H (CPU, about 20 s)")
    w = eigsh(LinearOperator((n, n), matvec=matvec, dtype=float), k=2, which="SA", tol=1e-12,
              v0=np.ones(n) / math.sqrt(n), return_eigenvectors=False)
    e0, e1 = sorted(w)
    assert e1 - e0 > 0.25, (e0, e1)
    banner(f"L=18 imaginary time, E0={e0:.12f}")
    hneg = device_ham(ctx, L, -fields, type(c)(c.lo, c.hi, -c.jx, -c.jy, -c.jz))
    hpos = device_ham(ctx, L, fields, c)
    states = [spx.StateVector(ctx, L, O.random_state(L, seed)) for seed in (5, 6)]
    prop = spx.Propagator(ctx, L, 1e-12)
    prop.evolve_many(hneg, spx.IMAGINARY, states, 0.1, 600, autonormalize=True)
    for s in states:
        assert abs(s.norm() - 1.0) < 1e-12
        e = spx.measure_energy(hpos, s)
        assert abs(e - e0) < 1e-9 * abs(e0), (e, e0)
    a, b = (D.to_complex(s.download()) for s in states)
    assert abs(np.vdot(a, b)) > 1.0 - 1e-9


@pytest.mark.parametrize("L", [9, 14, 16, 18])
def test_step_graph_replays_equal_eager_launches_bitwise(ctx, L):
    """Evolve calls that record one step as a HIP graph and replay it agree
    bitwise with the same calls launched order by order, over several calls
    with eager work (gates, measurements, host syncs) between them; also for
    several states at once and in imaginary time. At L = 18 the two-state
    calls take the Hadamard path (chained order launches)."""
    banner(f"graph replay vs eager L={L}")
    fields, c = dipolar_system(L, 21, transverse=(L == 9))
    h = device_ham(ctx, L, fields, c)
    psi0, ref0 = O.random_state(L, 5), O.random_state(L, 6)
    pg, pe = spx.Propagator(ctx, L, 1e-10), spx.Propagator(ctx, L, 1e-10)
    pg.set_graphs(True)
    pe.set_graphs(False)
    sg, se = spx.StateVector(ctx, L, psi0), spx.StateVector(ctx, L, psi0)
    rg, re = spx.StateVector(ctx, L, ref0), spx.StateVector(ctx, L, ref0)
    for rep in range(3):
        pg.evolve(h, spx.REAL, sg, 0.05, 6)
        pe.evolve(h, spx.REAL, se, 0.05, 6)
        ctx.synchronize()
        assert np.array_equal(sg.download(), se.download()), rep
        spx.rotate(sg, (0, 1, 0), math.pi / 3, list(range(L)))
        spx.rotate(se, (0, 1, 0), math.pi / 3, list(range(L)))
        assert spx.measure_mag(0, 0, sg) == spx.measure_mag(0, 0, se)
        pg.evolve_many(h, spx.REAL, [sg, rg], 0.05, 5)
        pe.evolve_many(h, spx.REAL, [se, re], 0.05, 5)
        assert np.array_equal(sg.download(), se.download()) and np.array_equal(rg.download(), re.download())
    pg.evolve(h, spx.IMAGINARY, sg, 0.02, 8)
    pe.evolve(h, spx.IMAGINARY, se, 0.02, 8)
    assert np.array_equal(sg.download(), se.download())


def test_streaming_batch_mid_size_matches_oracle_and_repeats(ctx):
    """Realisations above the LDS-resident size (L = 13, 14) go through the
    streaming batch path (worker contexts running realisations concurrently):
    every realisation's rows match the oracle's within 1e-10, and a second
    run is bitwise identical."""
    banner("streaming batch L=13/14")
    obs = ((O.OBS_CORR_X, 0), (O.OBS_NORM, 0), (O.OBS_MAG_Z, 2))
    Ls = [13, 14, 13, 14, 13, 14]
    b = spx.Batch(ctx, cpmg(spx.Program(), 0, cycles=3), obs, epsilon=1e-7, reference_axis=0)
    want = []
    for i, L in enumerate(Ls):
        fields, c, psi0 = realisation_inputs(L, 20240726, 500 + i)
        b.add(L, fields, c.lo, c.hi, c.jx, c.jy, c.jz, psi0)
        cfg = O.Config(L=L, program=cpmg(O.Program(), L, cycles=3), observables=obs, field_disorder=0.5,
                       min_separation=0.5, angular=True)
        want.append(O.run_realization(cfg, 500 + i)[1])
    t0 = time.perf_counter()
    _, gv = b.run()
    banner(f"streaming batch of {len(Ls)}: {time.perf_counter() - t0:.3f} s")
    for i in range(len(Ls)):
        err = float(np.max(np.abs(gv[i] - want[i])))
        assert err < 1e-10, (i, Ls[i], err)
    _, gv2 = b.run()
    assert np.array_equal(gv, gv2)


def test_streaming_batch_l15_group_of_eight_matches_oracle(ctx):
    """Eight L = 15 realisations walk as one lockstep group, sixteen vectors
    per launch, which at L = 15 takes the Hadamard path for the flip-flop
    couplings between pulses (the gather kernel below that width): the first
    two realisations match the oracle within 1e-10, and a second run is
    bitwise identical."""
    L, n = 15, 8
    banner(f"streaming batch L={L} x {n}")
    obs = ((O.OBS_CORR_X, 0), (O.OBS_NORM, 0))
    b = spx.Batch(ctx, cpmg(spx.Program(), L, cycles=2), obs, epsilon=1e-7, reference_axis=0)
    for i in range(n):
        fields, c, psi0 = realisation_inputs(L, 20240726, 900 + i)
        b.add(L, fields, c.lo, c.hi, c.jx, c.jy, c.jz, psi0)
    gt, gv = b.run()
    cfg = O.Config(L=L, program=cpmg(O.Program(), L, cycles=2), observables=obs, field_disorder=0.5,
                   min_separation=0.5, angular=True)
    for i in range(2):
        t, v = O.run_realization(cfg, 900 + i)
        assert np.allclose(gt, t, rtol=0, atol=1e-12)
        err = float(np.max(np.abs(gv[i] - v)))
        banner(f"L={L} realisation {i}: max |device - oracle| {err:.2e}")
        assert err < 1e-10, (i, err)
    _, gv2 = b.run()
    assert np.array_equal(gv, gv2)
