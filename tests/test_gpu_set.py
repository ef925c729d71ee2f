"""Propagator.evolve_set: realisations (one Hamiltonian each, as in
LocalRunner::run, src/ensemble/EnsembleRunner.cpp:27-88) stepped together,
up to sixteen per order launch, each with its own spectral bound and
expansion order. Every state must come out bitwise as when evolved alone."""
import numpy as np
import pytest

from oracle import oracle as O
from spinsim_amd import spx

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def realisation(L, index, scale=1.0, transverse=False):
    c, hz, _ = O.dipolar_lattice(L, 20240726, index, min_sep=0.5, angular=True, disorder=0.5)
    fields = np.zeros((L, 3))
    fields[:, 2] = hz * scale
    if transverse:
        fields[:, 0] = 0.4
    c.jx, c.jy, c.jz = c.jx * scale, c.jy * scale, c.jz * scale
    return fields, c


def device_ham(ctx, L, fields, c, dense):
    h = spx.Hamiltonian(ctx, L)
    h.set_fields(fields)
    h.set_couplings(c.lo, c.hi, c.jx, c.jy, c.jz)
    h.set_schedule(dense=dense)
    return h


def check_set(ctx, L, specs, dense, nsteps=3, tau=0.05):
    """specs: (index, scale, transverse) per realisation."""
    prop = spx.Propagator(ctx, L, 1e-12)
    hams = [device_ham(ctx, L, *realisation(L, i, s, t), dense) for i, s, t in specs]
    psis = [O.random_state(L, 300 + k) for k in range(len(specs))]
    orders, singles = [], []
    for h, p in zip(hams, psis):
        s = spx.StateVector(ctx, L, p)
        prop.evolve(h, spx.REAL, s, tau, nsteps)
        orders.append(prop.last_order())
        singles.append(s.download())
    states = [spx.StateVector(ctx, L, p) for p in psis]
    prop.evolve_set(hams, spx.REAL, states, tau, nsteps)
    assert prop.last_order() == max(orders)
    for k, (s, want) in enumerate(zip(states, singles)):
        assert np.array_equal(s.download(), want), (k, orders[k])
    return orders, [h.kernel_name() for h in hams]


@pytest.mark.parametrize("dense", [2, 1])
def test_uneven_orders_bitwise_equal_to_single_evolves(ctx, dense):
    L = 16
    banner(f"evolve_set L={L} dense={dense}")
    # orders: x3 couplings (long), x1 (typical), x1e-6 (order 1), six states
    specs = [(0, 1.0, False), (1, 3.0, False), (2, 1e-6, False), (3, 1.0, False), (4, 0.5, False), (5, 1.0, False)]
    orders, names = check_set(ctx, L, specs, dense)
    banner(f"orders {orders}")
    assert 1 in orders and len(set(orders)) >= 3
    prefix = "k_wht_" if dense == 2 else "k_dense"
    assert all(n.startswith(prefix) for n in names), names


@pytest.mark.parametrize("L,count", [(16, 8), (15, 16)])
def test_xcd_placed_launches_bitwise_equal_to_single_evolves(ctx, L, count):
    """Launches of 8 or 16 vectors place each vector's workgroups on one XCD
    (up to L = 17); the launches narrow below 8 as the shorter expansions
    finish. Every state equals its single evolve (one vector, no placement)
    bit for bit, transverse fields included."""
    banner(f"evolve_set L={L} dense=2, {count} states (XCD placement)")
    scales = [1.0, 3.0, 0.5, 1.0, 2.0, 1e-6, 1.0, 0.7]
    specs = [(k, scales[k % 8], k % 3 == 0) for k in range(count)]
    orders, names = check_set(ctx, L, specs, 2, nsteps=2)
    banner(f"orders {orders}")
    assert len(set(orders)) >= 3
    assert all(n.startswith("k_wht_") for n in names), names


def test_chained_steps_bitwise_equal_to_single_evolves(ctx):
    """Several steps of one launch group on the Hadamard path: each state's
    last order leaves the row transforms of the finished state for the next
    step's first order, which skips its forward row pass. The states must
    equal single evolves (unchained) bit for bit, transverse field included."""
    L = 16
    banner(f"evolve_set L={L} dense=2, chained steps")
    specs = [(0, 1.0, False), (1, 3.0, False), (3, 1.0, False), (4, 0.5, False), (5, 1.0, True)]
    orders, names = check_set(ctx, L, specs, 2, nsteps=4)
    banner(f"orders {orders}")
    assert min(orders) >= 2 and len(set(orders)) >= 2
    assert all(n.startswith("k_wht_") for n in names), names


def test_more_states_than_one_launch_takes(ctx):
    L = 14
    banner("evolve_set L=14, ten realisations (two launch groups)")
    specs = [(i, 1.0 + 0.25 * (i % 3), False) for i in range(10)]
    check_set(ctx, L, specs, 2, nsteps=2)


def test_mixed_paths_fall_back_to_one_by_one(ctx):
    L = 14
    banner("evolve_set mixed paths L=14")
    # dense=1: gathers for the flip-flop members, table kernels for the ones
    # with a transverse field; the two cannot share launches
    specs = [(0, 1.0, False), (1, 1.0, True), (2, 1.0, False), (3, 2.0, True), (4, 1.0, False)]
    orders, names = check_set(ctx, L, specs, 1)
    assert names[1].startswith("k_order3") or names[1].startswith("k_horder"), names


def test_l18_hadamard_set_of_four(ctx):
    L = 18
    banner("evolve_set L=18 dense=2, four realisations")
    check_set(ctx, L, [(i, 1.0, False) for i in range(4)], 2, nsteps=2)


@pytest.mark.parametrize("count", [4, 7, 8, 12, 18])
def test_two_stream_halves_bitwise_equal_to_one_stream(ctx, count):
    """A set of 4..8 states under distinct Hamiltonians on the Hadamard path
    runs as two halves on two streams (the default); with one stream it runs
    as one launch group. Larger sets alternate their launch groups between
    the streams (12: two groups of 6; 18: groups of 8, 8 and 2). The streams
    must give the same bits as one stream, also over chained steps, and the
    context's stream must see their results (the download below is ordered
    on it)."""
    L = 16
    banner(f"evolve_set L={L} dense=2, {count} states, two streams vs one")
    specs = [(i, 1.0 + 0.5 * (i % 3), i == 2) for i in range(count)]
    hams = [device_ham(ctx, L, *realisation(L, i, s, t), 2) for i, s, t in specs]
    psis = [O.random_state(L, 700 + k) for k in range(count)]
    out = {}
    for streams in (1, 2):
        prop = spx.Propagator(ctx, L, 1e-12)
        prop.set_streams(streams)
        states = [spx.StateVector(ctx, L, p) for p in psis]
        prop.evolve_set(hams, spx.REAL, states, 0.05, 3)
        out[streams] = [st.download() for st in states]
    for k in range(count):
        assert np.array_equal(out[1][k], out[2][k]), k
    assert all(h.kernel_name().startswith("k_wht_") for h in hams)
    with pytest.raises(spx.InvalidArgument):
        spx.Propagator(ctx, L, 1e-12).set_streams(3)


@pytest.mark.parametrize("pairs", [2, 3, 4])
def test_two_stream_halves_with_state_and_reference_pairs(ctx, pairs):
    """Each realisation's state and reference under its one Hamiltonian, side
    by side (as the batch walker passes them): the halves split between
    pairs (three pairs: four states and two) and match one stream bitwise
    over chained steps."""
    L = 16
    banner(f"evolve_set L={L} dense=2, {pairs} state/reference pairs, two streams vs one")
    hams = [device_ham(ctx, L, *realisation(L, i, 1.0 + 0.5 * i, i == 1), 2) for i in range(pairs)]
    hs = [h for h in hams for _ in range(2)]
    psis = [O.random_state(L, 800 + k) for k in range(2 * pairs)]
    out = {}
    for streams in (1, 2):
        prop = spx.Propagator(ctx, L, 1e-12)
        prop.set_streams(streams)
        states = [spx.StateVector(ctx, L, p) for p in psis]
        prop.evolve_set(hs, spx.REAL, states, 0.05, 3)
        out[streams] = [st.download() for st in states]
    for k in range(2 * pairs):
        assert np.array_equal(out[1][k], out[2][k]), k


def test_same_hamiltonian_twice_and_argument_errors(ctx):
    L = 14
    fields, c = realisation(L, 9)
    h = device_ham(ctx, L, fields, c, 2)
    prop = spx.Propagator(ctx, L, 1e-12)
    p0, p1 = O.random_state(L, 1), O.random_state(L, 2)
    a, b = spx.StateVector(ctx, L, p0), spx.StateVector(ctx, L, p1)
    prop.evolve_set([h, h], spx.REAL, [a, b], 0.05, 2)
    for p, s in ((p0, a), (p1, b)):
        one = spx.StateVector(ctx, L, p)
        prop.evolve(h, spx.REAL, one, 0.05, 2)
        assert np.array_equal(one.download(), s.download())
    with pytest.raises(spx.InvalidArgument):
        prop.evolve_set([h], spx.REAL, [a, b], 0.05, 1)
    with pytest.raises(spx.InvalidArgument):
        prop.evolve_set([h, h], spx.REAL, [a, a], 0.05, 1)
    other = spx.StateVector(ctx, L + 1, O.random_state(L + 1, 3))
    with pytest.raises(spx.InvalidArgument):
        prop.evolve_set([h], spx.REAL, [other], 0.05, 1)
    # a state of another context: its stream would not wait for the set's launches
    ctx2 = spx.Context(0)
    foreign = spx.StateVector(ctx2, L, O.random_state(L, 4))
    with pytest.raises(spx.InvalidArgument, match="another context"):
        prop.evolve_set([h], spx.REAL, [foreign], 0.05, 1)
    # single-state evolves and prebuilt steps check the same, for the state
    # and for the Hamiltonian
    with pytest.raises(spx.InvalidArgument, match="another context"):
        prop.evolve(h, spx.REAL, foreign, 0.05, 1)
    coeffs, eff = spx.cheb_build(spx.REAL, h.spectral_bound(), 0.05, 1e-12)
    with pytest.raises(spx.InvalidArgument, match="another context"):
        prop.steps(h, spx.REAL, foreign, coeffs, eff, 1)
    h2 = device_ham(ctx2, L, fields, c, 2)
    with pytest.raises(spx.InvalidArgument, match="another context"):
        prop.evolve(h2, spx.REAL, a, 0.05, 1)
    with pytest.raises(spx.InvalidArgument, match="another context"):
        prop.evolve_many(h2, spx.REAL, [a, b], 0.05, 1)
    # the rejected calls left the states alone
    for p, s in ((p0, a), (p1, b)):
        one = spx.StateVector(ctx, L, p)
        prop.evolve(h, spx.REAL, one, 0.05, 2)
        assert np.array_equal(one.download(), s.download())


def test_steps_set_over_prebuilt_expansions_matches_evolve_set(ctx):
    L = 16
    banner("steps_set vs evolve_set L=16")
    hams = [device_ham(ctx, L, *realisation(L, i, s), 2) for i, s in ((0, 1.0), (1, 2.0), (2, 1.0))]
    prop = spx.Propagator(ctx, L, 1e-12)
    psis = [O.random_state(L, 400 + k) for k in range(3)]
    a = [spx.StateVector(ctx, L, p) for p in psis]
    prop.evolve_set(hams, spx.REAL, a, 0.05, 2)
    built = [spx.cheb_build(spx.REAL, h.spectral_bound(), 0.05, 1e-12) for h in hams]
    b = [spx.StateVector(ctx, L, p) for p in psis]
    prop.steps_set(hams, spx.REAL, b, [c for c, _ in built], [e for _, e in built], 2)
    for x, y in zip(a, b):
        assert np.array_equal(x.download(), y.download())


def test_batched_autotune_reports_the_hadamard_path(ctx):
    L = 16
    banner("autotune at width 4, L=16")
    h = device_ham(ctx, L, *realisation(L, 0), -1)
    report = h.autotune(h.spectral_bound(), nvec=4)
    banner(report[:300])
    assert report.startswith("chosen ") and "hadamard: " in report
    with pytest.raises(spx.InvalidArgument):
        h.autotune(h.spectral_bound(), nvec=17)


def test_default_path_follows_size_and_launch_width(ctx):
    """Without a schedule: all-pairs gathers for one state per launch at
    L=18, the Hadamard path when states share launches, and back."""
    L = 18
    banner("default flip-flop path by launch width, L=18")
    fields, c = realisation(L, 0)
    h = spx.Hamiltonian(ctx, L)
    h.set_fields(fields)
    h.set_couplings(c.lo, c.hi, c.jx, c.jy, c.jz)
    prop = spx.Propagator(ctx, L, 1e-12)
    s = spx.StateVector(ctx, L, O.random_state(L, 1))
    t = spx.StateVector(ctx, L, O.random_state(L, 2))
    prop.evolve(h, spx.REAL, s, 0.05, 1)
    assert h.kernel_name().startswith("k_dense"), h.kernel_name()
    prop.evolve_many(h, spx.REAL, [s, t], 0.05, 1)
    assert h.kernel_name().startswith("k_wht_"), h.kernel_name()
    prop.evolve(h, spx.REAL, s, 0.05, 1)
    assert h.kernel_name().startswith("k_dense"), h.kernel_name()


@pytest.mark.parametrize("L,count,transverse", [(16, 8, False), (17, 5, False), (16, 8, True)])
def test_two_stream_guard_without_a_schedule(ctx, L, count, transverse):
    """Sets without a tuned schedule: flip-flop sets at L = 16 (8 states)
    and L = 17 (5 states) take the Hadamard path at their one-stream width,
    whose halves would fall to the gathers, so they keep one stream; a set
    with transverse fields stays on the Hadamard path at every width and
    splits. Either way every state is on the Hadamard path and bitwise equal
    to the same set on one stream."""
    banner(f"two-stream guard L={L} x{count} transverse={transverse}")
    specs = [realisation(L, i, 1.0, transverse) for i in range(count)]
    psis = [O.random_state(L, 700 + k) for k in range(count)]
    out = []
    for streams in (2, 1):
        hams = [device_ham(ctx, L, f, c, -1) for f, c in specs]
        prop = spx.Propagator(ctx, L, 1e-12)
        prop.set_streams(streams)
        states = [spx.StateVector(ctx, L, p) for p in psis]
        prop.evolve_set(hams, spx.REAL, states, 0.05, 3)
        names = [h.kernel_name() for h in hams]
        assert all(n.startswith("k_wht_") for n in names), names
        out.append([s.download() for s in states])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
