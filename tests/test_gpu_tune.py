"""The autotuner: it times the eligible order-kernel schedules, keeps the
fastest, caches it per problem shape, and every schedule gives the oracle's
result to rounding."""
import numpy as np
import pytest

from oracle import oracle as O
from spinsim_amd import spx

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def system(L, index):
    c, hz, _ = O.dipolar_lattice(L, 20240726, index, min_sep=0.5, angular=True, disorder=0.5)
    fields = np.zeros((L, 3))
    fields[:, 2] = hz
    return fields, c


def device_ham(ctx, L, fields, c):
    h = spx.Hamiltonian(ctx, L)
    h.set_fields(fields)
    h.set_couplings(c.lo, c.hi, c.jx, c.jy, c.jz)
    return h


@pytest.mark.parametrize("L", [16, 12])
def test_every_schedule_matches_the_oracle_and_the_tuner_keeps_one(ctx, L):
    banner(f"autotune L={L}")
    fields, c = system(L, 5)
    ho = O.Ham(L, fields, c)
    psi0 = O.random_state(L, 77)
    want = psi0.copy()
    rc, order = ho.evolve(want, 0.05, 2, 0, False, 1e-12)
    assert rc == 0
    scale = max(1.0, float(np.max(np.abs(want))))
    prop = spx.Propagator(ctx, L, 1e-12)
    schedules = [dict(dense=0, tile_bits=10), dict(dense=0, tile_bits=9, xcd=1), dict(dense=0, tile_bits=8),
                 dict(dense=0, tile_bits=10, block=1024)]
    if L >= 14:
        schedules += [dict(dense=1, dense_tile_bits=10), dict(dense=1, dense_tile_bits=9, xcd=1), dict(dense=2)]
    names = set()
    for sch in schedules:
        h = device_ham(ctx, L, fields, c)
        h.set_schedule(**sch)
        s = spx.StateVector(ctx, L, psi0)
        prop.evolve(h, spx.REAL, s, 0.05, 2)
        names.add(h.kernel_name())
        err = float(np.max(np.abs(s.download() - want)))
        assert err < 1e-12 * scale, (sch, err)
    assert len(names) >= 3, names
    h = device_ham(ctx, L, fields, c)
    report = h.autotune(h.spectral_bound() * 1.0)
    banner(report[:400])
    assert report.startswith("chosen ") and report.count(" us;") >= 4
    s = spx.StateVector(ctx, L, psi0)
    prop.evolve(h, spx.REAL, s, 0.05, 2)
    assert float(np.max(np.abs(s.download() - want))) < 1e-12 * scale
    chosen = h.kernel_name()
    # mode 1: another realisation of the same shape adopts the cached schedule
    spx.autotune_mode(1)
    try:
        f2, c2 = system(L, 6)
        h2 = device_ham(ctx, L, f2, c2)
        s2 = spx.StateVector(ctx, L, O.random_state(L, 78))
        prop.evolve(h2, spx.REAL, s2, 0.05, 1)
        assert h2.kernel_name() == chosen
    finally:
        spx.autotune_mode(0)
    with pytest.raises(spx.InvalidArgument):
        spx.autotune_mode(3)
    with pytest.raises(spx.InvalidArgument):
        h.set_schedule(tile_bits=3)
