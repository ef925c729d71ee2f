"""ctypes view of the CPU oracle (oracle/lib/liboracle.so) and, where it was
built, of the reference's own compiled kernels (oracle/_ref/libspinsim_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker or the CPU baseline, never
as the path being measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libspinsim_ref.so")

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_szp = C.POINTER(C.c_size_t)
_ip = C.POINTER(C.c_int)

OBS_CORR_X, OBS_CORR_Y, OBS_CORR_Z = 0, 1, 2
OBS_MAG_X, OBS_MAG_Y, OBS_MAG_Z = 3, 4, 5
OBS_NORM, OBS_ENERGY = 6, 7


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _u32(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def build():
    """Compile the oracle (and oracle/_ref when the reference is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/src/core"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.so_splitmix.restype = C.c_uint64
        L.so_splitmix.argtypes = [C.c_uint64]
        L.so_mt_raw.argtypes = [C.c_uint64, C.c_size_t, _u64p]
        L.so_rng_draws.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, C.c_int, _dp]
        for fn in (L.so_bessel_j, L.so_bessel_i):
            fn.restype = C.c_int
            fn.argtypes = [C.c_double, C.c_size_t, _dp, _szp]
        L.so_cheb_build.restype = C.c_int
        L.so_cheb_build.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_size_t,
                                    _dp, _szp, _dp]
        L.so_state_norm.restype = C.c_double
        L.so_state_norm.argtypes = [_dp, C.c_size_t]
        L.so_state_inner.argtypes = [_dp, _dp, C.c_size_t, _dp, _dp]
        L.so_state_add_scaled.argtypes = [_dp, C.c_size_t, C.c_double, C.c_double, _dp]
        L.so_axpy_sigma.argtypes = [C.c_int, C.c_double, _dp, C.c_size_t, C.c_size_t, _dp]
        L.so_apply_sigma.argtypes = [C.c_int, _dp, C.c_size_t, C.c_size_t, _dp]
        L.so_axpy_coupling.argtypes = [C.c_double, C.c_double, C.c_double, _dp, C.c_size_t,
                                       C.c_size_t, C.c_size_t, _dp]
        L.so_axpy_diagonal.argtypes = [C.c_double, _dp, _dp, C.c_size_t, _dp]
        L.so_axpy_transverse.argtypes = [C.c_double, C.c_double, _dp, C.c_size_t, C.c_size_t, _dp]
        L.so_axpy_flip_flop.argtypes = [C.c_double, _dp, C.c_size_t, C.c_size_t, C.c_size_t, _dp]
        L.so_axpy_flip_both.argtypes = [C.c_double, _dp, C.c_size_t, C.c_size_t, C.c_size_t, _dp]
        L.so_ham_new.restype = C.c_void_p
        L.so_ham_new.argtypes = [C.c_size_t]
        L.so_ham_free.argtypes = [C.c_void_p]
        L.so_ham_set_field.restype = C.c_int
        L.so_ham_set_field.argtypes = [C.c_void_p, C.c_size_t, C.c_double, C.c_double, C.c_double]
        L.so_ham_set_coupling.restype = C.c_int
        L.so_ham_set_coupling.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_double,
                                          C.c_double, C.c_double]
        L.so_ham_bound.restype = C.c_double
        L.so_ham_bound.argtypes = [C.c_void_p]
        L.so_ham_diag.restype = _dp
        L.so_ham_diag.argtypes = [C.c_void_p]
        L.so_ham_accumulate.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.so_ham_apply_physical.argtypes = [C.c_void_p, _dp, _dp]
        L.so_ham_apply_twice_normalized.restype = C.c_int
        L.so_ham_apply_twice_normalized.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.so_ham_accumulate_twice_normalized.restype = C.c_int
        L.so_ham_accumulate_twice_normalized.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.so_evolve.restype = C.c_int
        L.so_evolve.argtypes = [C.c_void_p, C.c_int, _dp, C.c_double, C.c_size_t, C.c_int,
                                C.c_double, _szp]
        L.so_rotate.restype = C.c_int
        L.so_rotate.argtypes = [_dp, C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_double,
                                _u32p, C.c_size_t]
        L.so_two_spin_gate.restype = C.c_int
        L.so_two_spin_gate.argtypes = [_dp, C.c_size_t, _dp, C.c_size_t, C.c_size_t]
        L.so_apply_total_spin.argtypes = [C.c_int, _dp, C.c_size_t, _dp]
        L.so_corr.argtypes = [C.c_int, _dp, _dp, C.c_size_t, _dp, _dp, _dp]
        L.so_mag.restype = C.c_double
        L.so_mag.argtypes = [C.c_int, C.c_size_t, _dp, C.c_size_t, _dp]
        L.so_energy.restype = C.c_double
        L.so_energy.argtypes = [C.c_void_p, _dp, _dp]
        L.so_dipolar_lattice.restype = C.c_int
        L.so_dipolar_lattice.argtypes = [C.c_size_t, C.c_void_p, C.c_double, C.c_double, C.c_int,
                                         _u32p, _u32p, _dp, _dp, _dp]
        L.so_rng_for_realization.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.so_rng_seed.argtypes = [C.c_void_p, C.c_uint64]
        L.so_rng_uniform.restype = C.c_double
        L.so_rng_uniform.argtypes = [C.c_void_p]
        L.so_rng_gaussian.restype = C.c_double
        L.so_rng_gaussian.argtypes = [C.c_void_p]
        L.so_prepare_initial_state.restype = C.c_int
        L.so_prepare_initial_state.argtypes = [_dp, C.c_size_t, C.c_int, C.c_void_p, C.c_size_t]
        L.so_prepare_reference.argtypes = [_dp, C.c_size_t, C.c_int, _dp]
        L.so_field_disorder.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
        L.so_program_new.restype = C.c_void_p
        L.so_program_free.argtypes = [C.c_void_p]
        L.so_program_evolve.argtypes = [C.c_void_p, C.c_double, C.c_size_t, C.c_int, C.c_int,
                                        _dp, C.c_size_t]
        L.so_program_pulse.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                       _u32p, C.c_size_t]
        L.so_program_measure.argtypes = [C.c_void_p]
        L.so_program_repeat.argtypes = [C.c_void_p, C.c_size_t]
        L.so_program_end.argtypes = [C.c_void_p]
        L.so_program_measurements.restype = C.c_size_t
        L.so_program_measurements.argtypes = [C.c_void_p]
        L.so_config_width.restype = C.c_size_t
        L.so_config_width.argtypes = [C.c_void_p]
        L.so_run_realization.restype = C.c_int
        L.so_run_realization.argtypes = [C.c_void_p, C.c_uint64, _dp, _dp]
        L.so_run_ensemble.restype = C.c_int
        L.so_run_ensemble.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, _dp, _dp, _dp, _dp, _dp]
        L.so_acc_new.restype = C.c_void_p
        L.so_acc_new.argtypes = [C.c_size_t]
        L.so_acc_free.argtypes = [C.c_void_p]
        L.so_acc_add.argtypes = [C.c_void_p, _dp]
        L.so_acc_merge.argtypes = [C.c_void_p, C.c_void_p]
        L.so_acc_stats.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.so_time_orders.restype = C.c_double
        L.so_time_orders.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t]
        _lib = L
    return _lib


_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own kernels, compiled from its sources (or None)."""
    global _ref
    if _ref is None and ref_available():
        R = C.CDLL(REF_PATH)
        R.ref_rng_draws.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, C.c_int, _dp]
        R.ref_rng_seeded.argtypes = [C.c_uint64, C.c_size_t, C.c_int, _dp]
        R.ref_bessel.restype = C.c_int
        R.ref_bessel.argtypes = [C.c_int, C.c_double, C.c_size_t, _dp, _szp]
        R.ref_ham_apply.restype = C.c_int
        R.ref_ham_apply.argtypes = [C.c_int, C.c_size_t, _dp, C.c_size_t, _u32p, _u32p, _dp, _dp,
                                    _dp, C.c_double, _dp, _dp]
        R.ref_ham_bound.restype = C.c_double
        R.ref_ham_bound.argtypes = [C.c_size_t, _dp, C.c_size_t, _u32p, _u32p, _dp, _dp, _dp]
        R.ref_kernel.argtypes = [C.c_int, C.c_size_t, C.c_double, C.c_double, C.c_double,
                                 C.c_size_t, C.c_size_t, _dp, _dp, _dp]
        R.ref_state_ops.argtypes = [C.c_size_t, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        R.ref_one_step.restype = C.c_int
        R.ref_one_step.argtypes = [C.c_int, C.c_size_t, _dp, C.c_size_t, _u32p, _u32p, _dp, _dp, _dp,
                                   _dp, C.c_size_t, C.c_double, C.c_size_t, _dp]
        R.ref_time_orders.restype = C.c_double
        R.ref_time_orders.argtypes = [C.c_size_t, _dp, C.c_size_t, _u32p, _u32p, _dp, _dp, _dp,
                                      C.c_size_t, C.c_size_t]
        _ref = R
    return _ref


# --------------------------------------------------------------------------- helpers

class RngState(C.Structure):
    _fields_ = [("mt", C.c_uint64 * 312), ("mti", C.c_size_t), ("spare", C.c_double),
                ("has_spare", C.c_int)]


def rng_for_realization(seed: int, index: int) -> RngState:
    r = RngState()
    lib().so_rng_for_realization(C.byref(r), seed, index)
    return r


def draws(seed: int, index: int, n: int, gaussian: bool) -> np.ndarray:
    out = np.zeros(n)
    lib().so_rng_draws(seed, index, n, int(gaussian), _d(out))
    return out


def bessel(kind: str, x: float, count: int):
    out = np.zeros(max(count, 1))
    nz = C.c_size_t(0)
    fn = lib().so_bessel_j if kind == "J" else lib().so_bessel_i
    rc = fn(x, count, _d(out), C.byref(nz))
    return rc, out[:count], nz.value


def cheb_build(mode: int, bound: float, tau: float, eps: float, max_order: int = 6000):
    coeffs = np.zeros(max_order + 4)
    order = C.c_size_t(0)
    eff = C.c_double(0)
    rc = lib().so_cheb_build(mode, bound, tau, eps, max_order, _d(coeffs), C.byref(order), C.byref(eff))
    return rc, coeffs[: order.value + 1].copy(), eff.value


@dataclass
class Couplings:
    lo: np.ndarray
    hi: np.ndarray
    jx: np.ndarray
    jy: np.ndarray
    jz: np.ndarray

    @property
    def n(self) -> int:
        return int(self.lo.shape[0])


def dipolar_lattice(L: int, seed: int, index: int, scale: float = 1.0, min_sep: float = 0.05,
                    angular: bool = False, disorder: float = 0.0):
    """Geometry (and optional Hz disorder) drawn in runRealization order."""
    r = rng_for_realization(seed, index)
    n = L * (L - 1) // 2
    c = Couplings(np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n), np.zeros(n), np.zeros(n))
    lib().so_dipolar_lattice(L, C.byref(r), scale, min_sep, int(angular), _u32(c.lo), _u32(c.hi),
                             _d(c.jx), _d(c.jy), _d(c.jz))
    hz = np.zeros(L)
    if disorder != 0.0:
        for s in range(L):
            hz[s] = hz[s] + disorder * lib().so_rng_gaussian(C.byref(r))
    return c, hz, r


class Ham:
    """Owning handle to an oracle Hamiltonian."""

    def __init__(self, L: int, fields: Optional[np.ndarray] = None, couplings: Optional[Couplings] = None):
        self.L = L
        self.h = lib().so_ham_new(L)
        if fields is not None:
            f = np.asarray(fields, dtype=float).reshape(L, 3)
            for s in range(L):
                lib().so_ham_set_field(self.h, s, f[s, 0], f[s, 1], f[s, 2])
        if couplings is not None:
            for p in range(couplings.n):
                lib().so_ham_set_coupling(self.h, int(couplings.lo[p]), int(couplings.hi[p]),
                                          couplings.jx[p], couplings.jy[p], couplings.jz[p])

    def __del__(self):
        try:
            lib().so_ham_free(self.h)
        except Exception:
            pass

    def bound(self) -> float:
        return lib().so_ham_bound(self.h)

    def diag(self) -> np.ndarray:
        p = lib().so_ham_diag(self.h)
        return np.ctypeslib.as_array(p, shape=(1 << self.L,)).copy()

    def apply_physical(self, psi: np.ndarray) -> np.ndarray:
        out = np.zeros_like(psi)
        lib().so_ham_apply_physical(self.h, _d(psi), _d(out))
        return out

    def accumulate(self, scale: float, psi: np.ndarray, out: np.ndarray) -> None:
        lib().so_ham_accumulate(self.h, scale, _d(psi), _d(out))

    def evolve(self, psi: np.ndarray, tau: float, nsteps: int = 1, mode: int = 0,
               autonorm: bool = False, eps: float = 1e-7):
        order = C.c_size_t(0)
        rc = lib().so_evolve(self.h, mode, _d(psi), tau, nsteps, int(autonorm), eps, C.byref(order))
        return rc, order.value


def random_state(L: int, seed: int) -> np.ndarray:
    """Deterministic normalised test state (numpy generator, not the reference stream)."""
    g = np.random.default_rng(seed)
    v = g.standard_normal(2 << L)
    return v / np.linalg.norm(v)


# ------------------------------------------------------------------- programs/configs

class Program:
    def __init__(self):
        self.p = lib().so_program_new()

    def __del__(self):
        try:
            lib().so_program_free(self.p)
        except Exception:
            pass

    def evolve(self, tau, nsteps=1, mode=0, autonorm=False, fields=None, L=0):
        f = None if fields is None else np.ascontiguousarray(fields, dtype=float).reshape(-1)
        lib().so_program_evolve(self.p, tau, nsteps, mode, int(autonorm), None if f is None else _d(f),
                                0 if f is None else f.size // 3)
        return self

    def pulse(self, axis, angle, targets):
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        lib().so_program_pulse(self.p, axis[0], axis[1], axis[2], angle, _u32(t), t.size)
        return self

    def measure(self):
        lib().so_program_measure(self.p)
        return self

    def repeat(self, count):
        lib().so_program_repeat(self.p, count)
        return self

    def end(self):
        lib().so_program_end(self.p)
        return self

    def measurements(self) -> int:
        return lib().so_program_measurements(self.p)


class _Config(C.Structure):
    _fields_ = [("L", C.c_size_t), ("base_fields", _dp), ("initial_state", C.c_int),
                ("basis_index", C.c_size_t), ("seed", C.c_uint64), ("generate_geometry", C.c_int),
                ("dipolar_scale", C.c_double), ("min_separation", C.c_double), ("angular", C.c_int),
                ("field_disorder", C.c_double), ("reference_axis", C.c_int), ("epsilon", C.c_double),
                ("obs_codes", _ip), ("obs_spins", _ip), ("nobs", C.c_size_t), ("program", C.c_void_p)]


@dataclass
class Config:
    L: int
    program: Program
    observables: Sequence[tuple] = ((OBS_CORR_X, 0), (OBS_NORM, 0))
    seed: int = 20240726
    base_fields: Optional[np.ndarray] = None
    initial_state: int = 0
    basis_index: int = 0
    generate_geometry: bool = True
    dipolar_scale: float = 1.0
    min_separation: float = 0.05
    angular: bool = False
    field_disorder: float = 0.0
    reference_axis: int = 0
    epsilon: float = 1e-7
    _keep: list = field(default_factory=list)

    def cstruct(self) -> _Config:
        codes = np.array([o[0] for o in self.observables], dtype=np.int32)
        spins = np.array([o[1] for o in self.observables], dtype=np.int32)
        bf = None if self.base_fields is None else np.ascontiguousarray(self.base_fields, dtype=float)
        self._keep = [codes, spins, bf]
        return _Config(self.L, None if bf is None else _d(bf), self.initial_state, self.basis_index,
                       self.seed, int(self.generate_geometry), self.dipolar_scale, self.min_separation,
                       int(self.angular), self.field_disorder, self.reference_axis, self.epsilon,
                       codes.ctypes.data_as(_ip), spins.ctypes.data_as(_ip), len(codes), self.program.p)


def run_realization(cfg: Config, index: int):
    cs = cfg.cstruct()
    points = cfg.program.measurements()
    width = lib().so_config_width(C.byref(cs))
    times = np.zeros(max(points, 1))
    values = np.zeros(max(points * width, 1))
    rc = lib().so_run_realization(C.byref(cs), index, _d(times), _d(values))
    if rc != 0:
        raise RuntimeError(f"oracle realisation {index} failed with code {rc}")
    return times[:points], values[: points * width].reshape(points, width)


def run_ensemble(cfg: Config, n: int, threads: int = 1):
    cs = cfg.cstruct()
    points = cfg.program.measurements()
    width = lib().so_config_width(C.byref(cs))
    times = np.zeros(max(points, 1))
    per = np.zeros(max(n * points * width, 1))
    mean = np.zeros(max(points * width, 1))
    sd = np.zeros_like(mean)
    se = np.zeros_like(mean)
    rc = lib().so_run_ensemble(C.byref(cs), n, threads, _d(times), _d(per), _d(mean), _d(sd), _d(se))
    if rc != 0:
        raise RuntimeError(f"oracle ensemble failed with code {rc}")
    sh = (points, width)
    return (times[:points], per[: n * points * width].reshape(n, points, width),
            mean[: points * width].reshape(sh), sd[: points * width].reshape(sh),
            se[: points * width].reshape(sh))


_std = None


def ref_test_state(L: int, seed: int) -> np.ndarray:
    """The reference test-suite's randomState(L, seed) (tests/DenseReference.cpp:204-214),
    reproduced with libstdc++'s normal_distribution over mt19937_64(seed)."""
    global _std
    if _std is None:
        path = os.path.join(HERE, "lib", "libstdinputs.so")
        if not os.path.exists(path):
            build()
        _std = C.CDLL(path)
        _std.std_random_state.argtypes = [C.c_size_t, C.c_uint64, _dp]
    out = np.zeros(2 << L)
    _std.std_random_state(L, seed, _d(out))
    return out


# ------------------------------------------------------- experiment programs
# The pulse programs of the ensemble workloads, restated for the oracle (the
# device path builds its own in spinsim_amd.ensembles and bench.py).

def cpmg_program(L: int, points: int, steps: int = 10, dt: float = 0.05) -> Program:
    """pi/2_x on every spin, then points x [steps x dt, ideal pi_y, steps x dt, measure]."""
    every = list(range(L))
    p = Program().pulse((1, 0, 0), math.pi / 2, every).repeat(points)
    return p.evolve(dt, steps).pulse((0, 1, 0), math.pi, every).evolve(dt, steps).measure().end()


def pulse_fields(L: int, axis: str, field: float) -> np.ndarray:
    f = np.zeros((L, 3))
    f[:, 0 if axis == "x" else 1] = field
    return f


XY8_AXES = "xyxyyxyx"


def xy8_program(L: int, blocks: int = 16, omega: float = 50.0, free_steps: int = 10) -> Program:
    """A measure, then blocks x [XY-8 with finite pulses (pi / omega under
    hx or hy = omega, hz 0), each after free_steps x 0.05, measure]."""
    p = Program().measure().repeat(blocks)
    for axis in XY8_AXES:
        p.evolve(0.05, free_steps).evolve(math.pi / omega, 1, fields=pulse_fields(L, axis, omega))
    return p.measure().end()


def xy16_program(L: int, omega: float = 50.0, free: float = 0.05, per_measure: int = 4) -> Program:
    """XY-16 (XY-8 and its phase-inverted copy), one free step of `free`
    before each finite pulse, a measure after every `per_measure` pulses."""
    seq = [(a, 1.0) for a in XY8_AXES] + [(a, -1.0) for a in XY8_AXES]
    p = Program().repeat(1)
    for k, (axis, sign) in enumerate(seq):
        p.evolve(free, 1).evolve(math.pi / omega, 1, fields=pulse_fields(L, axis, sign * omega))
        if (k + 1) % per_measure == 0:
            p.measure()
    return p.end()


def dipolar_config(L: int, program: Program, observables=((OBS_CORR_X, 0), (OBS_NORM, 0)),
                   seed: int = 20240726) -> "Config":
    """runRealization for the dipolar ensemble of the workloads: geometry at
    min spacing 0.5 with the secular angle, Sz disorder 0.5, a random initial
    state, reference axis x, epsilon 1e-7."""
    return Config(L=L, program=program, observables=observables, seed=seed, initial_state=0,
                  generate_geometry=True, dipolar_scale=1.0, min_separation=0.5, angular=True,
                  field_disorder=0.5, reference_axis=0, epsilon=1e-7)
