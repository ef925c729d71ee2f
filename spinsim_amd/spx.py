"""ctypes binding of libspx (include/spx.h), mirroring the reference's
spinsim classes closely enough that tests read like the reference's own.

The compiled library is required: importing works without it, but every
object raises SpxMissing on construction if spinsim_amd/lib/libspx.so was not
built (there is no CPU fallback on this path).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Iterable, Optional, Sequence

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libspx.so")

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_szp = C.POINTER(C.c_size_t)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_fp = C.POINTER(C.c_float)

REAL, IMAGINARY = 0, 1
# observable codes of the batched runner (include/spx.h)
OBS_CORR_X, OBS_CORR_Y, OBS_CORR_Z = 0, 1, 2
OBS_MAG_X, OBS_MAG_Y, OBS_MAG_Z = 3, 4, 5
OBS_NORM, OBS_ENERGY = 6, 7
AXES = {"x": 0, "y": 1, "z": 2}

OK, EINVAL, ERANGE, EDOMAIN, ECONVERGE, EOVERFLOW, ERUNTIME, EDEVICE, ENOMEM = range(9)


class SpxMissing(RuntimeError):
    pass


class SpxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class InvalidArgument(SpxError, ValueError):
    pass


class OutOfRange(SpxError, IndexError):
    pass


class DomainError(SpxError, ValueError):
    pass


class ConvergenceError(SpxError):
    def __init__(self, code, msg, bound, tau, eps, order):
        super().__init__(code, msg)
        self.spectral_bound, self.tau, self.epsilon, self.attempted_order = bound, tau, eps, order


class OverflowErr(SpxError, OverflowError):
    pass


# The exported surface, name -> (restype, argtypes); also the list the CPU
# test checks against include/spx.h.
SIGNATURES = {
    "spx_last_error": (C.c_char_p, []),
    "spx_last_convergence": (C.c_int, [_dp, _dp, _dp, _szp]),
    "spx_device_count": (C.c_int, [_ip]),
    "spx_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "spx_ctx_destroy": (C.c_int, [_vp]),
    "spx_ctx_synchronize": (C.c_int, [_vp]),
    "spx_ctx_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "spx_ctx_event_record": (C.c_int, [_vp, C.c_int]),
    "spx_ctx_event_elapsed": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "spx_state_create": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "spx_state_destroy": (C.c_int, [_vp]),
    "spx_state_nspins": (C.c_int, [_vp, _szp]),
    "spx_state_upload": (C.c_int, [_vp, _dp]),
    "spx_state_download": (C.c_int, [_vp, _dp]),
    "spx_state_copy": (C.c_int, [_vp, _vp]),
    "spx_state_set_zero": (C.c_int, [_vp]),
    "spx_state_set_basis": (C.c_int, [_vp, C.c_size_t]),
    "spx_state_scale": (C.c_int, [_vp, C.c_double]),
    "spx_state_negate": (C.c_int, [_vp]),
    "spx_state_add_scaled": (C.c_int, [_vp, C.c_double, C.c_double, _vp]),
    "spx_state_norm": (C.c_int, [_vp, _dp]),
    "spx_state_normalize": (C.c_int, [_vp, _dp]),
    "spx_state_inner": (C.c_int, [_vp, _vp, _dp, _dp]),
    "spx_bessel": (C.c_int, [C.c_int, C.c_double, C.c_size_t, _dp, _szp]),
    "spx_cheb_build": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_size_t, _dp,
                                 C.c_size_t, _szp, _dp]),
    "spx_ham_create": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "spx_ham_destroy": (C.c_int, [_vp]),
    "spx_ham_set_field": (C.c_int, [_vp, C.c_size_t, C.c_double, C.c_double, C.c_double]),
    "spx_ham_get_field": (C.c_int, [_vp, C.c_size_t, _dp]),
    "spx_ham_set_coupling": (C.c_int, [_vp, C.c_size_t, C.c_size_t, C.c_double, C.c_double,
                                       C.c_double]),
    "spx_ham_num_couplings": (C.c_int, [_vp, _szp]),
    "spx_ham_get_coupling": (C.c_int, [_vp, C.c_size_t, _u32p, _u32p, _dp]),
    "spx_ham_spectral_bound": (C.c_int, [_vp, _dp]),
    "spx_ham_apply": (C.c_int, [_vp, C.c_int, C.c_double, _vp, _vp]),
    "spx_ham_diagonal": (C.c_int, [_vp, _dp]),
    "spx_ham_schedule": (C.c_int, [_vp, C.c_char_p, C.c_size_t]),
    "spx_ham_autotune": (C.c_int, [_vp, C.c_double, C.c_char_p, C.c_size_t]),
    "spx_autotune_mode": (C.c_int, [C.c_int]),
    "spx_ham_set_schedule": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "spx_prop_create": (C.c_int, [_vp, C.c_size_t, C.c_double, C.POINTER(_vp)]),
    "spx_prop_destroy": (C.c_int, [_vp]),
    "spx_prop_evolve": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_double, C.c_size_t, C.c_int]),
    "spx_prop_evolve_many": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(_vp), C.c_size_t, C.c_double,
                                       C.c_size_t, C.c_int]),
    "spx_prop_evolve_set": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.POINTER(_vp), C.c_size_t, C.c_double,
                                      C.c_size_t, C.c_int]),
    "spx_prop_steps_set": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.POINTER(_vp), C.c_size_t, C.POINTER(_dp),
                                     _szp, _dp, C.c_size_t, C.c_int]),
    "spx_ham_autotune_batch": (C.c_int, [_vp, C.c_double, C.c_int, C.c_char_p, C.c_size_t]),
    "spx_prop_last_order": (C.c_int, [_vp, _szp]),
    "spx_prop_set_graphs": (C.c_int, [_vp, C.c_int]),
    "spx_prop_set_streams": (C.c_int, [_vp, C.c_int]),
    "spx_prop_epsilon": (C.c_int, [_vp, _dp]),
    "spx_prop_steps": (C.c_int, [_vp, _vp, C.c_int, _vp, _dp, C.c_size_t, C.c_double, C.c_size_t,
                                 C.c_int]),
    "spx_rotate": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, _u32p,
                             C.c_size_t]),
    "spx_two_spin_gate": (C.c_int, [_vp, _dp, C.c_size_t, C.c_size_t]),
    "spx_apply_sigma": (C.c_int, [C.c_int, C.c_double, _vp, C.c_size_t, _vp, C.c_int]),
    "spx_apply_total_spin": (C.c_int, [C.c_int, _vp, _vp]),
    "spx_measure_corr": (C.c_int, [C.c_int, _vp, _vp, _dp, _dp]),
    "spx_measure_mag": (C.c_int, [C.c_int, C.c_size_t, _vp, _dp]),
    "spx_measure_energy": (C.c_int, [_vp, _vp, _dp]),
    "spx_program_create": (C.c_int, [C.POINTER(_vp)]),
    "spx_program_destroy": (C.c_int, [_vp]),
    "spx_program_evolve": (C.c_int, [_vp, C.c_double, C.c_size_t, C.c_int, C.c_int, _dp,
                                     C.c_size_t]),
    "spx_program_evolve_uniform": (C.c_int, [_vp, C.c_double, C.c_size_t, C.c_int, C.c_int,
                                             C.c_double, C.c_double, C.c_double]),
    "spx_program_pulse": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, _u32p,
                                    C.c_size_t]),
    "spx_program_measure": (C.c_int, [_vp]),
    "spx_program_repeat": (C.c_int, [_vp, C.c_size_t]),
    "spx_program_end": (C.c_int, [_vp]),
    "spx_program_measurements": (C.c_int, [_vp, _szp]),
    "spx_batch_create": (C.c_int, [_vp, _vp, C.c_double, C.c_int, _ip, _ip, C.c_size_t,
                                   C.POINTER(_vp)]),
    "spx_batch_destroy": (C.c_int, [_vp]),
    "spx_batch_add": (C.c_int, [_vp, C.c_size_t, _dp, C.c_size_t, _u32p, _u32p, _dp, _dp, _dp,
                                _dp]),
    "spx_batch_add_dipolar": (C.c_int, [_vp, C.c_uint64, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double,
                                        C.c_double, C.c_int, C.c_double]),
    "spx_batch_run": (C.c_int, [_vp, _dp, _dp]),
    "spx_prop_evolve_noalloc": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_size_t, C.c_double, _szp]),
    "spx_batch_width": (C.c_int, [_vp, _szp]),
    "spx_batch_clear": (C.c_int, [_vp]),
    "spx_gnn_create": (C.c_int, [_vp, C.c_size_t, C.c_uint64, C.c_double, C.c_double, C.POINTER(_vp)]),
    "spx_gnn_destroy": (C.c_int, [_vp]),
    "spx_gnn_num_params": (C.c_int, [_vp, _szp]),
    "spx_gnn_get_params": (C.c_int, [_vp, _fp]),
    "spx_gnn_set_params": (C.c_int, [_vp, _fp]),
    "spx_gnn_predict": (C.c_int, [_vp, C.c_size_t, C.c_size_t, _dp, _dp, _fp]),
    "spx_gnn_set_training_set": (C.c_int, [_vp, C.c_size_t, C.c_size_t, _dp, _dp, _fp]),
    "spx_gnn_train_epoch": (C.c_int, [_vp, C.c_size_t, _u32p, C.c_float, _fp]),
    "spx_gnn_predict_dipolar": (C.c_int, [_vp, C.c_uint64, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double,
                                          C.c_double, C.c_int, C.c_double, _fp]),
    "spx_dipolar_realisation": (C.c_int, [C.c_uint64, C.c_size_t, C.c_size_t, C.c_double, C.c_double, C.c_int,
                                          C.c_double, _u32p, _u32p, _dp, _dp, _dp, _dp]),
    "spx_dipolar_graphs": (C.c_int, [C.c_uint64, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double, C.c_double,
                                     C.c_int, C.c_double, _dp, _dp]),
    "spx_dipolar_state": (C.c_int, [C.c_uint64, C.c_size_t, C.c_size_t, C.c_double, C.c_double, C.c_int,
                                    C.c_double, _dp]),
    "spx_gnn_gradient": (C.c_int, [_vp, C.c_size_t, C.c_size_t, _dp, _dp, _fp, _fp, _fp]),
    "spx_gnn_train_step": (C.c_int, [_vp, C.c_size_t, C.c_size_t, _dp, _dp, _fp, C.c_float, _fp]),
}

_lib = None


def lib():
    """The loaded libspx; raises SpxMissing when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SpxMissing(f"{LIB_PATH} is missing: build it with `python -m spinsim_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def check(rc: int) -> None:
    if rc == OK:
        return
    msg = (lib().spx_last_error() or b"").decode()
    if rc == EINVAL:
        raise InvalidArgument(rc, msg)
    if rc == ERANGE:
        raise OutOfRange(rc, msg)
    if rc == EDOMAIN:
        raise DomainError(rc, msg)
    if rc == ECONVERGE:
        b, t, e, o = C.c_double(), C.c_double(), C.c_double(), C.c_size_t()
        lib().spx_last_convergence(C.byref(b), C.byref(t), C.byref(e), C.byref(o))
        raise ConvergenceError(rc, msg, b.value, t.value, e.value, o.value)
    if rc == EOVERFLOW:
        raise OverflowErr(rc, msg)
    raise SpxError(rc, msg)


# ------------------------------------------------------------------- host numerics

def bessel(kind: str, x: float, count: int):
    out = np.zeros(max(count, 1))
    nz = C.c_size_t()
    check(lib().spx_bessel(0 if kind == "J" else 1, x, count, _d(out), C.byref(nz)))
    return out[:count], nz.value


def cheb_build(mode: int, bound: float, tau: float, eps: float, max_order: int = 6000):
    cap = max_order + 8
    buf = np.zeros(cap)
    order = C.c_size_t()
    eff = C.c_double()
    check(lib().spx_cheb_build(mode, bound, tau, eps, max_order, _d(buf), cap, C.byref(order),
                               C.byref(eff)))
    return buf[: order.value + 1].copy(), eff.value


# ------------------------------------------------------------------------ objects

class Context:
    def __init__(self, device: int = 0):
        h = _vp()
        check(lib().spx_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().spx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().spx_ctx_synchronize(self.h))

    def record(self, slot: int):
        check(lib().spx_ctx_event_record(self.h, slot))

    def elapsed_ms(self, s0: int, s1: int) -> float:
        ms = C.c_float()
        check(lib().spx_ctx_event_elapsed(self.h, s0, s1, C.byref(ms)))
        return ms.value


class StateVector:
    """Device state of 2^L amplitudes (include/spinsim/core/StateVector.hpp)."""

    def __init__(self, ctx: Context, nspins: int, data: Optional[np.ndarray] = None):
        self.ctx = ctx
        self.nspins = nspins
        h = _vp()
        check(lib().spx_state_create(ctx.h, nspins, C.byref(h)))
        self.h = h
        if data is not None:
            self.upload(data)

    def __del__(self):
        try:
            if self.h:
                lib().spx_state_destroy(self.h)
        except Exception:
            pass

    @property
    def size(self) -> int:
        return 1 << self.nspins

    def upload(self, data: np.ndarray):
        a = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
        if a.size != 2 * self.size:
            raise ValueError("host buffer does not match 2*2^L doubles")
        check(lib().spx_state_upload(self.h, _d(a)))

    def download(self) -> np.ndarray:
        out = np.zeros(2 * self.size)
        check(lib().spx_state_download(self.h, _d(out)))
        return out

    def copy_from(self, other: "StateVector"):
        check(lib().spx_state_copy(self.h, other.h))

    def set_zero(self):
        check(lib().spx_state_set_zero(self.h))

    def set_basis(self, index: int):
        check(lib().spx_state_set_basis(self.h, index))

    def scale(self, f: float):
        check(lib().spx_state_scale(self.h, f))

    def negate(self):
        check(lib().spx_state_negate(self.h))

    def add_scaled(self, cr: float, ci: float, other: "StateVector"):
        check(lib().spx_state_add_scaled(self.h, cr, ci, other.h))

    def norm(self) -> float:
        v = C.c_double()
        check(lib().spx_state_norm(self.h, C.byref(v)))
        return v.value

    def normalize(self) -> float:
        v = C.c_double()
        check(lib().spx_state_normalize(self.h, C.byref(v)))
        return v.value

    def inner(self, other: "StateVector"):
        re, im = C.c_double(), C.c_double()
        check(lib().spx_state_inner(self.h, other.h, C.byref(re), C.byref(im)))
        return re.value, im.value


class Hamiltonian:
    """include/spinsim/core/Hamiltonian.hpp, device-backed."""

    def __init__(self, ctx: Context, nspins: int):
        self.ctx = ctx
        self.nspins = nspins
        h = _vp()
        check(lib().spx_ham_create(ctx.h, nspins, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().spx_ham_destroy(self.h)
        except Exception:
            pass

    def set_field(self, spin, x=0.0, y=0.0, z=0.0):
        check(lib().spx_ham_set_field(self.h, spin, x, y, z))

    def field(self, spin):
        out = np.zeros(3)
        check(lib().spx_ham_get_field(self.h, spin, _d(out)))
        return out

    def set_fields(self, fields: np.ndarray):
        f = np.asarray(fields, dtype=float).reshape(self.nspins, 3)
        for s in range(self.nspins):
            self.set_field(s, *f[s])

    def set_coupling(self, a, b, jx, jy, jz):
        check(lib().spx_ham_set_coupling(self.h, a, b, jx, jy, jz))

    def set_couplings(self, lo, hi, jx, jy, jz):
        for p in range(len(lo)):
            self.set_coupling(int(lo[p]), int(hi[p]), float(jx[p]), float(jy[p]), float(jz[p]))

    def num_couplings(self) -> int:
        n = C.c_size_t()
        check(lib().spx_ham_num_couplings(self.h, C.byref(n)))
        return n.value

    def spectral_bound(self) -> float:
        v = C.c_double()
        check(lib().spx_ham_spectral_bound(self.h, C.byref(v)))
        return v.value

    def autotune(self, bound: float, nvec: int = 1) -> str:
        """Time the eligible order-kernel schedules at this step bound, with
        `nvec` states sharing each launch, and keep the fastest; returns the
        timing report."""
        buf = C.create_string_buffer(4096)
        check(lib().spx_ham_autotune_batch(self.h, bound, nvec, buf, 4096))
        return buf.value.decode()

    def set_schedule(self, tile_bits=0, block=0, xcd=-1, dense=-1, dense_tile_bits=0):
        check(lib().spx_ham_set_schedule(self.h, tile_bits, block, xcd, dense, dense_tile_bits))

    def kernel_name(self) -> str:
        """The order kernel the most recently prepared scale dispatches to."""
        buf = C.create_string_buffer(128)
        check(lib().spx_ham_schedule(self.h, buf, 128))
        return buf.value.decode()

    def apply_physical(self, inp: StateVector, out: StateVector):
        check(lib().spx_ham_apply(self.h, 0, 0.0, inp.h, out.h))

    def apply_twice_normalized(self, bound: float, inp: StateVector, out: StateVector):
        check(lib().spx_ham_apply(self.h, 1, bound, inp.h, out.h))

    def accumulate_twice_normalized(self, bound: float, inp: StateVector, out: StateVector):
        check(lib().spx_ham_apply(self.h, 2, bound, inp.h, out.h))

    def diagonal(self) -> np.ndarray:
        out = np.zeros(1 << self.nspins)
        check(lib().spx_ham_diagonal(self.h, _d(out)))
        return out


class Propagator:
    """include/spinsim/core/Propagator.hpp, device-backed."""

    def __init__(self, ctx: Context, nspins: int, epsilon: float = 1e-7):
        self.ctx = ctx
        h = _vp()
        check(lib().spx_prop_create(ctx.h, nspins, epsilon, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().spx_prop_destroy(self.h)
        except Exception:
            pass

    def evolve(self, ham: Hamiltonian, mode: int, psi: StateVector, tau: float, nsteps: int = 1,
               autonormalize: bool = False):
        check(lib().spx_prop_evolve(self.h, ham.h, mode, psi.h, tau, nsteps, int(autonormalize)))

    def set_graphs(self, enable: bool) -> None:
        check(lib().spx_prop_set_graphs(self.h, int(enable)))

    def set_streams(self, streams: int) -> None:
        """2 (default): a set on the Hadamard path runs on two streams whose
        launches overlap (4..8 states as two halves with no Hamiltonian in
        both; more than 8 as launch groups alternating between the streams);
        1: one stream."""
        check(lib().spx_prop_set_streams(self.h, int(streams)))

    def evolve_many(self, ham: Hamiltonian, mode: int, states: Sequence[StateVector], tau: float,
                    nsteps: int = 1, autonormalize: bool = False):
        arr = (_vp * len(states))(*[s.h for s in states])
        check(lib().spx_prop_evolve_many(self.h, ham.h, mode, arr, len(states), tau, nsteps,
                                         int(autonormalize)))

    def evolve_set(self, hams: Sequence[Hamiltonian], mode: int, states: Sequence[StateVector], tau: float,
                   nsteps: int = 1, autonormalize: bool = False):
        """State i under hams[i], each with its own expansion; up to sixteen
        states share every order launch where their Hamiltonians allow."""
        if len(hams) != len(states):
            raise InvalidArgument(EINVAL, "one Hamiltonian per state")
        ha = (_vp * len(hams))(*[h.h for h in hams])
        sa = (_vp * len(states))(*[s.h for s in states])
        check(lib().spx_prop_evolve_set(self.h, ha, mode, sa, len(states), tau, nsteps, int(autonormalize)))

    def steps_set(self, hams: Sequence[Hamiltonian], mode: int, states: Sequence[StateVector],
                  coeffs: Sequence[np.ndarray], bounds: Sequence[float], nsteps: int, autonormalize: bool = False):
        """evolve_set over prebuilt expansions: state i takes coeffs[i] at bounds[i]."""
        n = len(states)
        if not (len(hams) == len(coeffs) == len(bounds) == n):
            raise InvalidArgument(EINVAL, "one Hamiltonian, expansion and bound per state")
        cs = [np.ascontiguousarray(c, dtype=float) for c in coeffs]
        ha = (_vp * n)(*[h.h for h in hams])
        sa = (_vp * n)(*[s.h for s in states])
        ca = (_dp * n)(*[_d(c) for c in cs])
        orders = np.array([c.size - 1 for c in cs], dtype=np.uintp)
        b = np.ascontiguousarray(bounds, dtype=float)
        check(lib().spx_prop_steps_set(self.h, ha, mode, sa, n, ca, orders.ctypes.data_as(_szp), _d(b), nsteps,
                                       int(autonormalize)))

    def steps(self, ham: Hamiltonian, mode: int, psi: StateVector, coeffs: np.ndarray, bound: float,
              nsteps: int, autonormalize: bool = False):
        c = np.ascontiguousarray(coeffs, dtype=float)
        check(lib().spx_prop_steps(self.h, ham.h, mode, psi.h, _d(c), c.size - 1, bound, nsteps,
                                   int(autonormalize)))

    def last_order(self) -> int:
        v = C.c_size_t()
        check(lib().spx_prop_last_order(self.h, C.byref(v)))
        return v.value

    def epsilon(self) -> float:
        v = C.c_double()
        check(lib().spx_prop_epsilon(self.h, C.byref(v)))
        return v.value


def autotune_mode(mode: int) -> None:
    """0: default schedules; 1: tune each new problem shape on first evolve."""
    check(lib().spx_autotune_mode(mode))


def rotate(psi: StateVector, axis, angle: float, targets: Iterable[int]):
    t = np.ascontiguousarray(list(targets), dtype=np.uint32)
    check(lib().spx_rotate(psi.h, axis[0], axis[1], axis[2], angle, t.ctypes.data_as(_u32p), t.size))


def two_spin_gate(psi: StateVector, u: np.ndarray, a: int, b: int):
    m = np.ascontiguousarray(np.asarray(u, dtype=complex).reshape(16)).view(np.float64)
    check(lib().spx_two_spin_gate(psi.h, _d(m), a, b))


def apply_sigma(axis: int, c: float, inp: StateVector, spin: int, out: StateVector, accumulate: bool):
    check(lib().spx_apply_sigma(axis, c, inp.h, spin, out.h, int(accumulate)))


def apply_total_spin(axis: int, inp: StateVector, out: StateVector):
    check(lib().spx_apply_total_spin(axis, inp.h, out.h))


def measure_corr(axis: int, state: StateVector, ref: StateVector):
    re, im = C.c_double(), C.c_double()
    check(lib().spx_measure_corr(axis, state.h, ref.h, C.byref(re), C.byref(im)))
    return re.value, im.value


def measure_mag(axis: int, spin: int, state: StateVector) -> float:
    v = C.c_double()
    check(lib().spx_measure_mag(axis, spin, state.h, C.byref(v)))
    return v.value


def measure_energy(ham: Hamiltonian, state: StateVector) -> float:
    v = C.c_double()
    check(lib().spx_measure_energy(ham.h, state.h, C.byref(v)))
    return v.value


class Program:
    """Flattened Sequence AST (include/spinsim/sequence/Sequence.hpp)."""

    def __init__(self):
        h = _vp()
        check(lib().spx_program_create(C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().spx_program_destroy(self.h)
        except Exception:
            pass

    def evolve(self, tau, nsteps=1, mode=REAL, autonorm=False, fields=None):
        if fields is None:
            check(lib().spx_program_evolve(self.h, tau, nsteps, mode, int(autonorm), None, 0))
        else:
            f = np.ascontiguousarray(fields, dtype=float).reshape(-1)
            check(lib().spx_program_evolve(self.h, tau, nsteps, mode, int(autonorm), _d(f), f.size // 3))
        return self

    def evolve_uniform(self, tau, nsteps=1, hx=0.0, hy=0.0, hz=0.0, mode=REAL, autonorm=False):
        check(lib().spx_program_evolve_uniform(self.h, tau, nsteps, mode, int(autonorm), hx, hy, hz))
        return self

    def pulse(self, axis, angle, targets=()):
        t = np.ascontiguousarray(list(targets), dtype=np.uint32)
        check(lib().spx_program_pulse(self.h, axis[0], axis[1], axis[2], angle,
                                      t.ctypes.data_as(_u32p) if t.size else None, t.size))
        return self

    def measure(self):
        check(lib().spx_program_measure(self.h))
        return self

    def repeat(self, count):
        check(lib().spx_program_repeat(self.h, count))
        return self

    def end(self):
        check(lib().spx_program_end(self.h))
        return self

    def measurements(self) -> int:
        v = C.c_size_t()
        check(lib().spx_program_measurements(self.h, C.byref(v)))
        return v.value


class Batch:
    """Many realisations of one experiment (spx_batch_*)."""

    def __init__(self, ctx: Context, program: Program, observables, epsilon=1e-7, reference_axis=0):
        self.ctx = ctx
        self.program = program
        codes = np.array([o[0] for o in observables], dtype=np.int32)
        spins = np.array([o[1] for o in observables], dtype=np.int32)
        h = _vp()
        check(lib().spx_batch_create(ctx.h, program.h, epsilon, reference_axis,
                                     codes.ctypes.data_as(_ip), spins.ctypes.data_as(_ip), codes.size,
                                     C.byref(h)))
        self.h = h
        self.count = 0

    def __del__(self):
        try:
            if self.h:
                lib().spx_batch_destroy(self.h)
        except Exception:
            pass

    def width(self) -> int:
        v = C.c_size_t()
        check(lib().spx_batch_width(self.h, C.byref(v)))
        return v.value

    def add(self, nspins, fields, lo, hi, jx, jy, jz, psi0):
        f = np.ascontiguousarray(fields, dtype=float).reshape(-1)
        lo = np.ascontiguousarray(lo, dtype=np.uint32)
        hi = np.ascontiguousarray(hi, dtype=np.uint32)
        jx, jy, jz = (np.ascontiguousarray(v, dtype=float) for v in (jx, jy, jz))
        p = np.ascontiguousarray(psi0, dtype=float)
        check(lib().spx_batch_add(self.h, nspins, _d(f), lo.size, lo.ctypes.data_as(_u32p),
                                  hi.ctypes.data_as(_u32p), _d(jx), _d(jy), _d(jz), _d(p)))
        self.count += 1

    def add_dipolar(self, seed: int, first: int, count: int, nspins: int, scale: float = 1.0,
                    min_separation: float = 0.5, angular: bool = True, disorder: float = 0.5):
        """Realisations first .. first+count-1 of the dipolar ensemble with
        their random initial states, drawn inside libspx (runRealization order)."""
        check(lib().spx_batch_add_dipolar(self.h, seed, first, count, nspins, scale, min_separation,
                                          int(angular), disorder))
        self.count += count

    def run(self):
        points = self.program.measurements()
        w = self.width()
        times = np.zeros(max(points, 1))
        values = np.zeros(max(self.count * points * w, 1))
        check(lib().spx_batch_run(self.h, _d(times), _d(values)))
        return times[:points], values[: self.count * points * w].reshape(self.count, points, w)

    def clear(self):
        check(lib().spx_batch_clear(self.h))
        self.count = 0


class Dipolar:
    """One pair record of a drawn realisation: lo, hi, jx, jy, jz, hz."""

    def __init__(self, lo, hi, jx, jy, jz, hz):
        self.lo, self.hi, self.jx, self.jy, self.jz, self.hz = lo, hi, jx, jy, jz, hz


def dipolar_realisation(seed: int, index: int, nspins: int, scale: float = 1.0, min_separation: float = 0.5,
                        angular: bool = True, disorder: float = 0.5) -> Dipolar:
    """Realisation (seed, index): couplings in generation order and Sz disorder."""
    n = nspins * (nspins - 1) // 2
    lo = np.zeros(n, np.uint32)
    hi = np.zeros(n, np.uint32)
    jx, jy, jz = np.zeros(n), np.zeros(n), np.zeros(n)
    hz = np.zeros(nspins)
    check(lib().spx_dipolar_realisation(seed, index, nspins, scale, min_separation, int(angular), disorder,
                                        lo.ctypes.data_as(_u32p), hi.ctypes.data_as(_u32p), _d(jx), _d(jy),
                                        _d(jz), _d(hz)))
    return Dipolar(lo, hi, jx, jy, jz, hz)


def evolve_noalloc(ham: "Hamiltonian", mode: int, psi: "StateVector", tau: float, nsteps: int = 1,
                   epsilon: float = 1e-7) -> int:
    """spx_prop_evolve_noalloc: Propagator::evolve on one state from the
    library's static scratch (no device allocation; at most 16 spins).
    Returns the expansion order."""
    order = C.c_size_t()
    check(lib().spx_prop_evolve_noalloc(ham.h, mode, psi.h, tau, nsteps, epsilon, C.byref(order)))
    return order.value


def dipolar_state(seed: int, index: int, nspins: int, scale: float = 1.0, min_separation: float = 0.5,
                  angular: bool = True, disorder: float = 0.5) -> np.ndarray:
    """The random initial state of realisation (seed, index), interleaved."""
    psi = np.zeros(2 << nspins)
    check(lib().spx_dipolar_state(seed, index, nspins, scale, min_separation, int(angular), disorder, _d(psi)))
    return psi


def dipolar_graphs(seed: int, first: int, count: int, nspins: int, scale: float = 1.0,
                   min_separation: float = 0.5, angular: bool = True, disorder: float = 0.5):
    """Realisations first .. first+count-1 as J [count, L, L] and h [count, L]."""
    J = np.zeros((count, nspins, nspins))
    h = np.zeros((count, nspins))
    check(lib().spx_dipolar_graphs(seed, first, count, nspins, scale, min_separation, int(angular), disorder,
                                   _d(J), _d(h)))
    return J, h


class Surrogate:
    """The graph-network surrogate of an echo curve (spx_gnn_*): couplings
    J [G, L, L] (symmetric) and fields h [G, L] in, P curve points out."""

    def __init__(self, ctx: Context, outputs: int, seed: int = 0, scale_h: float = 1.0, scale_j: float = 1.0):
        self.ctx = ctx
        self.outputs = outputs
        h = _vp()
        check(lib().spx_gnn_create(ctx.h, outputs, seed, scale_h, scale_j, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().spx_gnn_destroy(self.h)
        except Exception:
            pass

    def num_params(self) -> int:
        v = C.c_size_t()
        check(lib().spx_gnn_num_params(self.h, C.byref(v)))
        return v.value

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        check(lib().spx_gnn_get_params(self.h, out.ctypes.data_as(_fp)))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params():
            raise ValueError("parameter vector has the wrong length")
        check(lib().spx_gnn_set_params(self.h, p.ctypes.data_as(_fp)))

    @staticmethod
    def _graphs(J, h):
        J = np.ascontiguousarray(J, dtype=float)
        h = np.ascontiguousarray(h, dtype=float)
        if J.ndim != 3 or J.shape[1] != J.shape[2] or h.shape != J.shape[:2]:
            raise ValueError("J must be [G, L, L] and h [G, L]")
        return J, h

    def predict(self, J, h) -> np.ndarray:
        J, h = self._graphs(J, h)
        out = np.zeros((J.shape[0], self.outputs), dtype=np.float32)
        check(lib().spx_gnn_predict(self.h, J.shape[0], J.shape[1], _d(J), _d(h), out.ctypes.data_as(_fp)))
        return out

    def predict_dipolar(self, seed: int, first: int, count: int, nspins: int, scale: float = 1.0,
                        min_separation: float = 0.5, angular: bool = True, disorder: float = 0.5) -> np.ndarray:
        out = np.zeros((count, self.outputs), dtype=np.float32)
        check(lib().spx_gnn_predict_dipolar(self.h, seed, first, count, nspins, scale, min_separation,
                                            int(angular), disorder, out.ctypes.data_as(_fp)))
        return out

    def _targets(self, J, target):
        t = np.ascontiguousarray(target, dtype=np.float32)
        if t.shape != (J.shape[0], self.outputs):
            raise ValueError("targets must be [G, outputs]")
        return t

    def gradient(self, J, h, target):
        J, h = self._graphs(J, h)
        t = self._targets(J, target)
        loss = C.c_float()
        grad = np.zeros(self.num_params(), dtype=np.float32)
        check(lib().spx_gnn_gradient(self.h, J.shape[0], J.shape[1], _d(J), _d(h), t.ctypes.data_as(_fp),
                                     C.byref(loss), grad.ctypes.data_as(_fp)))
        return loss.value, grad

    def set_training_set(self, J, h, target) -> None:
        J, h = self._graphs(J, h)
        t = self._targets(J, target)
        check(lib().spx_gnn_set_training_set(self.h, J.shape[0], J.shape[1], _d(J), _d(h),
                                             t.ctypes.data_as(_fp)))
        self.set_n = J.shape[0]

    def train_epoch(self, order, batch: int, lr: float) -> float:
        o = np.ascontiguousarray(order, dtype=np.uint32)
        if o.size != self.set_n:
            raise ValueError("order must be a permutation of the training set")
        loss = C.c_float()
        check(lib().spx_gnn_train_epoch(self.h, batch, o.ctypes.data_as(_u32p), lr, C.byref(loss)))
        return loss.value

    def train_step(self, J, h, target, lr: float) -> float:
        J, h = self._graphs(J, h)
        t = self._targets(J, target)
        loss = C.c_float()
        check(lib().spx_gnn_train_step(self.h, J.shape[0], J.shape[1], _d(J), _d(h), t.ctypes.data_as(_fp),
                                       lr, C.byref(loss)))
        return loss.value
