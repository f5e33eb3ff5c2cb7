"""The HIP backend behind the reference's Backend trait (src/exec.rs:134-162).

`HipBackend` is a thin ctypes handle over the zb_* C ABI; every method maps
to one trait method.  There is no CPU fallback: construction fails loudly
when the library or a HIP device is missing.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import lib, ptr


class ZenoError(Exception):
    """Crate error (src/lib.rs:77-107)."""


class Unsupported(ZenoError):
    pass


class MemoryBudget(ZenoError):
    pass


class DeviceError(ZenoError):
    pass


def _raise(rc: int, what: str):
    msg = f"{what}: {_lib.last_error()}"
    if rc == _lib.ZB_ERR_UNSUPPORTED:
        raise Unsupported(msg)
    if rc == _lib.ZB_ERR_MEMORY:
        raise MemoryBudget(msg)
    if rc == _lib.ZB_ERR_DEVICE:
        raise DeviceError(msg)
    raise ValueError(msg)


def _check(rc: int, what: str) -> int:
    if rc < 0:
        _raise(rc, what)
    return rc


def device_count() -> int:
    """Visible HIP devices (0 without a GPU)."""
    return int(lib().zb_device_count())


def device_memory():
    """(free, total) device memory in bytes."""
    f, t = C.c_uint64(), C.c_uint64()
    _check(lib().zb_mem_info(C.byref(f), C.byref(t)), "mem_info")
    return int(f.value), int(t.value)


def calibration_copy(n_doubles: int, reps: int = 3) -> float:
    """Average ms of a plain f64 device copy of n_doubles elements (8 B read
    and 8 B written each)."""
    ms = C.c_double()
    _check(lib().zb_calibration_copy(n_doubles, reps, C.byref(ms)), "calibration_copy")
    return float(ms.value)


def copy_probe(n_doubles: int, kind: int, reps: int = 3) -> float:
    """Average ms of a streaming copy probe (zb_copy_probe kinds 0..3)."""
    ms = C.c_double()
    _check(lib().zb_copy_probe(n_doubles, reps, kind, C.byref(ms)), "copy_probe")
    return float(ms.value)


class HipBackend:
    """make_backend(n, precision, BackendChoice::Hip) (src/exec.rs:306-331)."""

    EXACT, TILED = 0, 1

    def __init__(self, n_qubits: int, precision: int = 64, mode: int = TILED):
        h = C.c_void_p()
        _check(lib().zb_create(n_qubits, precision, C.byref(h)), "zb_create")
        self._h = h
        self.n = n_qubits
        self.precision = precision
        self.set_mode(mode)

    def close(self):
        if getattr(self, "_h", None):
            lib().zb_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def name(self) -> str:
        return lib().zb_name(self._h).decode()

    def set_mode(self, mode: int):
        _check(lib().zb_set_mode(self._h, mode), "zb_set_mode")

    def apply_unitary(self, qs, mat):
        qs = np.ascontiguousarray(qs, np.uint32)
        m = np.ascontiguousarray(mat, np.complex128).reshape(-1).view(np.float64)
        _check(lib().zb_apply_unitary(self._h, ptr(qs), len(qs), ptr(m)), "apply_unitary")

    def apply_diagonal(self, qs, d):
        qs = np.ascontiguousarray(qs, np.uint32)
        m = np.ascontiguousarray(d, np.complex128).view(np.float64)
        _check(lib().zb_apply_diagonal(self._h, ptr(qs), len(qs), ptr(m)), "apply_diagonal")

    def apply_cx(self, control: int, target: int):
        _check(lib().zb_apply_cx(self._h, control, target), "apply_cx")

    def prob_one(self, q: int) -> float:
        r = lib().zb_prob_one(self._h, q)
        if math.isnan(r):
            _raise(_lib.ZB_ERR_ARG, "prob_one")
        return r

    def measure(self, q: int, u: float) -> bool:
        return bool(_check(lib().zb_measure(self._h, q, u), "measure"))

    def reset_all(self):
        _check(lib().zb_reset_all(self._h), "reset_all")

    def sample(self, shots: int, seed: int) -> np.ndarray:
        out = np.zeros(max(shots, 1), np.uint64)
        _check(lib().zb_sample(self._h, shots, seed, ptr(out)), "sample")
        return out[:shots]

    def statevector(self) -> np.ndarray:
        out = np.empty(2 << self.n, np.float64)
        _check(lib().zb_statevector(self._h, ptr(out)), "statevector")
        return out.view(np.complex128)

    def norm_sqr(self) -> float:
        r = lib().zb_norm_sqr(self._h)
        if math.isnan(r):
            _raise(_lib.ZB_ERR_DEVICE, "norm_sqr")
        return r

    def finish(self):
        _check(lib().zb_finish(self._h), "finish")

    # -- extensions -------------------------------------------------------
    def raw(self):
        dt = np.float64 if self.precision == 64 else np.float32
        re = np.empty(1 << self.n, dt)
        im = np.empty(1 << self.n, dt)
        _check(lib().zb_state_raw(self._h, ptr(re), ptr(im)), "state_raw")
        return re, im

    def set_raw(self, re, im):
        dt = np.float64 if self.precision == 64 else np.float32
        re = np.ascontiguousarray(re, dt)
        im = np.ascontiguousarray(im, dt)
        _check(lib().zb_state_set_raw(self._h, ptr(re), ptr(im)), "state_set_raw")

    def chunk_sums(self) -> np.ndarray:
        ln = 1 << self.n
        ch = min(ln, 1 << 16)
        out = np.zeros((ln + ch - 1) // ch, np.float64)
        _check(lib().zb_chunk_sums(self._h, ptr(out)), "chunk_sums")
        return out

    def stream(self) -> int:
        """The hipStream_t the backend enqueues on (for events)."""
        return lib().zb_stream(self._h)

    def max_abs_diff(self, other: "HipBackend") -> float:
        out = C.c_double(0.0)
        _check(lib().zb_max_abs_diff(self._h, other._h, C.byref(out)), "max_abs_diff")
        return out.value

    def layout(self):
        """phys[q]: the state-index bit that holds circuit qubit q."""
        out = np.zeros(max(self.n, 1), np.uint32)
        lib().zb_layout(self._h, ptr(out))
        return [int(x) for x in out[: self.n]]

    def tune(self, key: str, value: int):
        """Set a tiled-kernel knob ("prefetch", "win_grid")."""
        _check(lib().zb_tune(self._h, key.encode(), int(value)), "tune")

    def profile_begin(self):
        """Record HIP events around every state-sweep launch from now on."""
        _check(lib().zb_profile_begin(self._h), "profile_begin")

    def profile_end(self, cap: int = 1 << 16):
        """[(kernel name, ms, algorithmic bytes)] for the launches since
        profile_begin."""
        k = np.zeros(cap, np.int32)
        ms = np.zeros(cap, np.float64)
        by = np.zeros(cap, np.float64)
        n = _check(lib().zb_profile_end(self._h, ptr(k), ptr(ms), ptr(by), cap), "profile_end")
        return [(_lib.KERNEL_NAMES.get(int(k[i]), str(int(k[i]))), float(ms[i]), float(by[i])) for i in range(n)]

    def profile_detail(self, cap: int = 1 << 16):
        """Per launch of the last profile_end: (ops, stages, diagonal ops,
        CX ops) for windows, zeros for other kernels."""
        d = np.zeros(4 * cap, np.int32)
        n = int(lib().zb_profile_detail(self._h, ptr(d), cap))
        return [tuple(int(x) for x in d[4 * i:4 * i + 4]) for i in range(n)]

    def apply_compiled(self, compiled, lo: int = 0, hi: int | None = None):
        hi = len(compiled) if hi is None else hi
        _check(lib().zr_apply_ops(self._h, compiled.handle, lo, hi), "apply_ops")
