"""CPU oracle for the zeno state-vector path — TEST INFRASTRUCTURE ONLY.

This package is the checker: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it; the zeno_amd product never does.  It wraps
oracle/libzeno_oracle.so (oracle/zeno_oracle.c), a plain-C restatement of
the reference's CPU path: gates.rs, circuit.rs, compiler.rs, state.rs,
sample.rs, exec.rs and noise.rs, plus RefSim from tests/reference.rs:69-116.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libzeno_oracle.so")

INSTR_DTYPE = np.dtype(
    [
        ("name", "S8"),
        ("kind", "<i4"),
        ("nq", "<i4"),
        ("np", "<i4"),
        ("q", "<u4", (3,)),
        ("clbit", "<u4"),
        ("cond_creg", "<i4"),
        ("cond_value", "<u8"),
        ("p", "<f8", (3,)),
    ],
    align=True,
)
assert INSTR_DTYPE.itemsize == 72

OP_U, OP_D, OP_CX, OP_MEAS, OP_RESET = 0, 1, 2, 3, 4


class Noise(C.Structure):
    _fields_ = [
        (f, C.c_double)
        for f in (
            "depolarizing_1q",
            "depolarizing_2q",
            "bit_flip",
            "phase_flip",
            "amplitude_damping",
            "readout_flip_0to1",
            "readout_flip_1to0",
        )
    ]


def build(force: bool = False) -> str:
    """Compile oracle/zeno_oracle.c with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P, U64, I64, I32, U32, D = C.c_void_p, C.c_uint64, C.c_int64, C.c_int, C.c_uint32, C.c_double
        sig = {
            "zo_splitmix64": (U64, [U64]),
            "zo_rng_u64_stream": (None, [U64, U64, P]),
            "zo_rng_f64_stream": (None, [U64, U64, P]),
            "zo_xoshiro_from_state": (None, [P, U64, P]),
            "zo_rng_range_stream": (None, [U64, U64, U64, U64, P]),
            "zo_rng_range_from_state": (None, [P, U64, U64, U64, P]),
            "zo_gate_build": (I32, [C.c_char_p, P, I32, P, P]),
            "zo_gate_lookup": (I32, [C.c_char_p, P, P]),
            "zo_gen_random": (I64, [U32, U32, U64, P, I64]),
            "zo_gen_qft": (I64, [U32, P, I64]),
            "zo_gen_u3_qft": (I64, [U32, U64, P, I64]),
            "zo_gen_matching": (I64, [U32, U32, U64, U32, P, I64]),
            "zo_compile": (P, [P, I64, U32, P, U32, I32, I32]),
            "zo_compiled_free": (None, [P]),
            "zo_compiled_nops": (I64, [P]),
            "zo_compiled_op": (I32, [P, I64, P, P, P, P, P, P]),
            "zo_compiled_stats": (None, [P, P]),
            "zo_compiled_final_measures": (I64, [P, P]),
            "zo_error": (C.c_char_p, []),
            "zo_state_new": (P, [U32, I32]),
            "zo_state_free": (None, [P]),
            "zo_set_threads": (None, [I32]),
            "zo_get_threads": (I32, []),
            "zo_reset_all": (None, [P]),
            "zo_apply_unitary": (None, [P, P, I32, P]),
            "zo_apply_diagonal": (None, [P, P, I32, P]),
            "zo_apply_cx": (None, [P, U32, U32]),
            "zo_prob_one": (D, [P, U32]),
            "zo_measure": (I32, [P, U32, D]),
            "zo_norm_sqr": (D, [P]),
            "zo_statevector": (None, [P, P]),
            "zo_state_raw": (None, [P, P, P]),
            "zo_state_set_raw": (None, [P, P, P]),
            "zo_apply_compiled": (None, [P, P, I64, I64]),
            "zo_sample": (None, [P, U64, U64, P]),
            "zo_chunk_sums": (None, [P, P]),
            "zo_run": (I64, [P, I32, U64, U64, P, P, P, I64, P]),
            "zo_run_shots": (I32, [P, I32, U64, U64, P, I32, P]),
            "zo_refsim_apply": (None, [P, U32, P, I32, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ RNG

def splitmix64(x: int) -> int:
    return lib().zo_splitmix64(x)


def rng_u64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib().zo_rng_u64_stream(seed, n, _ptr(out))
    return out


def xoshiro_from_state(state, n: int) -> np.ndarray:
    st = np.asarray(state, np.uint64)
    out = np.empty(n, np.uint64)
    lib().zo_xoshiro_from_state(_ptr(st), n, _ptr(out))
    return out


def rng_range_from_state(state, lo: int, hi: int, n: int) -> np.ndarray:
    st = np.asarray(state, np.uint64)
    out = np.empty(n, np.uint64)
    lib().zo_rng_range_from_state(_ptr(st), lo, hi, n, _ptr(out))
    return out


def rng_f64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().zo_rng_f64_stream(seed, n, _ptr(out))
    return out


def rng_range(seed: int, lo: int, hi: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib().zo_rng_range_stream(seed, lo, hi, n, _ptr(out))
    return out


# ---------------------------------------------------------------- gates

def gate(name: str, params=()):
    """(matrix or diagonal as complex128, is_diag) or None for unknown."""
    p = np.asarray(params, np.float64)
    out = np.zeros(2 * 64, np.float64)
    isd = C.c_int(0)
    k = lib().zo_gate_build(name.encode(), _ptr(p) if len(p) else None, len(p), _ptr(out), C.byref(isd))
    if k < 0:
        return None
    n = (1 << k) if isd.value else (1 << (2 * k))
    z = out[: 2 * n].view(np.complex128).copy()
    if not isd.value:
        z = z.reshape(1 << k, 1 << k)
    return z, bool(isd.value)


# ------------------------------------------------------------- programs

def instr(name, qubits, params=(), cond=None):
    r = np.zeros(1, INSTR_DTYPE)
    r["name"] = name.encode()
    r["kind"] = 0
    r["nq"] = len(qubits)
    r["np"] = len(params)
    r["q"][0, : len(qubits)] = qubits
    r["p"][0, : len(params)] = params
    r["cond_creg"] = -1 if cond is None else cond[0]
    r["cond_value"] = 0 if cond is None else cond[1]
    return r


def measure(q, clbit, cond=None):
    r = np.zeros(1, INSTR_DTYPE)
    r["kind"] = 1
    r["q"][0, 0] = q
    r["clbit"] = clbit
    r["cond_creg"] = -1 if cond is None else cond[0]
    r["cond_value"] = 0 if cond is None else cond[1]
    return r


def reset(q, cond=None):
    r = np.zeros(1, INSTR_DTYPE)
    r["kind"] = 2
    r["q"][0, 0] = q
    r["cond_creg"] = -1 if cond is None else cond[0]
    r["cond_value"] = 0 if cond is None else cond[1]
    return r


def barrier():
    r = np.zeros(1, INSTR_DTYPE)
    r["kind"] = 3
    r["cond_creg"] = -1
    return r


def program(rows) -> np.ndarray:
    return np.concatenate(rows) if rows else np.zeros(0, INSTR_DTYPE)


def measure_all(n):
    return program([measure(q, q) for q in range(n)])


def _gen(fn, *args):
    cnt = fn(*args, None, 0)
    out = np.zeros(max(cnt, 1), INSTR_DTYPE)
    fn(*args, _ptr(out), cnt)
    return out[:cnt]


def gen_random(n, depth, seed):
    return _gen(lib().zo_gen_random, n, depth, seed)


def gen_qft(n):
    return _gen(lib().zo_gen_qft, n)


def gen_u3_qft(n, seed):
    return _gen(lib().zo_gen_u3_qft, n, seed)


def gen_matching(n, depth, seed, rzz_period=0):
    return _gen(lib().zo_gen_matching, n, depth, seed, rzz_period)


# ------------------------------------------------------------- compiler

class CompileError(Exception):
    pass


class Compiled:
    def __init__(self, prog: np.ndarray, n_qubits: int, cregs=None, fusion: int = 1, diag_max: int = 10):
        prog = np.ascontiguousarray(prog, INSTR_DTYPE)
        cregs = np.asarray([n_qubits] if cregs is None else cregs, np.uint32)
        self.n_qubits = n_qubits
        self._h = lib().zo_compile(_ptr(prog), len(prog), n_qubits, _ptr(cregs), len(cregs), fusion, diag_max)
        if not self._h:
            raise CompileError(lib().zo_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().zo_compiled_free(self._h)
            self._h = None

    @property
    def stats(self):
        s = np.zeros(6, np.int64)
        lib().zo_compiled_stats(self._h, _ptr(s))
        keys = ("input_gates", "cancelled", "output_ops", "max_fused", "dynamic", "n_final_measures")
        return dict(zip(keys, (int(x) for x in s)))

    @property
    def final_measures(self):
        n = self.stats["n_final_measures"]
        pairs = np.zeros(2 * max(n, 1), np.uint32)
        lib().zo_compiled_final_measures(self._h, _ptr(pairs))
        return [(int(pairs[2 * i]), int(pairs[2 * i + 1])) for i in range(n)]

    def __len__(self):
        return lib().zo_compiled_nops(self._h)

    def op(self, i):
        kind, k, is_if = C.c_int32(), C.c_int32(), C.c_int32()
        qs = np.zeros(12, np.uint32)
        data = np.zeros(2 * 4096, np.float64)
        cond = np.zeros(3, np.uint64)
        rc = lib().zo_compiled_op(self._h, i, C.byref(kind), C.byref(k), _ptr(qs), _ptr(data), C.byref(is_if), _ptr(cond))
        if rc < 0:
            raise IndexError(i)
        kd, kk = kind.value, k.value
        d = {"kind": kd, "qubits": [int(x) for x in qs[:kk]], "is_if": bool(is_if.value), "cond": [int(x) for x in cond]}
        if kd == OP_U:
            d["mat"] = data[: 2 * (1 << 2 * kk)].view(np.complex128).reshape(1 << kk, 1 << kk).copy()
        elif kd == OP_D:
            d["diag"] = data[: 2 * (1 << kk)].view(np.complex128).copy()
        return d

    def ops(self):
        return [self.op(i) for i in range(len(self))]


# ---------------------------------------------------------------- state

class State:
    def __init__(self, n: int, prec: int = 64):
        self.n, self.prec = n, prec
        self._h = lib().zo_state_new(n, prec)
        if not self._h:
            raise MemoryError(lib().zo_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().zo_state_free(self._h)
            self._h = None

    def reset_all(self):
        lib().zo_reset_all(self._h)

    def apply_unitary(self, qs, mat):
        qs = np.asarray(qs, np.uint32)
        m = np.ascontiguousarray(mat, np.complex128).reshape(-1).view(np.float64)
        lib().zo_apply_unitary(self._h, _ptr(qs), len(qs), _ptr(m))

    def apply_diagonal(self, qs, d):
        qs = np.asarray(qs, np.uint32)
        m = np.ascontiguousarray(d, np.complex128).view(np.float64)
        lib().zo_apply_diagonal(self._h, _ptr(qs), len(qs), _ptr(m))

    def apply_cx(self, c, t):
        lib().zo_apply_cx(self._h, c, t)

    def prob_one(self, q):
        return lib().zo_prob_one(self._h, q)

    def measure(self, q, u):
        return bool(lib().zo_measure(self._h, q, u))

    def norm_sqr(self):
        return lib().zo_norm_sqr(self._h)

    def statevector(self):
        out = np.empty(2 << self.n, np.float64)
        lib().zo_statevector(self._h, _ptr(out))
        return out.view(np.complex128)

    def raw(self):
        dt = np.float64 if self.prec == 64 else np.float32
        re = np.empty(1 << self.n, dt)
        im = np.empty(1 << self.n, dt)
        lib().zo_state_raw(self._h, _ptr(re), _ptr(im))
        return re, im

    def set_raw(self, re, im):
        dt = np.float64 if self.prec == 64 else np.float32
        re = np.ascontiguousarray(re, dt)
        im = np.ascontiguousarray(im, dt)
        lib().zo_state_set_raw(self._h, _ptr(re), _ptr(im))

    def apply_compiled(self, comp: Compiled, lo: int = 0, hi: int = None):
        lib().zo_apply_compiled(self._h, comp._h, lo, len(comp) if hi is None else hi)

    def sample(self, shots, seed):
        out = np.zeros(max(shots, 1), np.uint64)
        lib().zo_sample(self._h, shots, seed, _ptr(out))
        return out[:shots]

    def chunk_sums(self):
        ln = 1 << self.n
        ch = min(ln, 1 << 16)
        out = np.zeros((ln + ch - 1) // ch, np.float64)
        lib().zo_chunk_sums(self._h, _ptr(out))
        return out


def set_threads(t: int):
    lib().zo_set_threads(t)


def statevector(prog, n, fusion=1, prec=64):
    comp = Compiled(prog, n, fusion=fusion)
    st = State(n, prec)
    st.apply_compiled(comp)
    return st.statevector()


def run(comp: Compiled, shots: int, seed: int, prec: int = 64, noise: dict | None = None):
    """Counts as {clbit word: count}, plus the simulated seconds."""
    cap = max(shots, 1)
    keys = np.zeros(cap, np.uint64)
    counts = np.zeros(cap, np.uint64)
    secs = C.c_double(0)
    nz = None
    if noise:
        nz = Noise(**{k: float(v) for k, v in noise.items()})
    m = lib().zo_run(comp._h, prec, shots, seed, C.byref(nz) if nz is not None else None, _ptr(keys), _ptr(counts), cap, C.byref(secs))
    if m < 0:
        raise RuntimeError(lib().zo_error().decode())
    return {int(k): int(c) for k, c in zip(keys[:m], counts[:m])}, secs.value


def run_shots(comp: Compiled, shots: int, seed: int, prec: int = 64, noise: dict | None = None,
              threads: int = 1) -> np.ndarray:
    """Per-shot clbit words of the per-shot path (src/exec.rs:459-528),
    shot-parallel over `threads` workers with a state each."""
    words = np.zeros(max(shots, 1), np.uint64)
    nz = Noise(**{k: float(v) for k, v in noise.items()}) if noise else None
    rc = lib().zo_run_shots(comp._h, prec, shots, seed, C.byref(nz) if nz is not None else None, threads,
                            _ptr(words))
    if rc < 0:
        raise RuntimeError(lib().zo_error().decode())
    return words[:shots]


# --------------------------------------------------------------- RefSim

def refsim_apply(amps: np.ndarray, n: int, qs, mat):
    qs = np.asarray(qs, np.uint32)
    m = np.ascontiguousarray(mat, np.complex128).reshape(-1).view(np.float64)
    assert amps.dtype == np.complex128 and amps.flags.c_contiguous
    lib().zo_refsim_apply(_ptr(amps), n, _ptr(qs), len(qs), _ptr(m))
