"""Compiler and executor front door (src/compiler.rs, src/exec.rs, src/lib.rs).

`compile` and `run_compiled` call the native host side (zr_*), which drives
the HIP backend through the zb_* boundary.  Counts keys follow the qiskit
convention of format_key (src/exec.rs:235-253).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Optional

import numpy as np

from . import _lib, ir
from ._lib import last_error, lib, ptr
from .backend import HipBackend, MemoryBudget, ZenoError, _check


class InvalidCircuit(ZenoError):
    pass


@dataclasses.dataclass
class NoiseModel:
    """noise::NoiseModel (src/noise.rs:37-46)."""

    depolarizing_1q: float = 0.0
    depolarizing_2q: float = 0.0
    bit_flip: float = 0.0
    phase_flip: float = 0.0
    amplitude_damping: float = 0.0
    readout_flip_0to1: float = 0.0
    readout_flip_1to0: float = 0.0

    def is_trivial(self) -> bool:
        return all(v == 0.0 for v in dataclasses.astuple(self))

    def validate(self):
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if not (0.0 <= v <= 1.0):
                raise ZenoError(f"invalid noise model: {f.name} = {v} is not a probability in [0, 1]")


class _CNoise(C.Structure):
    _fields_ = [(f.name, C.c_double) for f in dataclasses.fields(NoiseModel)]


OP_U, OP_D, OP_CX, OP_MEAS, OP_RESET = 0, 1, 2, 3, 4


class Compiled:
    """compiler::Compiled (src/compiler.rs:80-94)."""

    def __init__(self, prog: np.ndarray, n_qubits: int, cregs=None, fusion: int = 1, diag_max: int = 10):
        prog = np.ascontiguousarray(prog, ir.INSTR_DTYPE)
        cregs = np.asarray([n_qubits] if cregs is None else cregs, np.uint32)
        self.n_qubits = n_qubits
        self.cregs = [int(x) for x in cregs]
        self._h = lib().zr_compile(ptr(prog), len(prog), n_qubits, ptr(cregs), len(cregs), fusion, diag_max)
        if not self._h:
            raise InvalidCircuit(last_error())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().zr_compiled_free(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return lib().zr_compiled_nops(self._h)

    @property
    def stats(self):
        s = np.zeros(6, np.int64)
        lib().zr_compiled_stats(self._h, ptr(s))
        keys = ("input_gates", "cancelled", "output_ops", "max_fused", "dynamic", "n_final_measures")
        return dict(zip(keys, (int(x) for x in s)))

    @property
    def dynamic(self) -> bool:
        return bool(self.stats["dynamic"])

    @property
    def final_measures(self):
        n = self.stats["n_final_measures"]
        pairs = np.zeros(2 * max(n, 1), np.uint32)
        lib().zr_compiled_final_measures(self._h, ptr(pairs))
        return [(int(pairs[2 * i]), int(pairs[2 * i + 1])) for i in range(n)]

    def op(self, i):
        kind, k, is_if = C.c_int32(), C.c_int32(), C.c_int32()
        qs = np.zeros(12, np.uint32)
        data = np.zeros(2 * 4096, np.float64)
        cond = np.zeros(3, np.uint64)
        if lib().zr_compiled_op(self._h, i, C.byref(kind), C.byref(k), ptr(qs), ptr(data), C.byref(is_if), ptr(cond)) < 0:
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


def compile(prog, n_qubits: int, cregs=None, fusion: int = 1, diag_max: int = 10) -> Compiled:
    """compiler::compile (src/compiler.rs:118-200)."""
    return Compiled(prog, n_qubits, cregs, fusion, diag_max)


def format_key(clbits: int, cregs) -> str:
    """exec.rs:235-253: cregs joined by spaces, last-declared leftmost."""
    if not cregs:
        return ""
    parts, off = [], 0
    for size in cregs:
        mask = (1 << 64) - 1 if size >= 64 else (1 << size) - 1
        parts.append(format((clbits >> off) & mask, f"0{size}b") if size else "")
        off += size
    return " ".join(reversed(parts))


class Counts(dict):
    """exec::Counts (src/exec.rs:77-112): bitstring -> count."""

    def get(self, key, default=0):
        return super().get(key, default)

    def total(self) -> int:
        return sum(self.values())

    def probabilities(self):
        t = max(self.total(), 1)
        return sorted(((k, v / t) for k, v in self.items()), key=lambda kv: (-kv[1], kv[0]))


def counts_of(raw: dict, cregs) -> Counts:
    """Counts keyed by format_key (src/exec.rs:235-253) from clbit words."""
    out = Counts()
    for k, v in raw.items():
        key = format_key(k, cregs)
        out[key] = out.get(key) + v
    return out


@dataclasses.dataclass
class RunResult:
    """exec::RunResult (src/exec.rs:115-130).  `counts` (bitstring keys) is
    built from `raw_counts` (clbit words) on first use."""

    raw_counts: dict
    shots: int
    n_qubits: int
    precision: int
    backend: str
    seed: int
    sim_time: float
    statevector: Optional[np.ndarray]
    stats: dict
    mem_bytes: int
    notices: list
    cregs: list = dataclasses.field(default_factory=list)
    _counts: Optional[Counts] = dataclasses.field(default=None, repr=False)

    @property
    def counts(self) -> Counts:
        if self._counts is None:
            self._counts = counts_of(self.raw_counts, self.cregs)
        return self._counts


class _MemPlan(C.Structure):
    """zr_mem_plan: mem::MemPlan plus the fields of Error::Memory."""

    _fields_ = [("n_qubits", C.c_uint32), ("precision", C.c_int32), ("needed_bytes", C.c_uint64),
                ("budget_bytes", C.c_uint64), ("max_qubits", C.c_uint32)]


def _memory_error(p: "_MemPlan") -> MemoryBudget:
    e = MemoryBudget(_lib.last_error())
    e.n_qubits, e.precision, e.needed_bytes = int(p.n_qubits), int(p.precision), int(p.needed_bytes)
    e.budget_bytes, e.max_qubits = int(p.budget_bytes), int(p.max_qubits)
    return e


def state_bytes(n_qubits: int, precision: int) -> int:
    """mem::state_bytes (src/mem.rs:79-81): 8 B/amp in f32, 16 B in f64."""
    p = _MemPlan()
    lib().zr_plan(n_qubits, precision, 0, C.byref(p))
    return int(p.needed_bytes)


def max_qubits(budget: int, precision: int) -> int:
    """mem::max_qubits (src/mem.rs:84-93)."""
    p = _MemPlan()
    lib().zr_plan(0, precision, budget, C.byref(p))
    return int(p.max_qubits)


def budget_bytes(mem_limit: Optional[int] = None, fraction: float = 0.75) -> int:
    """mem::budget_bytes (src/mem.rs:103-109) on HBM: an explicit limit wins,
    otherwise `fraction` of the device's total HBM (`ZENO_MEM_BYTES`
    replaces that total, as it replaces RAM in src/mem.rs:11-17)."""
    return int(lib().zr_budget_bytes(mem_limit or 0, fraction))


def plan(n_qubits: int, precision: int, budget: int):
    """mem::plan (src/mem.rs:114-130): the needed bytes, or MemoryBudget
    carrying Error::Memory's fields."""
    p = _MemPlan()
    rc = lib().zr_plan(n_qubits, precision, budget, C.byref(p))
    if rc == _lib.ZB_ERR_MEMORY:
        raise _memory_error(p)
    _check(rc, "plan")
    return int(p.needed_bytes)


@dataclasses.dataclass
class RunOptions:
    """exec::RunOptions (src/exec.rs:29-72) for the HIP backend."""

    shots: int = 1024
    seed: Optional[int] = None
    precision: Optional[int] = None  # None = auto: f64 if it fits, else f32
    fusion_max: Optional[object] = None  # None = 1 (the CPU path's default); "auto" = autotuned width
    mem_limit: Optional[int] = None
    mem_fraction: float = 0.75  # src/exec.rs:65
    want_statevector: bool = False
    noise: Optional[NoiseModel] = None
    mode: int = HipBackend.TILED  # 0 exact, 1 tiled, 2 tiled with the per-shot loop on the host (zr_run)


def resolve_precision(n_qubits: int, opts: RunOptions):
    """resolve_precision (src/exec.rs:261-304) on the HBM budget, through
    zr_resolve_precision."""
    budget = budget_bytes(opts.mem_limit, opts.mem_fraction)
    prec, p, note = C.c_int(0), _MemPlan(), C.create_string_buffer(256)
    rc = lib().zr_resolve_precision(n_qubits, opts.precision or 0, budget, C.byref(prec), C.byref(p), note, 256)
    if rc == _lib.ZB_ERR_MEMORY:
        raise _memory_error(p)
    _check(rc, "resolve_precision")
    notices = [note.value.decode()] if note.value else []
    return int(prec.value), notices


def run_compiled(c: Compiled, opts: RunOptions) -> RunResult:
    """exec::run_compiled (src/exec.rs:334-549) on the HIP backend."""
    noise = opts.noise if (opts.noise is not None and not opts.noise.is_trivial()) else None
    if noise is not None:
        noise.validate()
    prec, notices = resolve_precision(c.n_qubits, opts)
    seed = opts.seed if opts.seed is not None else int.from_bytes(os.urandom(8), "little")
    if noise is not None and opts.want_statevector:
        raise InvalidCircuit(
            "a noise model is active: the noisy state is a mixture over trajectories and has no single "
            "final state vector (drop the statevector request or the noise model)"
        )
    if c.dynamic and opts.want_statevector:
        raise InvalidCircuit("circuit is dynamic (mid-circuit measure/reset/if): it has no single final state vector")
    statevector = None
    name = "hip-f64" if prec == 64 else "hip-f32"
    if opts.want_statevector:
        import time

        t0 = time.perf_counter()
        with HipBackend(c.n_qubits, prec, HipBackend.TILED if opts.mode == 2 else opts.mode) as be:
            be.apply_compiled(c)
            be.finish()
            statevector = be.statevector()
            raw = {}
            if c.final_measures and opts.shots > 0:
                from collections import Counter

                basis = Counter(be.sample(opts.shots, _splitmix64(seed)).tolist())
                for b, n in basis.items():
                    cl = 0
                    for q, cb in c.final_measures:
                        cl = (cl & ~(1 << cb)) | (((b >> q) & 1) << cb)
                    raw[cl] = raw.get(cl, 0) + n
        secs = time.perf_counter() - t0
    else:
        cap = max(opts.shots, 1)
        keys = np.zeros(cap, np.uint64)
        counts = np.zeros(cap, np.uint64)
        secs_c = C.c_double(0.0)
        cn = _CNoise(**dataclasses.asdict(noise)) if noise is not None else None
        m = lib().zr_run(
            c.handle, prec, opts.shots, seed, C.byref(cn) if cn is not None else None, opts.mode,
            ptr(keys), ptr(counts), cap, C.byref(secs_c),
        )
        _check(m, "run")
        raw = dict(zip(keys[:m].tolist(), counts[:m].tolist()))
        secs = secs_c.value
    if noise is not None:
        notices.append("noise: trajectory sampling — per-shot execution, fusion disabled")
    if not c.final_measures and not c.dynamic and opts.shots > 0:
        # run_per_shot's text on the noisy path (src/exec.rs:476-482), where a
        # statevector cannot be asked for; run_sampled's otherwise (:426-433)
        notices.append("circuit has no measurements: counts are empty" if noise is not None else
                       "circuit has no measurements: counts are empty (use --statevector to inspect the state)")
    return RunResult(
        raw_counts=raw, shots=opts.shots, n_qubits=c.n_qubits, precision=prec, backend=name,
        seed=seed, sim_time=secs, statevector=statevector, stats=c.stats,
        mem_bytes=state_bytes(c.n_qubits, prec), notices=notices, cregs=list(c.cregs),
    )


def trajectories_supported(c: Compiled) -> bool:
    """Whether `run_trajectories` takes the circuit: <= 26 qubits, dense ops
    on <= 6 qubits, diagonals on <= 12."""
    return bool(lib().zr_trajectories_supported(c.handle))


def run_trajectories(c: Compiled, shots: int, seed: int, precision: int = 64, noise: Optional[NoiseModel] = None):
    """Every shot of run_per_shot (src/exec.rs:459-549, src/noise.rs:91-192)
    in one persistent kernel launch.  Returns (clbit word per shot as
    uint64, kernel seconds); shot s equals the CPU per-shot path's shot s."""
    if noise is not None:
        noise.validate()
    out = np.zeros(max(shots, 1), np.uint64)
    secs = C.c_double(0.0)
    cn = _CNoise(**dataclasses.asdict(noise)) if noise is not None and not noise.is_trivial() else None
    rc = lib().zr_run_trajectories(
        c.handle, precision, shots, seed, C.byref(cn) if cn is not None else None, ptr(out), C.byref(secs)
    )
    _check(rc, "run_trajectories")
    return out[:shots], secs.value


def batch_supported(c: Compiled, precision: int, shots: int) -> bool:
    """Whether `run_batch` (and the server's batcher) takes the circuit:
    static, <= 13 qubits in f64 / 14 in f32, dense ops on <= 3 / 4 qubits,
    <= 2^17 shots."""
    return bool(lib().zr_batch_supported(c.handle, precision, shots))


def run_batch(runs):
    """Many small static circuits in one persistent launch: `runs` is a list
    of (Compiled, precision, shots, seed); returns ([raw counts per run],
    seconds).  Each run's counts equal run_compiled's for the same inputs."""
    runs = list(runs)
    m = len(runs)
    hs = (C.c_void_p * max(m, 1))(*[r[0].handle for r in runs])
    prec = np.asarray([r[1] for r in runs] or [64], np.int32)
    shots = np.asarray([r[2] for r in runs] or [0], np.uint64)
    seeds = np.asarray([r[3] for r in runs] or [0], np.uint64)
    words = np.zeros(max(int(shots[:m].sum()) if m else 0, 1), np.uint64)
    secs = C.c_double(0.0)
    _check(lib().zr_run_batch(hs, ptr(prec), ptr(shots), ptr(seeds), m, ptr(words), C.byref(secs)), "run_batch")
    out, at = [], 0
    for c, _p, s, _sd in runs:
        w = words[at:at + s] if c.final_measures else words[:0]
        at += s
        v, k = np.unique(w, return_counts=True)
        out.append({int(a): int(b) for a, b in zip(v, k)})
    return out, secs.value


def _splitmix64(x: int) -> int:
    M = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def compile_options(opts: RunOptions, n_qubits: Optional[int] = None) -> int:
    """lib.rs:224-240: noisy runs compile with fusion 0; the default dense
    fusion width is 1 (the reference CPU default); "auto" takes the width
    the autotuner measured fastest for this size (1 if none is cached)."""
    if opts.noise is not None and not opts.noise.is_trivial():
        return 0
    if opts.fusion_max == "auto":
        from .tune import tuned_fusion

        k = tuned_fusion(n_qubits, opts.precision or 64) if n_qubits is not None else None
        return 1 if k is None else k
    return 1 if opts.fusion_max is None else opts.fusion_max


def run_program(prog, n_qubits: int, cregs=None, opts: Optional[RunOptions] = None) -> RunResult:
    """lib.rs:242-250 run_program."""
    opts = opts or RunOptions()
    return run_compiled(compile(prog, n_qubits, cregs, fusion=compile_options(opts, n_qubits)), opts)


class Simulator:
    """Builder-style front door (src/lib.rs:139-222)."""

    def __init__(self):
        self.opts = RunOptions()

    def shots(self, s):
        self.opts.shots = s
        return self

    def seed(self, s):
        self.opts.seed = s
        return self

    def precision(self, p):
        self.opts.precision = p
        return self

    def fusion(self, k):
        self.opts.fusion_max = k
        return self

    def mem_limit(self, b):
        self.opts.mem_limit = b
        return self

    def noise(self, m: NoiseModel):
        self.opts.noise = m
        return self

    def mode(self, m):
        self.opts.mode = m
        return self

    def run(self, circuit) -> RunResult:
        prog, n, cregs = circuit.to_program()
        return run_program(prog, n, cregs, self.opts)

    def statevector(self, circuit) -> np.ndarray:
        o = dataclasses.replace(self.opts, want_statevector=True, shots=0)
        prog, n, cregs = circuit.to_program()
        return run_program(prog, n, cregs, o).statevector
