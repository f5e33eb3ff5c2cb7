"""Serving loop: continuous batching of circuit requests on one GPU.

The reference runs one circuit per invocation (src/lib.rs:242-250) with
mem::plan admission against RAM (src/mem.rs:79-130).  `Server` keeps one
HBM pool and a set of worker slots (one HIP stream each) in the native
library; requests are admitted first-fit in arrival order as soon as a slot
and pool space are free, and run concurrently with the ones already on the
GPU.  Results are the same as `run_compiled` on the same inputs: the same
kernels run, only placed in the pool and overlapped with other requests.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Optional

import numpy as np

from ._lib import lib, ptr
from .backend import HipBackend, _check
from .exec import Compiled, Counts, NoiseModel, _CNoise, format_key


@dataclasses.dataclass
class Served:
    """One completed request: counts as run_compiled gives them, plus the
    server-side times (seconds since the server started)."""

    id: int
    counts: Counts
    raw_counts: dict
    t_submit: float
    t_start: float
    t_end: float

    @property
    def latency(self) -> float:
        return self.t_end - self.t_submit

    @property
    def queued(self) -> float:
        return self.t_start - self.t_submit


class Server:
    """`batch`: small static requests (see `batch_supported`) run together,
    up to this many per persistent launch, beside the slots; 0 sends every
    request to a slot (with slots=1: one at a time)."""

    def __init__(self, pool_bytes: int = 0, slots: int = 4, mode: int = HipBackend.TILED,
                 batch: Optional[int] = None):
        h = C.c_void_p()
        _check(lib().zs_create(pool_bytes, slots, mode, C.byref(h)), "serve")
        self._h = h
        self._cregs = {}
        if batch is not None:
            _check(lib().zs_set_batch(self._h, batch), "serve batch")

    def close(self):
        if getattr(self, "_h", None):
            lib().zs_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def submit(self, c: Compiled, precision: int = 64, shots: int = 1024, seed: Optional[int] = None,
               noise: Optional[NoiseModel] = None) -> int:
        """Queue a run; returns the request id.  The server copies the
        compiled circuit, so `c` may be dropped afterwards."""
        if seed is None:
            seed = int.from_bytes(os.urandom(8), "little")
        cn = None
        if noise is not None and not noise.is_trivial():
            noise.validate()
            cn = _CNoise(**dataclasses.asdict(noise))
        rid = lib().zs_submit(self._h, c.handle, precision, shots, seed, C.byref(cn) if cn is not None else None)
        _check(rid, "submit")
        self._cregs[rid] = (list(c.cregs), max(shots, 1))
        return int(rid)

    def wait(self, rid: int) -> Served:
        cregs, cap = self._cregs.pop(rid)
        keys = np.zeros(cap, np.uint64)
        counts = np.zeros(cap, np.uint64)
        times = np.zeros(3, np.float64)
        m = lib().zs_wait(self._h, rid, ptr(keys), ptr(counts), cap, ptr(times))
        _check(m, "wait")
        raw = {int(k): int(v) for k, v in zip(keys[:m], counts[:m])}
        out = Counts()
        for k, v in raw.items():
            key = format_key(k, cregs)
            out[key] = out.get(key) + v
        return Served(rid, out, raw, float(times[0]), float(times[1]), float(times[2]))

    def status(self) -> dict:
        s = np.zeros(4, np.int64)
        lib().zs_status(self._h, ptr(s))
        return {"queued": int(s[0]), "running": int(s[1]), "largest_free": int(s[2]), "pool_bytes": int(s[3])}

    def counters(self) -> dict:
        s = np.zeros(4, np.int64)
        lib().zs_counters(self._h, ptr(s))
        return {"batches": int(s[0]), "batched": int(s[1]), "graph_replays": int(s[2]), "max_batch": int(s[3])}
