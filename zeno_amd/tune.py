"""Autotuner for the tiled HIP path, cached per (qubits, precision).

The reference hard-codes its fusion width per backend (src/lib.rs:224-240)
and its bench can only compare fusion widths side by side
(src/main.rs:1139-1149).  Here the launch shape of the tiled kernel is
chosen by measurement on the device at hand:

* knobs that never change results — tile width (12/13 bits), register
  slots per stage (3/4), register prefetch, window grid, CX deferral, qubit
  layout — are searched by coordinate descent and installed as process-wide
  defaults (zb_tune_default), so every backend of that shape created afterwards
  (run_compiled, the server, HipBackend) uses them;
* the dense fusion width K changes the op stream and hence the rounding
  (still within the reference's tolerances, but no longer bit-identical to
  the CPU path at its default K), so it is measured and reported as
  `fusion` but applied only where a caller opts in (RunOptions.fusion_max
  = "auto").

The probe workload is the benchmark's seeded random-matching circuit at the
requested size, shortened to a few layers; each candidate runs once to warm
up and is then timed from the HIP events the library records around its
kernels.  Results are cached as JSON (default
~/.cache/zeno_amd/tune-<arch>.json, or $ZENO_TUNE_CACHE).
"""
from __future__ import annotations

import json
import os
from typing import Dict, Optional

from ._lib import lib
from .backend import HipBackend, _check

# Searched in this order (coordinate descent from the library defaults).
SEARCH = {
    "tile_bits": [12, 13],
    "slots": [0, 4],
    "prefetch": [1, 0],
    "win_grid": [16384, 4096, 65536],
    "defer_cx": [1, 0],
    "relabel": [1, 0],
    "layout": [-1, 0],
}
FUSION = [1, 2, 3, 4]
ARCH = "gfx950"


def cache_path() -> str:
    env = os.environ.get("ZENO_TUNE_CACHE")
    if env:
        return env
    return os.path.join(os.path.expanduser("~"), ".cache", "zeno_amd", f"tune-{ARCH}.json")


def _key(n: int, precision: int) -> str:
    return f"{n}:{precision}"


def load_cache(path: Optional[str] = None) -> Dict[str, dict]:
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(entries: Dict[str, dict], path: Optional[str] = None) -> str:
    path = path or cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(entries, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
    return path


def install(n: int, precision: int, knobs: Dict[str, int]):
    """Make `knobs` the defaults of every later backend of this shape."""
    for k, v in knobs.items():
        if k in SEARCH:
            _check(lib().zb_tune_default(n, precision, k.encode(), int(v)), "tune_default")


def uninstall(n: int, precision: int):
    for k in SEARCH:
        lib().zb_tune_default(n, precision, k.encode(), -(1 << 63))


def load(path: Optional[str] = None) -> int:
    """Install every cached choice; returns how many shapes were loaded."""
    entries = load_cache(path)
    for key, e in entries.items():
        n, prec = (int(x) for x in key.split(":"))
        install(n, prec, e.get("knobs", {}))
    return len(entries)


def tuned_fusion(n: int, precision: int, path: Optional[str] = None) -> Optional[int]:
    e = load_cache(path).get(_key(n, precision))
    return int(e["fusion"]) if e and "fusion" in e else None


def _probe_program(n: int, depth: int, seed: int):
    from .circuit import matching_circuit

    return matching_circuit(n, depth, seed, rzz_period=5).to_program()[0]


def _time(be: HipBackend, comp, reps: int) -> float:
    """Kernel milliseconds of one run of `comp` (best of `reps`)."""
    best = float("inf")
    be.apply_compiled(comp)
    be.finish()
    for _ in range(reps):
        be.reset_all()
        be.profile_begin()
        be.apply_compiled(comp)
        recs = be.profile_end()
        best = min(best, sum(ms for _, ms, _ in recs))
    return best


def autotune(n: int, precision: int = 64, depth: int = 6, seed: int = 7, reps: int = 2,
             search: Optional[Dict[str, list]] = None, fusions=None, path: Optional[str] = None,
             save: bool = True, log=None) -> dict:
    """Measure, cache and install the fastest launch shape for this size.

    Returns {"knobs": {...}, "fusion": K, "ms": {...}} where "ms" holds the
    probe's kernel time for the chosen knobs and for each fusion width."""
    from .exec import compile as zcompile

    search = SEARCH if search is None else search
    fusions = FUSION if fusions is None else fusions
    prog = _probe_program(n, depth, seed)
    base = zcompile(prog, n, fusion=1)
    uninstall(n, precision)
    knobs: Dict[str, int] = {}

    def measure(cand: Dict[str, int], comp) -> float:
        with HipBackend(n, precision, HipBackend.TILED) as be:
            for k, v in cand.items():
                be.tune(k, v)
            return _time(be, comp, reps)

    best_ms = measure(knobs, base)
    if n >= 12:  # below 12 qubits every op runs the per-op kernels
        for key, values in search.items():
            for v in values:
                cand = dict(knobs, **{key: v})
                if cand == knobs:
                    continue
                ms = measure(cand, base)
                if log:
                    log(f"autotune {n}q f{precision} {cand}: {ms:.3f} ms")
                if ms < best_ms * 0.98:  # keep a change only for a clear gain
                    best_ms, knobs = ms, cand
    fusion_ms = {}
    for k in fusions:
        fusion_ms[str(k)] = measure(knobs, zcompile(prog, n, fusion=k))
    best_k = min(fusion_ms, key=fusion_ms.get)
    entry = {"knobs": knobs, "fusion": int(best_k), "ms": {"knobs": best_ms, "fusion": fusion_ms},
             "probe": f"matching circuit {n}q depth {depth} seed {seed}"}
    install(n, precision, knobs)
    if save:
        entries = load_cache(path)
        entries[_key(n, precision)] = entry
        save_cache(entries, path)
    return entry
