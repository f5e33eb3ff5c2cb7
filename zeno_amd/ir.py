"""Program records (ir::Instr, src/ir.rs:36-83) in the C layout zr_instr.

Qubits are little-endian (bit q of a basis index is qubit q); a gate's
qubit list is in argument order (controls first).
"""
from __future__ import annotations

import numpy as np

INSTR_DTYPE = np.dtype(
    [
        ("name", "S8"),
        ("kind", "<i4"),  # 0 gate, 1 measure, 2 reset, 3 barrier
        ("nq", "<i4"),
        ("np", "<i4"),
        ("q", "<u4", (3,)),
        ("clbit", "<u4"),
        ("cond_creg", "<i4"),  # -1: unconditional; else `if (creg == cond_value)`
        ("cond_value", "<u8"),
        ("p", "<f8", (3,)),
    ],
    align=True,
)
assert INSTR_DTYPE.itemsize == 72


def _row(kind, cond):
    r = np.zeros(1, INSTR_DTYPE)
    r["kind"] = kind
    r["cond_creg"] = -1 if cond is None else cond[0]
    r["cond_value"] = 0 if cond is None else cond[1]
    return r


def instr(name, qubits, params=(), cond=None):
    if len(name) > 7:
        raise ValueError(f"gate name too long: {name!r}")
    r = _row(0, cond)
    r["name"] = name.encode()
    r["nq"] = len(qubits)
    r["np"] = len(params)
    r["q"][0, : len(qubits)] = qubits
    r["p"][0, : len(params)] = params
    return r


def measure(q, clbit, cond=None):
    r = _row(1, cond)
    r["q"][0, 0] = q
    r["clbit"] = clbit
    return r


def reset(q, cond=None):
    r = _row(2, cond)
    r["q"][0, 0] = q
    return r


def barrier():
    return _row(3, None)


def program(rows) -> np.ndarray:
    rows = [r for r in rows if len(r)]
    return np.concatenate(rows).astype(INSTR_DTYPE) if rows else np.zeros(0, INSTR_DTYPE)


def measure_all(n):
    return program([measure(q, q) for q in range(n)])
