"""OpenQASM 2.0 front end (src/qasm.rs parse_str / parse_file; the subset
and rules of docs/QASM.md), parsed natively by zr_parse_qasm into the
program rows the compiler takes."""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import List, Optional

import numpy as np

from . import ir
from ._lib import last_error, lib, ptr


class QasmError(ValueError):
    """Parse error; the message starts with "line L:C:"."""


@dataclasses.dataclass
class QasmProgram:
    rows: np.ndarray  # ir.INSTR_DTYPE
    n_qubits: int
    cregs: List[int]


def parse(src: str) -> QasmProgram:
    data = src.encode()
    nq, ncr = C.c_uint32(), C.c_uint32()
    cap, ccap = 256, 64
    while True:
        rows = np.zeros(cap, ir.INSTR_DTYPE)
        cregs = np.zeros(ccap, np.uint32)
        n = lib().zr_parse_qasm(data, ptr(rows), cap, C.byref(nq), ptr(cregs), C.byref(ncr), ccap)
        if n < 0:
            raise QasmError(last_error())
        if n <= cap and ncr.value <= ccap:
            return QasmProgram(rows[:n].copy(), int(nq.value), [int(x) for x in cregs[: ncr.value]])
        cap, ccap = max(cap, int(n)), max(ccap, int(ncr.value))


def parse_file(path: str) -> QasmProgram:
    with open(path) as f:
        return parse(f.read())


def run_qasm(src: str, opts=None):
    """lib.rs run_program on a QASM source."""
    from .exec import RunOptions, run_program

    p = parse(src)
    return run_program(p.rows, p.n_qubits, p.cregs or [], opts or RunOptions())
