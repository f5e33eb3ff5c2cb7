"""Circuit builder and seeded generators (src/circuit.rs:18-174).

The generators run in the native library (zr_gen_*), which follows the
reference draw order: Xoshiro256PlusPlus::seed_from_u64(seed), then per
layer and qubit theta, phi, lambda = u * tau.
"""
from __future__ import annotations

import numpy as np

from . import ir
from ._lib import lib, ptr


class Circuit:
    """Single `q` register, single `c` register builder (circuit.rs:18-134)."""

    def __init__(self, n_qubits: int):
        self.n_qubits = n_qubits
        self._rows = []

    def gate(self, name, qubits, params=()):
        self._rows.append(ir.instr(name, list(qubits), list(params)))
        return self

    def extend(self, rows: np.ndarray):
        self._rows.append(np.asarray(rows, ir.INSTR_DTYPE))
        return self

    # gate0!, gate0_2q!, gate1p! (circuit.rs:68-70) and the rest
    def __getattr__(self, name):
        one = ("x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg")
        two = ("cx", "cy", "cz", "ch", "swap")
        par = ("rx", "ry", "rz", "p")
        if name in one:
            return lambda q: self.gate(name, [q])
        if name in two:
            return lambda a, b: self.gate(name, [a, b])
        if name in par:
            return lambda q, theta: self.gate(name, [q], [theta])
        raise AttributeError(name)

    def u(self, q, theta, phi, lam):
        return self.gate("u3", [q], [theta, phi, lam])

    def cp(self, c, t, lam):
        return self.gate("cp", [c, t], [lam])

    def crz(self, c, t, theta):
        return self.gate("crz", [c, t], [theta])

    def rzz(self, a, b, theta):
        return self.gate("rzz", [a, b], [theta])

    def ccx(self, c0, c1, t):
        return self.gate("ccx", [c0, c1, t])

    def cswap(self, c, a, b):
        return self.gate("cswap", [c, a, b])

    def measure(self, q, clbit):
        self._rows.append(ir.measure(q, clbit))
        return self

    def measure_all(self):
        for q in range(self.n_qubits):
            self.measure(q, q)
        return self

    def reset(self, q):
        self._rows.append(ir.reset(q))
        return self

    def barrier(self):
        self._rows.append(ir.barrier())
        return self

    def to_program(self):
        """(instructions, n_qubits, creg sizes) — circuit.rs:121-133."""
        return ir.program(self._rows), self.n_qubits, [self.n_qubits]

    def __len__(self):
        return int(sum(len(r) for r in self._rows))


def _gen(fn, *args):
    cnt = fn(*args, None, 0)
    out = np.zeros(max(cnt, 1), ir.INSTR_DTYPE)
    fn(*args, ptr(out), cnt)
    return out[:cnt]


def _wrap(n, rows):
    c = Circuit(n)
    c.extend(rows)
    return c


def random_circuit(n: int, depth: int, seed: int) -> Circuit:
    """circuit.rs:152-174: U3 layer + shifting CX brickwork per layer."""
    return _wrap(n, _gen(lib().zr_gen_random, n, depth, seed))


def qft(n: int) -> Circuit:
    """circuit.rs:138-148 (no final swaps)."""
    return _wrap(n, _gen(lib().zr_gen_qft, n))


def u3_qft(n: int, seed: int) -> Circuit:
    """A seeded U3 layer, then qft(n)."""
    return _wrap(n, _gen(lib().zr_gen_u3_qft, n, seed))


def matching_circuit(n: int, depth: int, seed: int, rzz_period: int = 0) -> Circuit:
    """U3 layer + CX on a seeded random perfect matching per layer, with an
    RZZ ladder on all neighbour pairs every `rzz_period`-th layer."""
    return _wrap(n, _gen(lib().zr_gen_matching, n, depth, seed, rzz_period))
