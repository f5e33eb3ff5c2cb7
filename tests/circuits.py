"""Seeded circuit fixtures shared by the tests (oracle INSTR_DTYPE rows).

random_mixed restates the adversarial pool generator of
tests/reference.rs:361-440 (splitmix64 TestRng, pool filtered by arity,
distinct qubits, angles in (-pi, pi], optional barriers with prob 1/8).
"""
import numpy as np

from refgates import TestRng

POOL = [
    ("x", 1, 0), ("y", 1, 0), ("z", 1, 0), ("h", 1, 0), ("s", 1, 0), ("sdg", 1, 0),
    ("t", 1, 0), ("tdg", 1, 0), ("sx", 1, 0), ("sxdg", 1, 0), ("rx", 1, 1), ("ry", 1, 1),
    ("rz", 1, 1), ("p", 1, 1), ("u1", 1, 1), ("u2", 1, 2), ("u3", 1, 3), ("u", 1, 3),
    ("cx", 2, 0), ("cy", 2, 0), ("cz", 2, 0), ("ch", 2, 0), ("swap", 2, 0), ("cp", 2, 1),
    ("cu1", 2, 1), ("crx", 2, 1), ("cry", 2, 1), ("crz", 2, 1), ("rxx", 2, 1),
    ("rzz", 2, 1), ("cu3", 2, 3), ("ccx", 3, 0), ("cswap", 3, 0),
]
DIAG_POOL = [
    ("z", 1, 0), ("s", 1, 0), ("sdg", 1, 0), ("t", 1, 0), ("tdg", 1, 0), ("rz", 1, 1),
    ("p", 1, 1), ("cz", 2, 0), ("cp", 2, 1), ("crz", 2, 1), ("rzz", 2, 1),
]


def pick_distinct(rng, n, k):
    qs = []
    while len(qs) < k:
        q = rng.below(n)
        if q not in qs:
            qs.append(q)
    return qs


def random_gates(mod, n, count, seed, pool=POOL, barriers=True):
    """Rows built with `mod.instr` / `mod.barrier` (mod = oracle or the
    product's program module; both share the 72-byte record layout)."""
    rng = TestRng(seed)
    pool = [g for g in pool if g[1] <= n]
    rows = []
    for _ in range(count):
        name, arity, npar = pool[rng.below(len(pool))]
        qs = pick_distinct(rng, n, arity)
        params = [rng.angle() for _ in range(npar)]
        rows.append(mod.instr(name, qs, params))
        if barriers and rng.below(8) == 0:
            rows.append(mod.barrier())
    return rows


def random_mixed(mod, n, count, seed):
    return mod.program(random_gates(mod, n, count, seed, POOL, True))


def diag_ladder(mod, n=6):
    """tests/reference.rs:652-669."""
    rows = [mod.instr("h", [q]) for q in range(n)]
    rows += random_gates(mod, n, 40, 0xD1A60001, DIAG_POOL, False)
    rows += [
        mod.instr("cp", [5, 2], [0.4]),
        mod.instr("crz", [4, 1], [0.8]),
        mod.instr("rzz", [3, 0], [1.1]),
        mod.instr("cu1", [4, 0], [0.6]),
        mod.instr("crz", [1, 5], [-0.9]),
    ]
    return mod.program(rows)


def disjoint_blocks(mod):
    """tests/reference.rs:674-715 (n = 8)."""
    rows = []
    for rnd in range(4):
        th = 0.3 + rnd * 0.17
        rows += [
            mod.instr("ccx", [0, 1, 2]), mod.instr("u3", [3], [th, 0.2, 0.9]),
            mod.instr("cx", [2, 3]), mod.instr("h", [0]), mod.instr("swap", [1, 3]),
            mod.instr("ccx", [4, 5, 6]), mod.instr("u3", [7], [th + 0.05, 1.1, 0.4]),
            mod.instr("cx", [6, 7]), mod.instr("h", [4]), mod.instr("swap", [5, 7]),
        ]
    return mod.program(rows)


def max_delta(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) if len(a) else 0.0
