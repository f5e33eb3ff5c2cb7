"""Independent gate matrices and dense simulators for the tests.

Hand-coded standard definitions (qiskit U(θ,φ,λ) convention, controls
first = low bits of the local index), restated in numpy from the
reference's own independent oracle, tests/reference.rs:157-283.  None of
this calls the oracle's or the product's gate library.
"""
import numpy as np

PI = np.pi
SQ = 1.0 / np.sqrt(2.0)


def eip(t):
    return complex(np.cos(t), np.sin(t))


def dmat(d):
    return np.diag(np.asarray(d, np.complex128))


def ctrl1(u):
    u = np.asarray(u, np.complex128).reshape(2, 2)
    m = np.zeros((4, 4), np.complex128)
    m[0, 0] = 1
    m[2, 2] = 1
    for to in range(2):
        for ti in range(2):
            m[1 + 2 * to, 1 + 2 * ti] = u[to, ti]
    return m


def u3m(th, ph, la):
    ct, st = np.cos(th / 2), np.sin(th / 2)
    return np.array([[ct, -eip(la) * st], [eip(ph) * st, eip(ph + la) * ct]], np.complex128)


def rxm(th):
    ct, st = np.cos(th / 2), np.sin(th / 2)
    return np.array([[ct, -1j * st], [-1j * st, ct]], np.complex128)


def rym(th):
    ct, st = np.cos(th / 2), np.sin(th / 2)
    return np.array([[ct, -st], [st, ct]], np.complex128)


def rzm(th):
    return dmat([eip(-th / 2), eip(th / 2)])


def ref_gate(name, p=()):
    """Dense matrix in argument-order convention (tests/reference.rs:188-283)."""
    H = np.array([[SQ, SQ], [SQ, -SQ]], np.complex128)
    Y = np.array([[0, -1j], [1j, 0]], np.complex128)
    if name in ("id", "u0"):
        return np.eye(2, dtype=np.complex128)
    if name == "x":
        return np.array([[0, 1], [1, 0]], np.complex128)
    if name == "y":
        return Y
    if name == "z":
        return dmat([1, -1])
    if name == "h":
        return H
    if name == "s":
        return dmat([1, 1j])
    if name == "sdg":
        return dmat([1, -1j])
    if name == "t":
        return dmat([1, eip(PI / 4)])
    if name == "tdg":
        return dmat([1, eip(-PI / 4)])
    if name == "sx":
        return np.array([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]])
    if name == "sxdg":
        return np.array([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]])
    if name == "rx":
        return rxm(p[0])
    if name == "ry":
        return rym(p[0])
    if name == "rz":
        return rzm(p[0])
    if name in ("u1", "p"):
        return dmat([1, eip(p[0])])
    if name == "u2":
        return u3m(PI / 2, p[0], p[1])
    if name in ("u3", "u"):
        return u3m(p[0], p[1], p[2])
    if name == "cx":
        m = np.zeros((4, 4), np.complex128)
        for col in range(4):
            cb, tb = col & 1, (col >> 1) & 1
            m[cb | ((tb ^ cb) << 1), col] = 1
        return m
    if name == "cy":
        return ctrl1(Y)
    if name == "cz":
        return dmat([1, 1, 1, -1])
    if name == "ch":
        return ctrl1(H)
    if name == "swap":
        m = np.zeros((4, 4), np.complex128)
        for col in range(4):
            a, b = col & 1, (col >> 1) & 1
            m[b | (a << 1), col] = 1
        return m
    if name in ("cp", "cu1"):
        return dmat([1, 1, 1, eip(p[0])])
    if name == "crx":
        return ctrl1(rxm(p[0]))
    if name == "cry":
        return ctrl1(rym(p[0]))
    if name == "crz":
        return ctrl1(rzm(p[0]))
    if name == "cu3":
        return ctrl1(u3m(p[0], p[1], p[2]))
    if name == "rxx":
        ct, st = np.cos(p[0] / 2), np.sin(p[0] / 2)
        m = np.zeros((4, 4), np.complex128)
        for col in range(4):
            m[col, col] = ct
            m[col ^ 3, col] = -1j * st
        return m
    if name == "rzz":
        return dmat([eip(p[0] / 2 if ((m & 1) ^ ((m >> 1) & 1)) else -p[0] / 2) for m in range(4)])
    if name == "ccx":
        m = np.zeros((8, 8), np.complex128)
        for col in range(8):
            m[col ^ 4 if col & 3 == 3 else col, col] = 1
        return m
    if name == "cswap":
        m = np.zeros((8, 8), np.complex128)
        for col in range(8):
            a, b = (col >> 1) & 1, (col >> 2) & 1
            m[(1 | (b << 1) | (a << 2)) if col & 1 else col, col] = 1
        return m
    raise KeyError(name)


NATIVE = [
    ("id", 1, 0), ("u0", 1, 0), ("x", 1, 0), ("y", 1, 0), ("z", 1, 0), ("h", 1, 0),
    ("s", 1, 0), ("sdg", 1, 0), ("t", 1, 0), ("tdg", 1, 0), ("sx", 1, 0), ("sxdg", 1, 0),
    ("rx", 1, 1), ("ry", 1, 1), ("rz", 1, 1), ("u1", 1, 1), ("p", 1, 1), ("u2", 1, 2),
    ("u3", 1, 3), ("u", 1, 3), ("cx", 2, 0), ("cy", 2, 0), ("cz", 2, 0), ("ch", 2, 0),
    ("swap", 2, 0), ("cp", 2, 1), ("cu1", 2, 1), ("crx", 2, 1), ("cry", 2, 1), ("crz", 2, 1),
    ("rxx", 2, 1), ("rzz", 2, 1), ("cu3", 2, 3), ("ccx", 3, 0), ("cswap", 3, 0),
]


class TestRng:
    """splitmix64 test RNG of tests/reference.rs:36-64."""

    __test__ = False

    M = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.M

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.M
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) / float(1 << 53)

    def angle(self):
        return (self.uniform() * 2.0 - 1.0) * PI

    def below(self, n):
        return self.next_u64() % n


def kron_apply(psi, n, qs, mat):
    """Dense application by tensor contraction (numpy), an implementation
    independent of RefSim: axis for qubit q is n-1-q in a C-order reshape."""
    k = len(qs)
    t = psi.reshape([2] * n)
    m = np.asarray(mat, np.complex128).reshape([2] * (2 * k))
    # local index bit b <-> qs[b]; in the reshaped matrix the row bits are
    # ordered most-significant first, i.e. axis j <-> bit k-1-j.
    row_axes = list(range(k))
    col_axes = list(range(k, 2 * k))
    tgt_axes = [n - 1 - qs[k - 1 - j] for j in range(k)]
    out = np.tensordot(m, t, axes=(col_axes, tgt_axes))
    # out axes: row axes (k) then remaining state axes in order
    rest = [a for a in range(n) if a not in tgt_axes]
    order = [None] * n
    for j, a in enumerate(tgt_axes):
        order[a] = j
    for i, a in enumerate(rest):
        order[a] = k + i
    return np.transpose(out, order).reshape(-1).copy()


def ref_of_program(prog, n):
    """Replay an instruction array (oracle INSTR_DTYPE layout) through the
    numpy tensor-contraction simulator with the hand-coded matrices."""
    psi = np.zeros(1 << n, np.complex128)
    psi[0] = 1
    for r in prog:
        if r["kind"] == 3:
            continue
        assert r["kind"] == 0
        name = r["name"].decode()
        qs = [int(x) for x in r["q"][: r["nq"]]]
        psi = kron_apply(psi, n, qs, ref_gate(name, list(r["p"][: r["np"]])))
    return psi
