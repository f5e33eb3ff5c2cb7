"""OpenQASM 2.0 front end (src/qasm.rs, docs/QASM.md) on the CPU: parsed
rows against hand-built programs, expression semantics, broadcasting,
user gates, and error positions."""
import math

import numpy as np
import pytest

import zeno_amd as z
from zeno_amd.qasm import QasmError, parse

HDR = 'OPENQASM 2.0;\ninclude "qelib1.inc";\n'


def rows_of(p):
    out = []
    for r in p.rows:
        kind = int(r["kind"])
        if kind == 0:
            out.append(("g", r["name"].decode(), [int(x) for x in r["q"][: r["nq"]]],
                        [float(x) for x in r["p"][: r["np"]]], int(r["cond_creg"]), int(r["cond_value"])))
        elif kind == 1:
            out.append(("m", int(r["q"][0]), int(r["clbit"]), int(r["cond_creg"]), int(r["cond_value"])))
        elif kind == 2:
            out.append(("r", int(r["q"][0]), int(r["cond_creg"])))
        else:
            out.append(("b",))
    return out


def test_bell_and_registers():
    p = parse(HDR + "qreg q[2];\ncreg c[2];\nh q[0];\ncx q[0],q[1];\nmeasure q -> c;\n")
    assert p.n_qubits == 2 and p.cregs == [2]
    assert rows_of(p) == [("g", "h", [0], [], -1, 0), ("g", "cx", [0, 1], [], -1, 0),
                          ("m", 0, 0, -1, 0), ("m", 1, 1, -1, 0)]


def test_register_offsets_and_builtins():
    src = HDR + "qreg a[2];\nqreg b[3];\ncreg x[1];\ncreg y[2];\nU(0.1,0.2,0.3) b[1];\nCX a[1],b[2];\nu(1,2,3) a[0];\nmeasure b[2] -> y[1];\n"
    p = parse(src)
    assert p.n_qubits == 5 and p.cregs == [1, 2]
    assert rows_of(p) == [("g", "u3", [3], [0.1, 0.2, 0.3], -1, 0), ("g", "cx", [1, 4], [], -1, 0),
                          ("g", "u", [0], [1.0, 2.0, 3.0], -1, 0), ("m", 4, 2, -1, 0)]


def test_broadcasting():
    p = parse(HDR + "qreg q[3];\nqreg r[3];\nh q;\ncx q, r;\ncx q[0], r;\nreset r;\n")
    got = rows_of(p)
    assert got[:3] == [("g", "h", [i], [], -1, 0) for i in range(3)]
    assert got[3:6] == [("g", "cx", [i, 3 + i], [], -1, 0) for i in range(3)]
    assert got[6:9] == [("g", "cx", [0, 3 + i], [], -1, 0) for i in range(3)]
    assert got[9:] == [("r", 3 + i, -1) for i in range(3)]


def test_expressions():
    src = HDR + "qreg q[1];\nrz(pi/2) q[0];\nrz(2^3^2) q[0];\nrz(-2^2) q[0];\nrz(2^-3) q[0];\n" \
                "u3(sin(0.3)+cos(0.2)*2, exp(1)-ln(2), sqrt(2)/tan(0.4)) q[0];\nrx(.5e1 - 1.5E0) q[0];\n"
    got = rows_of(parse(src))
    assert got[0][3] == [math.pi / 2]
    assert got[1][3] == [512.0]
    assert got[2][3] == [-4.0]
    assert got[3][3] == [0.125]
    assert got[4][3] == [math.sin(0.3) + math.cos(0.2) * 2, math.exp(1) - math.log(2), math.sqrt(2) / math.tan(0.4)]
    assert got[5][3] == [3.5]


def test_user_gates_expand_inline_with_parameters():
    src = HDR + """qreg q[3];
gate rot(a, b) x { rz(a) x; ry(b/2) x; }
gate pair(t) a, b { rot(t, 2*t) a; cx a, b; barrier a, b; rot(-t, pi) b; }
pair(0.25) q[2], q[0];
"""
    assert rows_of(parse(src)) == [
        ("g", "rz", [2], [0.25], -1, 0), ("g", "ry", [2], [0.25], -1, 0), ("g", "cx", [2, 0], [], -1, 0),
        ("b",), ("g", "rz", [0], [-0.25], -1, 0), ("g", "ry", [0], [math.pi / 2], -1, 0)]


def test_if_and_measure_forms():
    src = HDR + "qreg q[2];\ncreg c0[1];\ncreg c1[2];\nmeasure q[0] -> c0[0];\nif (c0 == 1) x q;\n" \
                "if (c1 == 3) measure q[1] -> c1[0];\nif(c0==0) reset q[1];\n"
    got = rows_of(parse(src))
    assert got == [("m", 0, 0, -1, 0), ("g", "x", [0], [], 0, 1), ("g", "x", [1], [], 0, 1),
                   ("m", 1, 1, 1, 3), ("r", 1, 0)]


def test_parsed_program_compiles_like_the_builder(orc):
    """The QASM route and the oracle's own builders give the same compiled
    ops and the same counts on the CPU oracle."""
    src = HDR + "qreg q[3];\ncreg c[3];\nh q[0];\ncx q[0],q[1];\ncx q[1],q[2];\nrzz(0.3) q[0],q[2];\nmeasure q -> c;\n"
    p = parse(src)
    I = orc.instr
    ref = orc.program([I("h", [0]), I("cx", [0, 1]), I("cx", [1, 2]), I("rzz", [0, 2], [0.3]), orc.measure_all(3)])
    a, _ = orc.run(orc.Compiled(p.rows, 3, cregs=p.cregs, fusion=2), 2000, 11)
    b, _ = orc.run(orc.Compiled(ref, 3, fusion=2), 2000, 11)
    assert a == b and set(a) <= {0, 7, 1, 6, 2, 5, 3, 4}


@pytest.mark.parametrize("src,line,col,text", [
    ("OPENQASM 3.1;\n", 1, 10, "unsupported OpenQASM version"),
    (HDR + "qreg q[2];\nH q[0];\n", 4, 1, "did you mean 'h'"),
    (HDR + "qreg a[2];\nqreg b[3];\ncx a, b;\n", 5, 1, "broadcast size mismatch"),
    (HDR + "qreg q[2];\ncx q[1], q;\n", 4, 1, "duplicate qubit 'q[1]'"),
    (HDR + "qreg q[2];\nx q[5];\n", 4, 5, "index 5 out of range"),
    (HDR + "qreg q[2];\nx q[0]\nx q[1];\n", 5, 1, "expected ';'"),
    (HDR + "qreg q[1];\ncreg c[1];\nif (c = 1) x q[0];\n", 5, 7, "single '='"),
    (HDR + "qreg q[1];\nrz(theta) q[0];\n", 4, 4, "unknown identifier 'theta'"),
    (HDR + "qreg q[1];\ngate g a { g a; }\n", 4, 12, "cannot call itself"),
    (HDR + "qreg q[1];\nopaque foo a;\n", 4, 1, "opaque"),
    ('OPENQASM 2.0;\ninclude "other.inc";\n', 2, 9, "only \"qelib1.inc\""),
    (HDR + "qreg q[49];\n", 3, 6, "too many qubits"),
])
def test_errors_carry_positions(src, line, col, text):
    with pytest.raises(QasmError) as e:
        parse(src)
    msg = str(e.value)
    assert msg.startswith(f"line {line}:{col}:"), msg
    assert text in msg, msg
