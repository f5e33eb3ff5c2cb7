"""CPU tests of the product's host side (no GPU): the C-ABI library loads
and exports every declared symbol, and the native generators and compiler
produce exactly the oracle's instruction and op streams."""
import os
import re

import numpy as np
import pytest

import zeno_amd as z
from zeno_amd import _lib
from circuits import diag_ladder, disjoint_blocks, random_mixed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "zeno_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(z[br]_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 30
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"


def test_backend_fails_loudly_without_device():
    if z.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(z.Unsupported):
        z.HipBackend(4)
    with pytest.raises(z.Unsupported):
        z.Server(1 << 30)
    with pytest.raises(z.Unsupported):
        z.device_memory()


def test_generators_match_oracle(orc):
    pairs = [
        (z.random_circuit(7, 5, 3), orc.gen_random(7, 5, 3)),
        (z.random_circuit(20, 12, 42), orc.gen_random(20, 12, 42)),
        (z.matching_circuit(9, 10, 5, 5), orc.gen_matching(9, 10, 5, 5)),
        (z.matching_circuit(30, 40, 7, 5), orc.gen_matching(30, 40, 7, 5)),
        (z.qft(11), orc.gen_qft(11)),
        (z.u3_qft(24, 3), orc.gen_u3_qft(24, 3)),
    ]
    for a, b in pairs:
        assert a.to_program()[0].tobytes() == b.tobytes()


def _same_ops(c1, c2):
    assert len(c1) == len(c2)
    assert c1.stats == c2.stats
    for x, y in zip(c1.ops(), c2.ops()):
        assert x["kind"] == y["kind"] and x["qubits"] == y["qubits"] and x["is_if"] == y["is_if"]
        for key in ("mat", "diag"):
            if key in x:
                assert np.array_equal(x[key].view(np.uint64), y[key].view(np.uint64)), "bit-identical fused data"


def test_compiler_matches_oracle_bit_for_bit(orc):
    progs = [
        (orc.gen_matching(12, 8, 5, 5), 12),
        (random_mixed(orc, 8, 120, 77), 8),
        (diag_ladder(orc), 6),
        (disjoint_blocks(orc), 8),
        (orc.gen_u3_qft(10, 4), 10),
    ]
    for prog, n in progs:
        for fusion in range(7):
            _same_ops(z.compile(prog, n, fusion=fusion), orc.Compiled(prog, n, fusion=fusion))


def test_compiler_errors_match_reference_messages():
    with pytest.raises(z.InvalidCircuit, match="unknown gate 'foo'"):
        z.compile(z.ir.program([z.ir.instr("foo", [0])]), 2)
    with pytest.raises(z.InvalidCircuit, match="duplicate qubits"):
        z.compile(z.ir.program([z.ir.instr("cx", [1, 1])]), 2)
    with pytest.raises(z.InvalidCircuit, match="out of range"):
        z.compile(z.ir.program([z.ir.instr("h", [3])]), 2)
    with pytest.raises(z.InvalidCircuit, match="expects 1 parameter"):
        z.compile(z.ir.program([z.ir.instr("rx", [0], [])]), 2)
    with pytest.raises(z.InvalidCircuit, match="no qubits"):
        z.compile(z.ir.program([]), 0)


def test_format_key_and_mem_planning():
    assert z.format_key(0b01110, [2, 3]) == "011 10"
    assert z.max_qubits(24 << 30, 64) == 30 and z.max_qubits(24 << 30, 32) == 31
    assert z.state_bytes(30, 64) == 16 << 30 and z.state_bytes(30, 32) == 8 << 30
    with pytest.raises(z.MemoryBudget):
        z.plan(30, 64, 1 << 30)
    z.plan(20, 64, 1 << 30)


def test_teleport_compiles_dynamic():
    I = z.ir.instr
    rows = [I("x", [0]), I("h", [1]), I("cx", [1, 2]), I("cx", [0, 1]), I("h", [0]),
            z.ir.measure(0, 0), z.ir.measure(1, 1), I("x", [2], cond=(1, 1)), I("z", [2], cond=(0, 1)),
            z.ir.measure(2, 2)]
    c = z.compile(z.ir.program(rows), 3, cregs=[1, 1, 1])
    assert c.dynamic and c.op(7)["is_if"] and c.op(7)["cond"] == [1, 1, 1]


def test_tune_cache_roundtrip(tmp_path):
    path = str(tmp_path / "tune.json")
    assert z.tune.load_cache(path) == {}
    entry = {"knobs": {"tile_bits": 13, "slots": 4}, "fusion": 3, "ms": {}}
    z.tune.save_cache({"20:64": entry}, path)
    assert z.tune.load_cache(path)["20:64"] == entry
    assert z.tune.tuned_fusion(20, 64, path) == 3
    assert z.tune.tuned_fusion(21, 64, path) is None
    assert z.tune.load(path) == 1  # installs into the native defaults table (no device needed)
    z.tune.uninstall(20, 64)
    lib = z._lib.lib()
    assert lib.zb_tune_default(20, 64, b"no_such_knob", 1) < 0
    assert lib.zb_tune_default(20, 16, b"slots", 3) < 0


def test_fusion_auto_uses_cache(tmp_path, monkeypatch):
    path = str(tmp_path / "tune.json")
    z.tune.save_cache({"9:64": {"knobs": {}, "fusion": 4}}, path)
    monkeypatch.setenv("ZENO_TUNE_CACHE", path)
    from zeno_amd.exec import compile_options

    assert compile_options(z.RunOptions(fusion_max="auto", precision=64), 9) == 4
    assert compile_options(z.RunOptions(fusion_max="auto", precision=64), 10) == 1
    assert compile_options(z.RunOptions(fusion_max=2), 9) == 2
    assert compile_options(z.RunOptions(fusion_max="auto", noise=z.NoiseModel(bit_flip=0.1)), 9) == 0
