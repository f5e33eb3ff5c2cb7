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


def test_plan_and_resolve_precision_through_the_c_abi(monkeypatch):
    """mem::plan / resolve_precision (src/mem.rs:114-130, src/exec.rs:261-304)
    behind zr_plan / zr_resolve_precision: Error::Memory's fields and text,
    the auto-precision fallback and its notice, and mem::budget_bytes'
    precedence (explicit limit, then ZENO_MEM_BYTES x fraction)."""
    with pytest.raises(z.MemoryBudget) as ei:
        z.plan(30, 64, 1 << 30)
    e = ei.value
    assert (e.n_qubits, e.precision, e.needed_bytes, e.budget_bytes, e.max_qubits) == (30, 64, 16 << 30, 1 << 30, 26)
    assert "needs 16.0 GiB but the budget is 1.0 GiB (max 26 qubits at this precision" in str(e)
    assert z.plan(26, 64, 1 << 30) == 1 << 30  # exactly at the budget fits (needed <= budget)
    # auto: f64 when it fits, else f32 with the notice; an explicit precision never falls back
    assert z.resolve_precision(29, z.RunOptions(mem_limit=8 << 30)) == (64, [])
    assert z.resolve_precision(30, z.RunOptions(mem_limit=(16 << 30) - 1)) == (
        32, ["auto-precision: 30 qubits exceed the f64 budget, using f32 (8 B/amplitude)"])
    with pytest.raises(z.MemoryBudget):
        z.resolve_precision(30, z.RunOptions(mem_limit=(16 << 30) - 1, precision=64))
    with pytest.raises(z.MemoryBudget) as ei:
        z.resolve_precision(31, z.RunOptions(mem_limit=(16 << 30) - 1))
    assert ei.value.precision == 32 and ei.value.max_qubits == 30
    # budget precedence: explicit limit wins; otherwise ZENO_MEM_BYTES * fraction
    monkeypatch.setenv("ZENO_MEM_BYTES", str(40 << 30))
    assert z.budget_bytes(123, 0.75) == 123
    assert z.budget_bytes(None, 0.75) == 30 << 30
    assert z.resolve_precision(31, z.RunOptions(mem_fraction=0.8)) == (64, [])  # 32 GiB: exactly fits
    assert z.resolve_precision(31, z.RunOptions()) == (32, [
        "auto-precision: 31 qubits exceed the f64 budget, using f32 (8 B/amplitude)"])


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
