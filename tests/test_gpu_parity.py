"""GPU parity of the HIP backend against the CPU oracle, through the C ABI.

Bar: bit-exact for every sweep (the backend reproduces the reference
expression trees without FMA contraction), for CX (a pure permutation), for
prob_one / norm_sqr (same fixed reduction order), for the sampler's indices
and hence for counts.  At sizes the oracle cannot finish quickly, the tests
use size-independent properties: norm, forward-then-adjoint returns |0>,
and the QFT's closed form.
"""
import math
import sys

import numpy as np
import pytest

import zeno_amd as z
from circuits import diag_ladder, disjoint_blocks, random_mixed

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _rand_state(n, prec, seed):
    rng = np.random.default_rng(seed)
    dt = np.float64 if prec == 64 else np.float32
    v = rng.normal(size=(2, 1 << n))
    v /= np.sqrt(np.sum(v**2))
    return v[0].astype(dt), v[1].astype(dt)


def _bits(a):
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def _assert_same(gpu, ora, what):
    gr, gi = gpu.raw()
    orr, oi = ora.raw()
    ok = np.array_equal(_bits(gr), _bits(orr)) and np.array_equal(_bits(gi), _bits(oi))
    if not ok:
        d = max(np.max(np.abs(gr.astype(np.float64) - orr)), np.max(np.abs(gi.astype(np.float64) - oi)))
        raise AssertionError(f"{what}: not bit-identical (max |d| = {d:.3e})")


def _pair(orc, n, prec, seed):
    re, im = _rand_state(n, prec, seed)
    g = z.HipBackend(n, prec, z.HipBackend.EXACT)
    g.set_raw(re, im)
    o = orc.State(n, prec)
    o.set_raw(re, im)
    return g, o


@pytest.mark.parametrize("prec", [64, 32])
def test_single_ops_bit_exact(orc, prec):
    n = 13
    rng = np.random.default_rng(1)
    g, o = _pair(orc, n, prec, 11)
    for q in range(n):  # 1q on every target
        m = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        _say(f"1q prec={prec} q={q}")
        g.apply_unitary([q], m)
        o.apply_unitary([q], m)
    _assert_same(g, o, "1q sweep")
    for k in range(2, 7):  # dense kq, low/high/mixed supports
        for qs in ([0, 1, 2, 3, 4, 5][:k], list(range(n - k, n)), sorted(rng.choice(n, k, replace=False).tolist())):
            m = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
            _say(f"kq prec={prec} k={k} qs={qs}")
            g.apply_unitary(qs, m)
            o.apply_unitary(qs, m)
            _assert_same(g, o, f"kq k={k} qs={qs}")
    for k in (1, 2, 3, 5, 8, 10, 12):
        qs = sorted(rng.choice(n, k, replace=False).tolist())
        d = np.exp(1j * rng.uniform(0, 2 * np.pi, 1 << k)) * rng.uniform(0.5, 1.5, 1 << k)
        _say(f"diag prec={prec} k={k} qs={qs}")
        g.apply_diagonal(qs, d)
        o.apply_diagonal(qs, d)
        _assert_same(g, o, f"diag k={k}")
    for c, t in [(0, 1), (1, 0), (0, n - 1), (n - 1, 0), (5, 9), (9, 5)]:
        _say(f"cx prec={prec} c={c} t={t}")
        g.apply_cx(c, t)
        o.apply_cx(c, t)
    _assert_same(g, o, "cx")
    assert g.norm_sqr() == o.norm_sqr()
    for q in range(n):
        assert g.prob_one(q) == o.prob_one(q), q
    g.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_exact_mode_diagonals_bit_exact(orc, prec):
    """The per-op diagonal kernels (exact mode): tables of 1 to 12 bits on
    qubits below and above index bit 8 (the table index splits into a lane
    part and a row part there), and a state shorter than one workgroup."""
    rng = np.random.default_rng(17)
    for n, ks in ((5, (1, 2, 3, 5)), (14, (1, 2, 4, 7, 9, 12))):
        g = z.HipBackend(n, prec, z.HipBackend.EXACT)
        o = orc.State(n, prec)
        prog = orc.gen_random(n, 2, 50 + n)
        g.apply_compiled(z.compile(prog, n, fusion=1))
        o.apply_compiled(orc.Compiled(prog, n, fusion=1))
        for k in ks:
            for qs in (sorted(rng.choice(n, k, replace=False).tolist()), list(range(n - k, n))):
                d = np.exp(1j * rng.uniform(0, 2 * np.pi, 1 << k)) * rng.uniform(0.5, 1.5, 1 << k)
                _say(f"exact diag prec={prec} n={n} k={k} qs={qs}")
                g.apply_diagonal(qs, d)
                o.apply_diagonal(qs, d)
        _assert_same(g, o, f"exact diagonals n={n}")
        g.close()


MODES = [z.HipBackend.EXACT, z.HipBackend.TILED]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("prec", [64, 32])
def test_random_circuits_bit_exact(orc, prec, mode):
    for n in (2, 3, 5, 8, 11, 12, 13, 14, 17, 20):
        prog = orc.gen_random(n, 12, 1000 + n)
        for fusion in (0, 1, 2, 4, 6):
            _say(f"random_circuit n={n} fusion={fusion} prec={prec} mode={mode}")
            comp = orc.Compiled(prog, n, fusion=fusion)
            o = orc.State(n, prec)
            o.apply_compiled(comp)
            with z.HipBackend(n, prec, mode) as g:
                g.apply_compiled(z.compile(prog, n, fusion=fusion))
                _assert_same(g, o, f"n={n} fusion={fusion}")


@pytest.mark.parametrize("tune", [{"win_grid": 3, "prefetch": 1}, {"win_grid": 5, "prefetch": 0},
                                  {"win_grid": 2, "slots": 4}, {"win_grid": 4, "slots": 3},
                                  {"tile_bits": 13, "win_grid": 3}, {"tile_bits": 13, "slots": 4, "prefetch": 0},
                                  {"tile_bits": 13}, {"defer_cx": 0, "win_grid": 3}])
@pytest.mark.parametrize("prec", [64, 32])
def test_tiled_multi_tile_workgroups_bit_exact(orc, prec, tune):
    """A small window grid makes every workgroup walk many tiles (the
    prefetch and write-back hand-over between consecutive tiles)."""
    for n in (12, 13, 14, 17):
        for prog in (orc.gen_random(n, 10, 77 + n), orc.gen_matching(n, 10, 5 + n, 3)):
            for fusion in (1, 3):
                _say(f"multi-tile n={n} fusion={fusion} prec={prec} tune={tune}")
                o = orc.State(n, prec)
                o.apply_compiled(orc.Compiled(prog, n, fusion=fusion))
                with z.HipBackend(n, prec, z.HipBackend.TILED) as g:
                    for k, v in tune.items():
                        g.tune(k, v)
                    g.apply_compiled(z.compile(prog, n, fusion=fusion))
                    _assert_same(g, o, f"n={n} fusion={fusion} tune={tune}")


@pytest.mark.parametrize("layout", [1, 5, 11])
@pytest.mark.parametrize("mode", MODES)
def test_qubit_layouts_bit_exact(orc, mode, layout):
    """Any qubit-to-bit layout (the spread one, seeded shuffles) gives the
    oracle's state, norm, probabilities, chunk sums, samples and collapse."""
    for n in (12, 14, 17, 20):
        for prec in (64, 32):
            prog = orc.gen_matching(n, 8, 40 + n, 3)
            for fusion in (0, 2, 4):
                _say(f"layout={layout} mode={mode} n={n} prec={prec} fusion={fusion}")
                o = orc.State(n, prec)
                o.apply_compiled(orc.Compiled(prog, n, fusion=fusion))
                with z.HipBackend(n, prec, mode) as g:
                    g.tune("layout", layout)
                    assert sorted(g.layout()) == list(range(n))
                    g.apply_compiled(z.compile(prog, n, fusion=fusion))
                    _assert_same(g, o, f"layout={layout} n={n} fusion={fusion}")
                    assert g.norm_sqr() == o.norm_sqr()
                    for q in (0, n // 2, n - 1):
                        assert g.prob_one(q) == o.prob_one(q), q
                    assert np.array_equal(g.chunk_sums().view(np.uint64), o.chunk_sums().view(np.uint64))
                    assert np.array_equal(g.sample(3000, 9), o.sample(3000, 9))
                    assert g.measure(n - 2, 0.4) == o.measure(n - 2, 0.4)
                    _assert_same(g, o, "after collapse")
                    with z.HipBackend(n, prec, mode) as h:  # identity layout, same circuit
                        h.tune("layout", 0)
                        h.apply_compiled(z.compile(prog, n, fusion=fusion))
                        h.measure(n - 2, 0.4)
                        assert g.max_abs_diff(h) == 0.0


@pytest.mark.parametrize("mode", MODES)
def test_adversarial_circuits_bit_exact(orc, mode):
    cases = [(random_mixed(orc, n, 80, 9000 + n), n) for n in (3, 6, 10, 13, 15)]
    cases += [(diag_ladder(orc), 6), (disjoint_blocks(orc), 8), (orc.gen_u3_qft(12, 5), 12),
              (orc.gen_u3_qft(17, 5), 17), (orc.gen_matching(16, 10, 3, 5), 16),
              (orc.gen_matching(19, 12, 4, 3), 19)]
    for prog, n in cases:
        for fusion in (0, 1, 3, 5):
            _say(f"adversarial n={n} fusion={fusion} mode={mode}")
            o = orc.State(n, 64)
            o.apply_compiled(orc.Compiled(prog, n, fusion=fusion))
            with z.HipBackend(n, 64, mode) as g:
                g.apply_compiled(z.compile(prog, n, fusion=fusion))
                _assert_same(g, o, f"n={n} fusion={fusion}")


@pytest.mark.parametrize("prec", [64, 32])
def test_relabelling_windows_bit_exact(orc, prec):
    """Batched runs relabel qubits on window write-back (the layout moves);
    the state, norm, probabilities and samples equal the oracle's and a run
    with relabelling off, bit for bit."""
    for n, prog, fusion in ((20, orc.gen_matching(20, 12, 41, 5), 1), (22, orc.gen_random(22, 10, 42), 2),
                            (24, orc.gen_matching(24, 8, 43, 3), 1)):
        _say(f"relabel prec={prec} n={n} fusion={fusion}")
        o = orc.State(n, prec)
        o.apply_compiled(orc.Compiled(prog, n, fusion=fusion))
        comp = z.compile(prog, n, fusion=fusion)
        with z.HipBackend(n, prec, z.HipBackend.TILED) as g, z.HipBackend(n, prec, z.HipBackend.TILED) as f:
            start = list(g.layout())
            f.tune("relabel", 0)
            g.apply_compiled(comp)
            f.apply_compiled(comp)
            assert list(g.layout()) != start, "no window relabelled"
            assert list(f.layout()) == start
            _assert_same(g, o, f"n={n} relabel")
            _assert_same(f, o, f"n={n} no relabel")
            assert g.norm_sqr() == o.norm_sqr()
            for q in (0, n // 2, n - 1):
                assert g.prob_one(q) == o.prob_one(q)
            assert np.array_equal(g.sample(5000, 3), o.sample(5000, 3))


@pytest.mark.parametrize("prec", [32, 64])
def test_relabelling_large_state_deterministic(prec):
    """At 28 qubits (65536 tiles, every wave of a workgroup busy) the
    permuted write-back scatter of a relabelling window reads positions that
    other waves overwrite: the result must equal relabelling off exactly, on
    every repeat."""
    n = 28
    comp = z.compile(z.matching_circuit(n, 5, 3, 5).to_program()[0], n, fusion=1)
    with z.HipBackend(n, prec) as f:
        f.tune("relabel", 0)
        f.apply_compiled(comp)
        ref = f.norm_sqr()
        for rep in range(3):
            _say(f"relabel n=28 prec={prec} repeat {rep}")
            with z.HipBackend(n, prec) as g:
                g.apply_compiled(comp)
                assert list(g.layout()) != list(f.layout()), "no window relabelled"
                assert g.max_abs_diff(f) == 0.0, rep
                assert g.norm_sqr() == ref, rep


def _unitary(rng, k):
    q, r = np.linalg.qr(rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _max_diff(g, o):
    gr, gi = g.raw()
    orr, oi = o.raw()
    return max(np.max(np.abs(gr.astype(np.float64) - orr)), np.max(np.abs(gi.astype(np.float64) - oi)))


MFMA_TOL = {64: 1e-12, 32: 1e-5}


@pytest.mark.parametrize("prec", [64, 32])
def test_dense_ops_on_lds_tiles_bit_exact(orc, prec):
    """Dense 2..6-qubit ops through LDS tiles (knob kq_tiled=2; by default
    from 12 qubits where the per-lane gather is slow) and gathered per lane
    (kq_tiled=0) both give the oracle's kern_kq bits: op qubits low, high,
    low and high mixed, spread, under a shuffled qubit layout (the op's
    state bits out of order), and at 17 qubits where each workgroup walks
    several tiles.  Under the knob fma=1 the two data paths still agree bit
    for bit with each other."""
    rng = np.random.default_rng(8)
    for n, layout in ((12, 0), (14, 0), (17, 0), (14, 5)):
        re, im = _rand_state(n, prec, 60 + n)
        o = orc.State(n, prec)
        o.set_raw(re, im)
        gs = {}
        for tiled in (2, 0):
            g = z.HipBackend(n, prec, z.HipBackend.EXACT)
            g.tune("layout", layout)
            g.tune("kq_tiled", tiled)
            g.set_raw(re, im)
            gs[tiled] = g
        for k in (2, 3, 4, 5, 6):
            for qs in (list(range(k)), list(range(n - k, n)), [0, 1, 2, n - 3, n - 2, n - 1][:k],
                       sorted(rng.choice(n, k, replace=False).tolist())):
                m = _unitary(rng, k)
                _say(f"kq tiles prec={prec} n={n} layout={layout} k={k} qs={qs}")
                o.apply_unitary(qs, m)
                for tiled, g in gs.items():
                    g.apply_unitary(qs, m)
                    _assert_same(g, o, f"n={n} layout={layout} k={k} qs={qs} kq_tiled={tiled}")
        for g in gs.values():
            g.close()
    n = 14
    re, im = _rand_state(n, prec, 99)
    gs = {}
    for tiled in (2, 0):
        g = z.HipBackend(n, prec, z.HipBackend.EXACT)
        g.tune("fma", 1)
        g.tune("kq_tiled", tiled)
        g.set_raw(re, im)
        gs[tiled] = g
    for k, qs in ((2, [0, 13]), (3, [0, 1, 2]), (4, [0, 5, 9, 13]), (5, [0, 1, 2, 3, 4]), (6, [0, 2, 7, 8, 11, 13])):
        m = _unitary(rng, k)
        _say(f"kq tiles fma prec={prec} k={k} qs={qs}")
        for g in gs.values():
            g.apply_unitary(qs, m)
    (ar, ai), (br, bi) = gs[2].raw(), gs[0].raw()
    assert np.array_equal(_bits(ar), _bits(br)) and np.array_equal(_bits(ai), _bits(bi))
    for g in gs.values():
        g.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_mfma_dense_ops_within_tolerance(orc, prec):
    """Dense 4..6-qubit ops on the matrix cores (knob mfma=1) against the
    oracle's scalar kern_kq on the same state: op qubits on the low run
    bits, the top bits, mixed, under a shuffled qubit layout (op bits out of
    order), and at 23 qubits where every workgroup walks several tiles."""
    rng = np.random.default_rng(5)
    tol = MFMA_TOL[prec]
    for n, layout in ((12, 0), (13, 0), (16, 0), (16, 7)):
        re, im = _rand_state(n, prec, 40 + n)
        o = orc.State(n, prec)
        o.set_raw(re, im)
        with z.HipBackend(n, prec, z.HipBackend.EXACT) as g:
            g.tune("layout", layout)
            g.tune("mfma", 1)
            g.set_raw(re, im)
            for k in (4, 5, 6):
                for qs in (list(range(k)), list(range(n - k, n)), sorted(rng.choice(n, k, replace=False).tolist())):
                    m = _unitary(rng, k)
                    _say(f"mfma prec={prec} n={n} layout={layout} k={k} qs={qs}")
                    g.apply_unitary(qs, m)
                    o.apply_unitary(qs, m)
                    d = _max_diff(g, o)
                    assert d <= tol, f"n={n} k={k} qs={qs}: max |d| = {d:.3e}"
    n = 23
    orc.set_threads(16)
    try:
        re, im = _rand_state(n, prec, 77)
        o = orc.State(n, prec)
        o.set_raw(re, im)
        with z.HipBackend(n, prec, z.HipBackend.EXACT) as g:
            g.tune("mfma", 1)
            g.set_raw(re, im)
            for k, qs in ((4, [0, 6, 13, 22]), (5, [1, 2, 3, 17, 20]), (6, [5, 8, 9, 10, 11, 21])):
                m = _unitary(rng, k)
                _say(f"mfma prec={prec} n={n} k={k} qs={qs}")
                g.apply_unitary(qs, m)
                o.apply_unitary(qs, m)
            d = _max_diff(g, o)
            assert d <= tol, f"n={n}: max |d| = {d:.3e}"
    finally:
        orc.set_threads(1)


@pytest.mark.parametrize("prec", [64, 32])
def test_mfma_fused_circuits_within_tolerance(orc, prec):
    """Compiled circuits at fusion 4..6 with the dense blocks on the matrix
    cores, in the tiled mode, against the oracle's scalar path."""
    tol = MFMA_TOL[prec]
    for n in (12, 14, 17):
        for prog in (orc.gen_random(n, 10, 300 + n), orc.gen_matching(n, 10, 310 + n, 3)):
            for fusion in (4, 5, 6):
                _say(f"mfma circuit prec={prec} n={n} fusion={fusion}")
                o = orc.State(n, prec)
                o.apply_compiled(orc.Compiled(prog, n, fusion=fusion))
                with z.HipBackend(n, prec, z.HipBackend.TILED) as g:
                    g.tune("mfma", 1)
                    g.apply_compiled(z.compile(prog, n, fusion=fusion))
                    d = _max_diff(g, o)
                    assert d <= tol, f"n={n} fusion={fusion}: max |d| = {d:.3e}"


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("prec", [64, 32])
def test_fma_knob_within_tolerance(orc, prec, mode):
    """Knob fma=1: the same expression trees with fused multiply-adds, in
    the per-op kernels and the tiled windows, against the oracle's separately
    rounded path.  The results must differ somewhere (the knob is live) and
    stay within tolerance."""
    tol = MFMA_TOL[prec]
    worst = 0.0
    for n in (5, 12, 14, 17):
        for prog in (orc.gen_random(n, 12, 500 + n), orc.gen_matching(n, 12, 510 + n, 3)):
            for fusion, knobs in ((1, {"fma": 1}), (2, {"fma": 1}), (4, {"fma": 1}), (5, {"fma": 1, "mfma": 1})):
                _say(f"fma prec={prec} mode={mode} n={n} fusion={fusion} knobs={knobs}")
                o = orc.State(n, prec)
                o.apply_compiled(orc.Compiled(prog, n, fusion=fusion))
                with z.HipBackend(n, prec, mode) as g:
                    for k, v in knobs.items():
                        g.tune(k, v)
                    g.apply_compiled(z.compile(prog, n, fusion=fusion))
                    d = _max_diff(g, o)
                    assert d <= tol, f"n={n} fusion={fusion}: max |d| = {d:.3e}"
                    worst = max(worst, d)
    assert worst > 0.0


def test_sampler_indices_and_chunk_sums_bit_exact(orc):
    for n, shots, seed in ((1, 100, 3), (4, 1000, 1), (10, 5000, 2), (16, 20000, 5), (17, 4096, 3), (20, 100000, 4)):
        _say(f"sample n={n} shots={shots}")
        prog = random_mixed(orc, n, 3 * n, 700 + n)
        o = orc.State(n, 64)
        o.apply_compiled(orc.Compiled(prog, n, fusion=3))
        with z.HipBackend(n, 64) as g:
            g.apply_compiled(z.compile(prog, n, fusion=3))
            assert np.array_equal(g.chunk_sums().view(np.uint64), o.chunk_sums().view(np.uint64))
            assert np.array_equal(g.sample(shots, seed), o.sample(shots, seed))


def test_sampler_many_chunks_bit_exact(orc):
    """At 27 qubits (2048 chunks, 8 per CU) the chunk walks take their
    one-wave shape, and so does the rescan of a 100000-shot sample (device
    sort, a block per chunk); 1000 shots rescan the touched chunks with the
    four-wave shape.  With chunk_multi=1024 pass 1 walks 8 chunks per
    workgroup instead (the default from 4096 chunks).  Chunk sums and draws
    equal the oracle's."""
    n = 27
    prog = orc.gen_random(n, 2, 2700)
    orc.set_threads(16)
    try:
        for prec in (32, 64):
            _say(f"sampler many chunks n={n} prec={prec}")
            o = orc.State(n, prec)
            o.apply_compiled(orc.Compiled(prog, n, fusion=2))
            want = o.chunk_sums().view(np.uint64)
            with z.HipBackend(n, prec) as g:
                g.apply_compiled(z.compile(prog, n, fusion=2))
                for multi in (0, 1024):
                    g.tune("chunk_multi", multi)
                    assert np.array_equal(g.chunk_sums().view(np.uint64), want), multi
                for shots, seed in ((100000, 5), (1000, 6)):
                    assert np.array_equal(g.sample(shots, seed), o.sample(shots, seed)), (prec, shots)
            del o
    finally:
        orc.set_threads(1)


def test_sampler_device_draws_at_block_and_jump_edges(orc):
    """From 32768 shots the draws come from the device: threads of 256
    draws each, started by jump maps over the bits of the thread index.  At
    the first device shot count, one past it, a count that ends inside a
    block, one whose thread count needs another jump level, and 2^20 + 3,
    the indices equal the oracle's sequential stream (host path below)."""
    n = 14
    prog = random_mixed(orc, n, 3 * n, 1414)
    for prec in (64, 32):
        o = orc.State(n, prec)
        o.apply_compiled(orc.Compiled(prog, n, fusion=2))
        with z.HipBackend(n, prec) as g:
            g.apply_compiled(z.compile(prog, n, fusion=2))
            for shots, seed in ((32767, 1), (32768, 2), (32769, 3), (65536 + 100, 4), (256 * 513, 5),
                                ((1 << 20) + 3, 6)):
                _say(f"device draws prec={prec} shots={shots}")
                assert np.array_equal(g.sample(shots, seed), o.sample(shots, seed)), (prec, shots)


def test_sampler_concentrated_distributions_bit_exact(orc):
    """Mass on a few indices: GHZ (index 0 and the last index, every chunk
    between empty) and H on qubits 0..2 only (8 indices inside the first
    16-value group), so runs of draws land on one index and most chunks
    contribute nothing; the draws equal the oracle's."""
    for n, kind in ((18, "ghz"), (21, "ghz"), (19, "h3")):
        if kind == "ghz":
            rows = [orc.instr("h", [0])] + [orc.instr("cx", [q, q + 1]) for q in range(n - 1)]
        else:
            rows = [orc.instr("h", [q]) for q in range(3)]
        prog = orc.program(rows)
        for prec in (64, 32):
            _say(f"sample concentrated n={n} {kind} prec={prec}")
            o = orc.State(n, prec)
            o.apply_compiled(orc.Compiled(prog, n, fusion=1))
            with z.HipBackend(n, prec) as g:
                g.apply_compiled(z.compile(prog, n, fusion=1))
                for shots, seed in ((1000, 11), (50000, 12)):
                    assert np.array_equal(g.sample(shots, seed), o.sample(shots, seed)), (n, kind, prec, shots)


def test_measure_and_collapse_bit_exact(orc):
    n = 12
    for prec in (64, 32):
        g, o = _pair(orc, n, prec, 21)
        for i, (q, u) in enumerate([(0, 0.3), (5, 0.9), (11, 0.1), (3, 0.5)]):
            _say(f"measure prec={prec} q={q}")
            assert g.measure(q, u) == o.measure(q, u)
            _assert_same(g, o, f"collapse #{i}")
        g.reset_all()
        o.reset_all()
        _assert_same(g, o, "reset")
        g.close()


def test_20_qubit_counts_identical(orc):
    """20 qubits, depth 12, fusion 2, 1024 shots, seed 42 — both the
    reference brickwork and the random-matching variant."""
    for prog in (orc.gen_random(20, 12, 42), orc.gen_matching(20, 12, 42)):
        full = orc.program([prog, orc.measure_all(20)])
        _say("20q d12 counts")
        want, _ = orc.run(orc.Compiled(full, 20, fusion=2), 1024, 42)
        r = z.run_program(full, 20, opts=z.RunOptions(shots=1024, seed=42, precision=64, fusion_max=2))
        assert r.raw_counts == want
        assert r.counts.total() == 1024 and r.backend == "hip-f64"


def test_counts_with_scrambled_repeated_and_unmeasured_clbits(orc):
    """Final measurements in any order into two registers: qubit q to clbit
    (7q + 3) mod 18 of 20 clbits, qubit 5 measured a second time into clbit
    0 (the later measure decides it), clbits 18 and 19 never written.  The
    clbit words come from per-byte tables of the sample index; counts equal
    the oracle's (exec.rs:442-447)."""
    n = 17
    rows = [orc.gen_random(n, 6, 808)]
    rows += [orc.measure(q, (7 * q + 3) % 18) for q in range(n)] + [orc.measure(5, 0)]
    prog = orc.program(rows)
    cregs = [12, 8]
    for shots, seed in ((4096, 3), (40000, 4)):
        _say(f"scrambled clbits n={n} shots={shots}")
        want, _ = orc.run(orc.Compiled(prog, n, cregs=cregs, fusion=1), shots, seed)
        got = z.run_program(prog, n, cregs=cregs, opts=z.RunOptions(shots=shots, seed=seed, precision=64, fusion_max=1))
        assert got.raw_counts == want


@pytest.mark.parametrize("mode", MODES)
def test_dynamic_and_noisy_runs_match_oracle(orc, mode):
    """Per-shot path: EXACT replays every shot through the zb_* calls,
    TILED runs all shots as one batched trajectory launch."""
    I = orc.instr
    rows = [I("x", [0]), I("h", [1]), I("cx", [1, 2]), I("cx", [0, 1]), I("h", [0]),
            orc.measure(0, 0), orc.measure(1, 1), I("x", [2], cond=(1, 1)), I("z", [2], cond=(0, 1)),
            orc.measure(2, 2)]
    prog = orc.program(rows)
    _say(f"teleport per-shot mode={mode}")
    want, _ = orc.run(orc.Compiled(prog, 3, cregs=[1, 1, 1]), 200, 5)
    got = z.run_program(prog, 3, cregs=[1, 1, 1], opts=z.RunOptions(shots=200, seed=5, mode=mode))
    assert got.raw_counts == want
    nz = {"depolarizing_1q": 0.02, "depolarizing_2q": 0.05, "amplitude_damping": 0.03,
          "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}
    full = orc.program([orc.gen_random(5, 4, 9), orc.measure_all(5)])
    for prec in (64, 32):
        _say(f"noisy per-shot prec={prec} mode={mode}")
        want, _ = orc.run(orc.Compiled(full, 5, fusion=0), 64, 17, prec=prec, noise=nz)
        got = z.run_program(full, 5, opts=z.RunOptions(shots=64, seed=17, precision=prec, noise=z.NoiseModel(**nz),
                                                        mode=mode))
        assert got.raw_counts == want


def _traj_counts(words):
    vals, cnt = np.unique(np.asarray(words, np.uint64), return_counts=True)
    return {int(v): int(c) for v, c in zip(vals, cnt)}


@pytest.mark.parametrize("prec", [32, 64])
def test_batched_trajectories_match_oracle(orc, prec):
    """Batched noisy trajectories (every noise channel, reset, conditionals,
    3-qubit gates) give the CPU trajectory sampler's counts exactly."""
    I = orc.instr
    nz = {"depolarizing_1q": 0.01, "depolarizing_2q": 0.03, "bit_flip": 0.01, "phase_flip": 0.01,
          "amplitude_damping": 0.02, "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.03}
    rows = [orc.gen_random(9, 6, 4), I("ccx", [0, 4, 8]), I("cswap", [1, 2, 7]), orc.reset(3),
            orc.measure(5, 5), I("rx", [3], [0.7], cond=(1, 1)), orc.barrier(), orc.measure_all(9)]
    prog = orc.program(rows)
    for fusion in (0, 3):
        c = orc.Compiled(prog, 9, cregs=[5, 4], fusion=fusion)
        want, _ = orc.run(c, 300, 123, prec=prec, noise=nz)
        zc = z.compile(prog, 9, cregs=[5, 4], fusion=fusion)
        assert z.trajectories_supported(zc)
        _say(f"batched trajectories 9q prec={prec} fusion={fusion}")
        words, _ = z.run_trajectories(zc, 300, 123, prec, z.NoiseModel(**nz))
        assert _traj_counts(words) == want
        # shot s of a run seeded S is the single shot of a run seeded S ^ s
        for shot in range(0, 300, 7):
            one, _ = orc.run(c, 1, 123 ^ shot, prec=prec, noise=nz)
            assert list(one) == [int(words[shot])], shot
    # the configuration of the noisy-trajectory benchmark, 16 qubits depth 30
    nz16 = {"depolarizing_1q": 1e-3, "depolarizing_2q": 1e-2, "amplitude_damping": 5e-3,
            "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}
    prog16 = orc.program([orc.gen_random(16, 30, 2024), orc.measure_all(16)])
    orc.set_threads(16)
    try:
        want, _ = orc.run(orc.Compiled(prog16, 16, fusion=0), 48, 99, prec=prec, noise=nz16)
    finally:
        orc.set_threads(1)
    _say(f"batched trajectories 16q d30 prec={prec}")
    words, _ = z.run_trajectories(z.compile(prog16, 16, fusion=0), 48, 99, prec, z.NoiseModel(**nz16))
    assert _traj_counts(words) == want


@pytest.mark.parametrize("prec", [32, 64])
def test_batched_trajectories_fused_damping_probability(orc, prec):
    """From 13 qubits a 1q gate's sweep also forms its qubit's damping
    probability (17 qubits: more than one group of 8 reduction blocks); with
    every channel at high rates (Paulis often land between the gate and its
    damping draw, which falls back to the separate sum) each shot still
    equals the CPU trajectory sampler's."""
    nz = {"depolarizing_1q": 0.05, "depolarizing_2q": 0.05, "bit_flip": 0.05, "phase_flip": 0.05,
          "amplitude_damping": 0.1, "readout_flip_0to1": 0.01, "readout_flip_1to0": 0.01}
    for n, shots in ((13, 40), (14, 40), (17, 16)):
        prog = orc.program([orc.gen_random(n, 6, 31 + n), orc.measure_all(n)])
        c = orc.Compiled(prog, n, fusion=0)
        want, _ = orc.run(c, shots, 77, prec=prec, noise=nz)
        _say(f"fused damping probability {n}q prec={prec}")
        words, _ = z.run_trajectories(z.compile(prog, n, fusion=0), shots, 77, prec, z.NoiseModel(**nz))
        assert _traj_counts(words) == want
        for shot in range(0, shots, 9):
            one, _ = orc.run(c, 1, 77 ^ shot, prec=prec, noise=nz)
            assert list(one) == [int(words[shot])], shot


def _adjoint_roundtrip(n, prec, prog, fusion, tol):
    comp = z.compile(prog, n, fusion=fusion)
    with z.HipBackend(n, prec) as g:
        g.apply_compiled(comp)
        nrm = g.norm_sqr()
        for op in reversed(comp.ops()):
            if op["kind"] == z.exec.OP_U:
                g.apply_unitary(op["qubits"], op["mat"].conj().T)
            elif op["kind"] == z.exec.OP_D:
                g.apply_diagonal(op["qubits"], op["diag"].conj())
            else:
                g.apply_cx(*op["qubits"])
        p0 = 1.0 - max(g.prob_one(q) for q in range(n))
        assert abs(nrm - 1) < tol, nrm
        assert p0 > 1 - tol, p0


def test_large_state_roundtrip_and_norm():
    n = 26
    _say("roundtrip 26q f64 fusion=4")
    _adjoint_roundtrip(n, 64, z.matching_circuit(n, 6, 11, 5).to_program()[0], 4, 1e-9)
    _say("roundtrip 26q f32 fusion=5")
    _adjoint_roundtrip(n, 32, z.matching_circuit(n, 6, 11, 5).to_program()[0], 5, 1e-4)


def test_qft24_closed_form():
    n, x = 24, 0x5A5A5A
    rows = [z.ir.instr("x", [q]) for q in range(n) if (x >> q) & 1]
    prog = z.ir.program(rows + [z.qft(n).to_program()[0]])
    for fusion in (1, 4):
        _say(f"qft24 fusion={fusion}")
        with z.HipBackend(n, 64) as g:
            g.apply_compiled(z.compile(prog, n, fusion=fusion))
            sv = g.statevector()
        j = np.arange(1 << n, dtype=np.int64)
        rev = np.zeros_like(j)
        for b in range(n):
            rev |= ((j >> b) & 1) << (n - 1 - b)
        want = np.exp(2j * np.pi * ((x * j) % (1 << n)) / (1 << n)) / math.sqrt(1 << n)
        assert np.max(np.abs(sv[rev] - want)) < 1e-10


def _run_pair(orc, prog, n, fusion, prec, shots, seed):
    """GPU run vs oracle run of the same compiled program; state bits and
    counts must be identical."""
    full = orc.program([prog, orc.measure_all(n)])
    oc = orc.Compiled(full, n, fusion=fusion)
    o = orc.State(n, prec)
    o.apply_compiled(oc)
    gc = z.compile(full, n, fusion=fusion)
    with z.HipBackend(n, prec) as g:
        g.apply_compiled(gc)
        _assert_same(g, o, f"n={n} fusion={fusion} prec={prec}")
        assert np.array_equal(g.sample(shots, seed), o.sample(shots, seed))


def test_qft24_and_random26_bit_exact_with_1m_shots(orc):
    """QFT on 24 qubits after a seeded U3 layer, and the random circuit at 26
    qubits depth 12, fusion K in {2, 3, 4}, f64, 1M shots."""
    orc.set_threads(16)
    try:
        for fusion in (2, 3, 4):
            _say(f"qft24 fusion={fusion}")
            _run_pair(orc, orc.gen_u3_qft(24, 11), 24, fusion, 64, 1_000_000, 42)
        for fusion in (2, 3, 4):
            _say(f"random26 fusion={fusion}")
            _run_pair(orc, orc.gen_random(26, 12, 26), 26, fusion, 64, 1_000_000, 7)
    finally:
        orc.set_threads(1)


def test_33_qubit_state_roundtrip():
    """The largest state: 33 qubits (128 GiB f64, 64 GiB f32), matching
    circuit with RZZ ladders, forward then adjoint returns |0...0>."""
    free, _ = z.device_memory()
    n = 33
    for prec, tol in ((64, 1e-9), (32, 1e-4)):
        if free < (1 << n) * (16 if prec == 64 else 8) + (8 << 30):
            pytest.skip("not enough device memory for a 33-qubit state")
        _say(f"33q roundtrip prec={prec}")
        _adjoint_roundtrip(n, prec, z.matching_circuit(n, 5, 3, 5).to_program()[0], 1, tol)


def _serve_cases(orc):
    """Mixed requests: sizes 10..24 qubits, both precisions, a dynamic and a
    noisy circuit among them."""
    cases = []
    rng = np.random.default_rng(5)
    for i in range(22):
        n = int(rng.integers(10, 25))
        depth = int(rng.integers(4, 16))
        prog = orc.program([orc.gen_random(n, depth, 300 + i), orc.measure_all(n)])
        cases.append((prog, n, None, 32 if i % 2 else 64, 1000, 77 + i, None))
    I = orc.instr
    tele = orc.program([I("x", [0]), I("h", [1]), I("cx", [1, 2]), I("cx", [0, 1]), I("h", [0]),
                        orc.measure(0, 0), orc.measure(1, 1), I("x", [2], cond=(1, 1)), I("z", [2], cond=(0, 1)),
                        orc.measure(2, 2)])
    cases.append((tele, 3, [1, 1, 1], 64, 500, 5, None))
    nz = {"depolarizing_1q": 0.01, "depolarizing_2q": 0.02, "amplitude_damping": 0.01,
          "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}
    cases.append((orc.program([orc.gen_random(8, 5, 9), orc.measure_all(8)]), 8, None, 32, 400, 11, nz))
    return cases


def test_serving_results_same_alone_and_in_mix(orc):
    """Each served result equals the same request run alone (and, for the
    small ones, the oracle), with requests overlapping on the GPU and
    admitted as pool space frees up."""
    cases = _serve_cases(orc)
    alone = []
    for prog, n, cregs, prec, shots, seed, nz in cases:
        c = z.compile(prog, n, cregs, fusion=0 if nz else 1)
        r = z.run_compiled(c, z.RunOptions(shots=shots, seed=seed, precision=prec,
                                           noise=z.NoiseModel(**nz) if nz else None))
        alone.append(r.raw_counts)
        if n <= 14:
            want, _ = orc.run(orc.Compiled(prog, n, cregs=cregs, fusion=0 if nz else 1), shots, seed, prec=prec,
                              noise=nz)
            assert r.raw_counts == want, (n, prec, seed)
    # a pool that holds a few 24q f64 states at a time forces queueing
    for pool, slots in ((1 << 30, 4), (3 << 30, 8)):
        _say(f"serving mix pool={pool >> 20} MiB slots={slots}")
        with z.Server(pool, slots) as srv:
            ids = []
            for prog, n, cregs, prec, shots, seed, nz in cases:
                c = z.compile(prog, n, cregs, fusion=0 if nz else 1)
                ids.append(srv.submit(c, prec, shots, seed, z.NoiseModel(**nz) if nz else None))
            got = [srv.wait(i) for i in ids]
            assert srv.status()["queued"] == 0 and srv.status()["running"] == 0
        for g, a in zip(got, alone):
            assert g.raw_counts == a
            assert g.t_submit <= g.t_start <= g.t_end
        overlap = sum(1 for a in got for b in got if a.id < b.id and a.t_start < b.t_end and b.t_start < a.t_end)
        assert overlap > 0, "no two requests ran at the same time"


def test_serving_large_requests_one_at_a_time():
    """Requests of 1 GiB of state and more run one at a time while smaller
    ones share the other slots; every result equals the request run alone."""
    large = z.compile(z.Circuit(26).extend(z.random_circuit(26, 6, 3).to_program()[0]).measure_all().to_program()[0],
                      26, fusion=1)
    small = z.compile(z.Circuit(16).extend(z.random_circuit(16, 20, 4).to_program()[0]).measure_all().to_program()[0],
                      16, fusion=1)
    plan = [(large, 11), (small, 12), (large, 13), (small, 14), (small, 15), (large, 16), (small, 17)]
    alone = [z.run_compiled(c, z.RunOptions(shots=500, seed=s, precision=64)).raw_counts for c, s in plan]
    _say("serving large one at a time")
    with z.Server(5 << 30, 4) as srv:
        ids = [srv.submit(c, 64, 500, s) for c, s in plan]
        got = [srv.wait(i) for i in ids]
    for g, a in zip(got, alone):
        assert g.raw_counts == a
    big = [g for g, (c, _) in zip(got, plan) if c is large]
    for a in big:
        for b in big:
            if a.id < b.id:
                assert a.t_end <= b.t_start or b.t_end <= a.t_start, "two large requests overlapped"
    lit = [g for g, (c, _) in zip(got, plan) if c is small]
    assert any(s.t_start < b.t_end and b.t_start < s.t_end for s in lit for b in big), "small ones waited"


def test_serving_admission_limits():
    with z.Server(64 << 20, 2) as srv:
        big = z.compile(z.random_circuit(23, 2, 1).to_program()[0], 23)
        with pytest.raises(z.MemoryBudget):
            srv.submit(big, 64, 10, 1)  # 128 MiB state > 64 MiB pool
        small = z.compile(z.random_circuit(21, 3, 1).to_program()[0], 21)
        ids = [srv.submit(small, 64, 0, k) for k in range(3)]  # 32 MiB each: at most 2 resident
        for i in ids:
            r = srv.wait(i)
            assert r.raw_counts == {}


def test_autotune_and_installed_defaults_keep_results(orc, tmp_path):
    """The autotuner's choices (and any installed knob defaults) leave every
    result bit-identical: only launch shapes change."""
    n = 14
    entry = z.tune.autotune(n, 64, depth=3, reps=1, path=str(tmp_path / "t.json"),
                            search={"tile_bits": [13], "slots": [4], "prefetch": [0]}, fusions=[1, 2])
    assert set(entry) >= {"knobs", "fusion", "ms"} and entry["fusion"] in (1, 2)
    prog = orc.gen_matching(n, 8, 3, 3)
    o = orc.State(n, 64)
    o.apply_compiled(orc.Compiled(prog, n, fusion=1))
    try:
        for knobs in (entry["knobs"], {"tile_bits": 13, "slots": 4, "layout": 5, "win_grid": 3}):
            z.tune.install(n, 64, knobs)
            with z.HipBackend(n, 64) as g:
                g.apply_compiled(z.compile(prog, n, fusion=1))
                _assert_same(g, o, f"installed {knobs}")
            r = z.run_program(orc.program([prog, orc.measure_all(n)]), n,
                              opts=z.RunOptions(shots=500, seed=3, precision=64))
            want, _ = orc.run(orc.Compiled(orc.program([prog, orc.measure_all(n)]), n, fusion=1), 500, 3)
            assert r.raw_counts == want
    finally:
        z.tune.uninstall(n, 64)


def test_qasm_programs_run_like_the_oracle(orc):
    """QASM front end -> compiler -> HIP backend, counts against the oracle
    on the same parsed rows (static, dynamic and user-gate programs)."""
    hdr = 'OPENQASM 2.0;\ninclude "qelib1.inc";\n'
    progs = [
        hdr + "qreg q[14];\ncreg c[14];\nh q[0];\n" + "".join(f"cx q[{i}],q[{i + 1}];\n" for i in range(13)) +
        "measure q -> c;\n",
        hdr + "qreg q[13];\ncreg c[13];\ngate cphase(t) a, b { cp(t) a, b; }\nh q;\n" +
        "".join(f"cphase(pi/{2 ** (j % 5 + 1)}) q[{j}], q[{(j + 5) % 13}];\n" for j in range(12)) +
        "u3(0.3, 0.2, 0.1) q;\nmeasure q -> c;\n",
        hdr + "qreg q[3];\ncreg a[1];\ncreg b[1];\ncreg r[1];\nx q[0];\nh q[1];\ncx q[1],q[2];\ncx q[0],q[1];\n"
        "h q[0];\nmeasure q[0] -> a[0];\nmeasure q[1] -> b[0];\nif (b == 1) x q[2];\nif (a == 1) z q[2];\n"
        "measure q[2] -> r[0];\n",
    ]
    for i, src in enumerate(progs):
        _say(f"qasm program {i}")
        p = z.qasm.parse(src)
        want, _ = orc.run(orc.Compiled(p.rows, p.n_qubits, cregs=p.cregs, fusion=1), 3000, 5 + i)
        got = z.qasm.run_qasm(src, z.RunOptions(shots=3000, seed=5 + i, precision=64))
        assert got.raw_counts == want
