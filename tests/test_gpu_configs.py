"""GPU tests at the benchmark configurations and on the paths that the
smaller parity tests do not reach, through the C ABI.

- The precision configuration: random-matching circuits with RZZ ladders
  every 5th layer at 30 and 33 qubits, depth 40.  Exact f64 at fusion 1 is
  the reference there (bit-identical to the CPU path at every size the CPU
  can check); f32 sweeps and fused 5-qubit blocks on the matrix cores are
  held to stated bounds against it, and f64 to its norm and to the adjoint
  returning |0...0>.
- The per-shot path executed one shot at a time in tiled mode, and the
  batched launch above 20 qubits and with fused blocks of more than 3
  qubits, against the oracle shot for shot.
- The HBM precision policy forcing f32, re-bound handles carrying no knob
  of an earlier shape, noisy requests served from the pool, and all 4096
  trajectories of the noisy-trajectory configuration against fixtures the
  oracle wrote.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import zeno_amd as z
from zeno_amd import _lib

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _adjoint(g, comp):
    for op in reversed(comp.ops()):
        if op["kind"] == z.exec.OP_U:
            g.apply_unitary(op["qubits"], op["mat"].conj().T)
        elif op["kind"] == z.exec.OP_D:
            g.apply_diagonal(op["qubits"], op["diag"].conj())
        else:
            g.apply_cx(*op["qubits"])


def _run(n, prec, comp, knobs=None):
    g = z.HipBackend(n, prec)
    for k, v in (knobs or {}).items():
        g.tune(k, v)
    g.apply_compiled(comp)
    g.finish()
    return g


def test_precision_config_30q_depth40():
    """30 qubits, depth 40, seed 7, RZZ ladder every 5th layer: f32 bulk
    within 1e-5 of exact f64; fusion K=5 on the matrix cores within 1e-5
    (f32) and 1e-9 (f64); the f64 norm within 1e-12; forward then adjoint
    in f64 returns |0...0> (every P(q=1) <= 1e-18)."""
    n = 30
    prog = z.matching_circuit(n, 40, 7, rzz_period=5).to_program()[0]
    c1 = z.compile(prog, n, fusion=1)
    c5 = z.compile(prog, n, fusion=5)
    assert max(len(o["qubits"]) for o in c5.ops() if o["kind"] == z.exec.OP_U) == 5  # 5-qubit dense blocks
    _say("30q d40 f64 fusion 1")
    ref = _run(n, 64, c1)
    nrm = ref.norm_sqr()
    _say(f"  norm - 1 = {nrm - 1:.3e}")
    assert abs(nrm - 1.0) <= 1e-12
    errs = {}
    for label, prec, comp, knobs, tol in (("f32 fusion 1", 32, c1, {}, 1e-5),
                                          ("f32 fusion 5 mfma", 32, c5, {"mfma": 1}, 1e-5),
                                          ("f64 fusion 5 mfma", 64, c5, {"mfma": 1}, 1e-9)):
        _say(f"30q d40 {label}")
        with _run(n, prec, comp, knobs) as g:
            errs[label] = ref.max_abs_diff(g)
        _say(f"  max |d amp| vs exact f64 = {errs[label]:.3e}")
        assert errs[label] <= tol, (label, errs[label])
    assert errs["f32 fusion 1"] > 0.0  # the f32 path is really f32
    _say("30q d40 f64 adjoint")
    _adjoint(ref, c1)
    p1 = max(ref.prob_one(q) for q in range(n))
    _say(f"  max P(q=1) after the adjoint = {p1:.3e}")
    assert p1 <= 1e-18
    assert abs(ref.norm_sqr() - 1.0) <= 1e-12
    ref.close()


@pytest.mark.parametrize("n,prec", [(30, 64), (29, 32), (28, 64)])
def test_sampler_pass1_walks_8_chunks_per_workgroup_bit_exact(n, prec):
    """From 4096 chunks sampler pass 1 walks 8 chunks per workgroup; its
    chunk sums and a 1000-shot sample equal the one-chunk-per-workgroup
    walk's (chunk_multi=0), which equals the oracle's (27-qubit test)."""
    prog = z.matching_circuit(n, 2, 11 + n, rzz_period=0).to_program()[0]
    with _run(n, prec, z.compile(prog, n, fusion=1)) as g:
        _say(f"sampler pass 1 n={n} f{prec}")
        multi, draws = g.chunk_sums().view(np.uint64), g.sample(1000, 21)
        g.tune("chunk_multi", 0)
        assert np.array_equal(multi, g.chunk_sums().view(np.uint64))
        assert np.array_equal(draws, g.sample(1000, 21))


@pytest.mark.parametrize("prec", [64, 32])
def test_fused_blocks_in_windows_on_the_spread_layout_bit_exact(prec):
    """26-qubit brickwork at fusion 2, 3 and 4 on the default (spread)
    layout, where a block's qubits land on state bits out of order: the
    window stages order their register slots by each block's matrix bits
    (4-qubit blocks with 4 slots), so every block runs inside a window, and
    the state equals the per-op exact kernels' bit for bit."""
    n = 26
    prog = z.random_circuit(n, 8, 26).to_program()[0]
    for fusion in (2, 3, 4):
        comp = z.compile(prog, n, fusion=fusion)
        _say(f"fused blocks n={n} f{prec} fusion={fusion}")
        with z.HipBackend(n, prec, z.HipBackend.TILED) as g, z.HipBackend(n, prec, z.HipBackend.EXACT) as e:
            assert g.layout() != list(range(n))  # the spread layout
            g.profile_begin()
            g.apply_compiled(comp)
            kinds = {name for name, _, _ in g.profile_end()}
            assert not kinds & {"k_kq<3>", "k_kq<4>"}, kinds  # 3- and 4-qubit blocks run in windows
            e.apply_compiled(comp)
            assert g.max_abs_diff(e) == 0.0


def test_precision_config_33q_depth40():
    """33 qubits, depth 40 (128 GiB f64 and 64 GiB f32 resident together):
    f32 bulk within 1e-5 of exact f64, f64 norm within 1e-12."""
    n = 33
    free, _ = z.device_memory()
    if free < (1 << n) * 24 + (8 << 30):
        pytest.skip("not enough device memory for 33-qubit f64 and f32 states at once")
    prog = z.matching_circuit(n, 40, 7, rzz_period=5).to_program()[0]
    c1 = z.compile(prog, n, fusion=1)
    _say("33q d40 f64 fusion 1")
    with _run(n, 64, c1) as ref:
        nrm = ref.norm_sqr()
        _say(f"  norm - 1 = {nrm - 1:.3e}")
        assert abs(nrm - 1.0) <= 1e-12
        _say("33q d40 f32 fusion 1")
        with _run(n, 32, c1) as g:
            err = ref.max_abs_diff(g)
        _say(f"  max |d amp| vs exact f64 = {err:.3e}")
        assert 0.0 < err <= 1e-5


def _shot_words(orc, prog, n, cregs, fusion, prec, shots, seed, nz, threads=1):
    return orc.run_shots(orc.Compiled(prog, n, cregs=cregs, fusion=fusion), shots, seed, prec, nz, threads)


def _counts(words):
    v, c = np.unique(np.asarray(words, np.uint64), return_counts=True)
    return {int(a): int(b) for a, b in zip(v, c)}


@pytest.mark.parametrize("prec", [64, 32])
def test_per_shot_loop_in_tiled_mode_with_fused_blocks(orc, prec):
    """A 13-qubit dynamic circuit (mid-circuit measure, reset, `if`)
    compiled at fusion 4.  With the per-shot loop on the host (zr_run mode
    2) every shot runs through zb_* calls in tiled mode — windows, held-back
    CX ops, and the prob_one / measure flushes between them — and the counts
    equal the oracle's; the batched launch (its wide instance: 4-qubit
    blocks) gives the oracle's word for every shot."""
    n = 13
    I = orc.instr
    rows = [orc.gen_random(n, 4, 41), orc.measure(3, 0), orc.reset(5), I("h", [5]),
            I("rx", [7], [0.9], cond=(0, 1)), orc.gen_random(n, 3, 42), orc.measure(8, 1), orc.reset(8),
            I("cx", [2, 9], cond=(1, 0)), orc.gen_random(n, 2, 43)]
    rows += [orc.measure(q, 2 + q) for q in range(n)]
    prog = orc.program(rows)
    cregs = [2, n]
    zc = z.compile(prog, n, cregs, fusion=4)
    assert zc.dynamic and z.trajectories_supported(zc)
    assert max(len(o["qubits"]) for o in zc.ops() if o["kind"] == z.exec.OP_U) == 4
    words = _shot_words(orc, prog, n, cregs, 4, prec, 12, 21, None)
    _say(f"per-shot tiled 13q dynamic fusion 4 prec={prec}")
    got = z.run_compiled(zc, z.RunOptions(shots=12, seed=21, precision=prec, mode=2))
    assert got.raw_counts == _counts(words)
    _say(f"batched 13q dynamic fusion 4 prec={prec}")
    bw, _ = z.run_trajectories(zc, 12, 21, prec)
    assert np.array_equal(bw, words)


def test_per_shot_loop_above_20_qubits(orc):
    """A 21-qubit noisy circuit with 4 shots: fewer concurrent states than
    CUs, so zr_run keeps the per-shot loop (mode 1) as mode 2 does, and the
    counts equal the oracle's."""
    n = 21
    nz = {"depolarizing_1q": 0.02, "depolarizing_2q": 0.05, "amplitude_damping": 0.03,
          "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}
    prog = orc.program([orc.gen_random(n, 2, 210), orc.measure_all(n)])
    zc = z.compile(prog, n, fusion=0)
    want = _counts(_shot_words(orc, prog, n, None, 0, 32, 4, 8, nz, threads=4))
    for mode in (1, 2):
        _say(f"per-shot tiled 21q noisy f32 mode={mode}")
        got = z.run_compiled(zc, z.RunOptions(shots=4, seed=8, precision=32, noise=z.NoiseModel(**nz), mode=mode))
        assert got.raw_counts == want


@pytest.mark.parametrize("n,prec", [(22, 32), (22, 64), (24, 32)])
def test_batched_trajectories_above_20_qubits(orc, n, prec):
    """The batched launch above 20 qubits, where every probability sum has
    more than 256 blocks (their totals reduced level by level in each
    workgroup's slot): a noisy circuit at fusion 0 and a dynamic one at
    fusion 5 (the wide instance: dense blocks on up to 5 qubits), 16 shots
    each, every shot's word equal to the oracle's."""
    I = orc.instr
    nz = {"depolarizing_1q": 0.02, "depolarizing_2q": 0.05, "amplitude_damping": 0.05, "bit_flip": 0.01,
          "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}
    noisy = orc.program([orc.gen_random(n, 1, 300 + n), I("ccx", [0, n // 2, n - 1]), orc.measure_all(n)])
    dyn = orc.program([orc.gen_random(n, 3, 310 + n), orc.measure(1, 0), orc.reset(n - 2),
                       I("rx", [n - 3], [0.7], cond=(0, 1)), orc.gen_random(n, 2, 320 + n)]
                      + [orc.measure(q, 1 + q) for q in range(n)])
    for label, prog, cregs, fusion, noise in (("noisy", noisy, None, 0, nz), ("dynamic", dyn, [1, n], 5, None)):
        zc = z.compile(prog, n, cregs, fusion=fusion)
        assert z.trajectories_supported(zc)
        _say(f"batched {n}q {label} fusion {fusion} prec={prec}")
        want = _shot_words(orc, prog, n, cregs, fusion, prec, 16, 77, noise, threads=16)
        got, secs = z.run_trajectories(zc, 16, 77, prec, z.NoiseModel(**noise) if noise else None)
        _say(f"  {secs:.3f} s")
        assert np.array_equal(got, want), np.nonzero(got != want)[0][:8]


def test_served_mode_choice_above_20_qubits(orc):
    """512 noisy 22-qubit shots: zr_run takes the batched launch (one
    trajectory per CU and more) in mode 1 and the per-shot loop in mode 2;
    both give the oracle's counts."""
    n, shots = 22, 512
    nz = {"amplitude_damping": 0.02, "readout_flip_0to1": 0.03}
    prog = orc.program([orc.gen_random(n, 1, 400), orc.measure_all(n)])
    zc = z.compile(prog, n, fusion=0)
    want = _counts(_shot_words(orc, prog, n, None, 0, 32, shots, 9, nz, threads=16))
    for mode in (1, 2):
        _say(f"zr_run 22q noisy f32 {shots} shots mode={mode}")
        got = z.run_compiled(zc, z.RunOptions(shots=shots, seed=9, precision=32, noise=z.NoiseModel(**nz), mode=mode))
        _say(f"  {got.sim_time:.3f} s")
        assert got.raw_counts == want


def test_auto_precision_falls_back_to_f32_on_the_hbm_budget(orc):
    """resolve_precision through zr_resolve_precision: with a budget one
    byte short of the f64 state the run takes f32 with the reference's
    notice, and its counts are the oracle's f32 counts."""
    n = 16
    prog = orc.program([orc.gen_random(n, 10, 160), orc.measure_all(n)])
    opts = z.RunOptions(shots=2000, seed=3, mem_limit=z.state_bytes(n, 64) - 1)
    r = z.run_program(prog, n, opts=opts)
    assert r.precision == 32 and r.backend == "hip-f32"
    assert "auto-precision: 16 qubits exceed the f64 budget, using f32 (8 B/amplitude)" in r.notices
    want, _ = orc.run(orc.Compiled(prog, n, fusion=1), 2000, 3, prec=32)
    assert r.raw_counts == want
    r64 = z.run_program(prog, n, opts=z.RunOptions(shots=2000, seed=3, mem_limit=z.state_bytes(n, 64)))
    assert r64.precision == 64 and not r64.notices
    with pytest.raises(z.MemoryBudget) as ei:
        z.run_program(prog, n, opts=z.RunOptions(shots=10, seed=3, mem_limit=z.state_bytes(n, 32) - 1))
    assert ei.value.max_qubits == n - 1


def _bits(a):
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def test_rebind_resets_knobs_of_the_earlier_shape(orc):
    """zb_rebind: a handle created for a shape with an installed fma default
    and then re-bound to another shape runs exact arithmetic again (state
    bits equal the oracle's), while the installed shape still gets fma."""
    L = _lib.lib()
    n_a, n_b = 14, 15
    prog = orc.gen_random(n_b, 6, 5)
    o = orc.State(n_b, 64)
    o.apply_compiled(orc.Compiled(prog, n_b, fusion=1))
    assert L.zb_tune_default(n_a, 64, b"fma", 1) == 0
    try:
        with z.HipBackend(n_b, 64) as mem:  # device memory for the re-bound state
            re, im = C.c_void_p(), C.c_void_p()
            assert L.zb_state_device_ptrs(mem.handle, C.byref(re), C.byref(im)) == 0
            g = z.HipBackend(n_a, 64)
            try:
                assert L.zb_rebind(g.handle, n_b, 64, re, im) == 0, _lib.last_error()
                g.n = n_b
                g.apply_compiled(z.compile(prog, n_b, fusion=1))
                gr, gi = g.raw()
                orr, oi = o.raw()
                assert np.array_equal(_bits(gr), _bits(orr)) and np.array_equal(_bits(gi), _bits(oi))
            finally:
                g.close()
        # the installed default is live for its own shape: fma results differ
        pa = orc.gen_random(n_a, 6, 6)
        oa = orc.State(n_a, 64)
        oa.apply_compiled(orc.Compiled(pa, n_a, fusion=1))
        with z.HipBackend(n_a, 64) as ga:
            ga.apply_compiled(z.compile(pa, n_a, fusion=1))
            gr, _ = ga.raw()
            orr, _ = oa.raw()
            assert not np.array_equal(_bits(gr), _bits(orr))
            assert np.max(np.abs(gr - orr)) <= 1e-12
    finally:
        L.zb_tune_default(n_a, 64, b"fma", -(1 << 63))


def test_served_noisy_requests_run_in_the_pool(orc):
    """Noisy and dynamic requests take their trajectory workspace from the
    serving pool: with a pool that holds only part of the default grid the
    server still runs them (fewer concurrent trajectories), alongside
    static ones, and every result equals the request run alone and the
    oracle."""
    nz = {"depolarizing_1q": 0.01, "depolarizing_2q": 0.03, "amplitude_damping": 0.02,
          "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}
    cases = []
    for i, (n, prec) in enumerate(((12, 32), (14, 64), (13, 32), (15, 64), (11, 64), (16, 32))):
        prog = orc.program([orc.gen_random(n, 5, 500 + i), orc.measure_all(n)])
        cases.append((prog, n, prec, 256, 900 + i, nz if i % 3 != 2 else None))
    want = []
    for prog, n, prec, shots, seed, noise in cases:
        oc = orc.Compiled(prog, n, fusion=0 if noise else 1)
        if noise:
            want.append(_counts(orc.run_shots(oc, shots, seed, prec, noise, 8)))
        else:
            want.append(orc.run(oc, shots, seed, prec=prec)[0])
    for pool, slots in ((96 << 20, 3), (2 << 30, 4)):
        _say(f"served noisy requests pool={pool >> 20} MiB slots={slots}")
        with z.Server(pool, slots) as srv:
            ids = [srv.submit(z.compile(p, n, fusion=0 if nzz else 1), prec, shots, seed,
                              z.NoiseModel(**nzz) if nzz else None)
                   for p, n, prec, shots, seed, nzz in cases]
            got = [srv.wait(i) for i in ids]
        for g, w, case in zip(got, want, cases):
            assert g.raw_counts == w, case[1:3]


def test_all_4096_noisy_trajectories_equal_the_oracle_fixture():
    """The noisy-trajectory configuration (16 qubits, depth 30, brickwork,
    p1 = 1e-3, p2 = 1e-2, damping 5e-3, readout 2%, seed 2024 circuit,
    shot seed 99): every one of the 4096 per-shot words equals the
    oracle's, in f32 and f64 (fixtures written by tests/golden/
    make_fixtures.py with the shot-parallel oracle)."""
    nz = {"depolarizing_1q": 1e-3, "depolarizing_2q": 1e-2, "amplitude_damping": 5e-3,
          "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}
    prog = z.Circuit(16).extend(z.random_circuit(16, 30, 2024).to_program()[0]).measure_all().to_program()[0]
    comp = z.compile(prog, 16, fusion=0)
    for prec in (32, 64):
        want = np.load(os.path.join(GOLDEN, f"trajectories_16q_d30_f{prec}.npy"))
        _say(f"4096 trajectories 16q d30 prec={prec}")
        words, _ = z.run_trajectories(comp, len(want), 99, prec, z.NoiseModel(**nz))
        bad = np.nonzero(words != want)[0]
        assert bad.size == 0, f"{bad.size} shots differ, first {bad[:8].tolist()}"


def _small_runs(orc, count, seed0):
    """Uneven small static circuits: 1..13 qubits in f64, 1..14 in f32,
    random brickwork, matching circuits with RZZ ladders (diagonals),
    fusion 1..3, shot counts from 1 to 5000, one circuit without
    measurements."""
    rng = np.random.default_rng(seed0)
    runs = []
    for i in range(count):
        prec = 64 if i % 2 else 32
        n = int(rng.integers(1, (13 if prec == 64 else 14) + 1))
        depth = int(rng.integers(1, 12))
        if n >= 2 and i % 3 == 0:
            body = orc.gen_matching(n, depth, 40 + i, 2)
        else:
            body = orc.gen_random(n, depth, 40 + i)
        measured = i != 5
        prog = orc.program([body, orc.measure_all(n)]) if measured else body
        fusion = int(rng.integers(1, 4))
        shots = int(rng.choice([1, 7, 1000, 5000]))
        runs.append((prog, n, prec, fusion, shots, 7000 + i))
    return runs


def test_batched_small_circuits_match_the_oracle_and_solo_runs(orc):
    """zr_run_batch: 48 uneven small circuits as instances of one launch per
    precision; each instance's counts equal the oracle's and zr_run's for the
    same circuit, precision, shots and seed.  The batch runs twice (its
    captured launch replays) with the same results."""
    runs = _small_runs(orc, 48, 3)
    comps = [z.compile(p, n, fusion=f) for p, n, prec, f, s, sd in runs]
    for c, (p, n, prec, f, s, sd) in zip(comps, runs):
        assert z.batch_supported(c, prec, s), (n, prec, f)
    args = [(c, prec, s, sd) for c, (p, n, prec, f, s, sd) in zip(comps, runs)]
    _say("batch of 48 small circuits")
    got, secs = z.run_batch(args)
    _say(f"  {secs * 1e3:.2f} ms")
    again, _ = z.run_batch(args)
    assert again == got
    for g, c, (p, n, prec, f, s, sd) in zip(got, comps, runs):
        want, _ = orc.run(orc.Compiled(p, n, fusion=f), s, sd, prec=prec)
        assert g == want, (n, prec, f, s)
        solo = z.run_compiled(c, z.RunOptions(shots=s, seed=sd, precision=prec))
        assert g == solo.raw_counts, (n, prec, f, s)


def test_batch_limits_are_enforced():
    """Outside the batch's range (noisy or dynamic circuits, 14 qubits in
    f64, dense ops on 4 qubits in f64, more than 2^17 shots) zr_run_batch
    refuses; those requests take the server's slots instead."""
    big = z.compile(z.Circuit(14).extend(z.random_circuit(14, 2, 1).to_program()[0]).measure_all()
                    .to_program()[0], 14, fusion=1)
    assert z.batch_supported(big, 32, 1000) and not z.batch_supported(big, 64, 1000)
    k4 = z.compile(z.random_circuit(8, 6, 2).to_program()[0], 8, fusion=4)
    assert any(len(o["qubits"]) == 4 for o in k4.ops() if o["kind"] == z.exec.OP_U)
    assert z.batch_supported(k4, 32, 10) and not z.batch_supported(k4, 64, 10)
    small = z.compile(z.random_circuit(6, 3, 2).to_program()[0], 6)
    assert not z.batch_supported(small, 64, (1 << 17) + 1)
    with pytest.raises(z.Unsupported):
        z.run_batch([(big, 64, 10, 1)])


def test_served_small_requests_are_batched(orc):
    """The server's batcher: a burst of small static requests beside larger
    ones runs as a few batched launches (the captured launch replayed), and
    every result equals the request run alone; with batching off
    (max_batch 0) the same requests go through the slots with the same
    results."""
    runs = _small_runs(orc, 40, 11)
    rows = [(z.compile(p, n, fusion=f), prec, s, sd) for p, n, prec, f, s, sd in runs]
    for i, n in enumerate((18, 20, 16)):  # larger ones take the slots meanwhile
        prog = orc.program([orc.gen_random(n, 8, 90 + i), orc.measure_all(n)])
        rows.append((z.compile(prog, n, fusion=1), 64, 1000, 95 + i))
    alone = [z.run_compiled(c, z.RunOptions(shots=s, seed=sd, precision=p)).raw_counts for c, p, s, sd in rows]
    for batch in (1024, 0):
        _say(f"served small requests batch={batch}")
        with z.Server(1 << 30, 3, batch=batch) as srv:
            for _rep in range(2):  # a second burst of the same shapes: the batch launch replays
                ids = [srv.submit(c, p, s, sd) for c, p, s, sd in rows]
                got = [srv.wait(i) for i in ids]
                for g, a in zip(got, alone):
                    assert g.raw_counts == a
            cnt = srv.counters()
        _say(f"  counters {cnt}")
        if batch:
            assert cnt["batched"] == 2 * len(runs) and 1 <= cnt["batches"] < 2 * len(runs)
            assert cnt["graph_replays"] >= 1
        else:
            assert cnt["batches"] == 0 and cnt["batched"] == 0


def _checksums(g, n):
    return [g.norm_sqr()] + [g.prob_one(q) for q in range(n)]


@pytest.mark.parametrize("n,prec", [(34, 64), (35, 32)])
def test_largest_states_tiled(n, prec):
    """The largest states one MI355X holds (34 qubits f64, 35 f32: 256 GiB)
    run in tiled windows: a depth-3 random-matching circuit with an RZZ
    ladder gives bit-identical norm and per-qubit probabilities (fixed-order
    f64 sums over the whole state) in tiled windows and in one exact sweep
    per op, and forward then adjoint returns |0...0>."""
    free, _ = z.device_memory()
    if free < (1 << n) * (16 if prec == 64 else 8) + (4 << 30):
        pytest.skip(f"not enough device memory for a {n}-qubit f{prec} state")
    comp = z.compile(z.matching_circuit(n, 3, 17, rzz_period=2).to_program()[0], n, fusion=1)
    with z.HipBackend(n, prec, z.HipBackend.TILED) as g:
        _say(f"{n}q f{prec} tiled")
        g.apply_compiled(comp)
        tiled = _checksums(g, n)
        _say(f"  norm - 1 = {tiled[0] - 1:.3e}")
        g.reset_all()
        g.set_mode(z.HipBackend.EXACT)
        _say(f"{n}q f{prec} exact")
        g.apply_compiled(comp)
        exact = _checksums(g, n)
        assert [x.hex() for x in tiled] == [x.hex() for x in exact]
        g.set_mode(z.HipBackend.TILED)
        _say(f"{n}q f{prec} adjoint")
        _adjoint(g, comp)
        p1 = max(g.prob_one(q) for q in range(n))
        _say(f"  max P(q=1) after the adjoint = {p1:.3e}")
        assert p1 <= (1e-18 if prec == 64 else 1e-8)
        assert abs(g.norm_sqr() - 1.0) <= (1e-12 if prec == 64 else 1e-4)
