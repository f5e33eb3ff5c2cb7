"""Pins for the CPU oracle (oracle/zeno_oracle.c) — no GPU needed.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign, swapped operand, wrong bit order) fails: published generator
test vectors, hand-coded standard gate matrices (tests/reference.rs:157-283
restated in numpy), an independent numpy tensor-contraction simulator,
closed forms (QFT = bit-reversed DFT, GHZ, Bell), the reference's compiler
and kernel KATs, brute-force inverse-CDF sampling and 6-sigma statistics.
"""
import math

import numpy as np
import pytest

from circuits import diag_ladder, disjoint_blocks, max_delta, random_mixed
from refgates import NATIVE, TestRng, kron_apply, ref_gate, ref_of_program, u3m

TAU = 2 * math.pi


# ------------------------------------------------------------------ RNG

def test_xoshiro256pp_published_vector(orc):
    # rand_xoshiro 0.6.0 test vector for Xoshiro256PlusPlus, state [1,2,3,4].
    want = [41943041, 58720359, 3588806011781223, 3591011842654386,
            9228616714210784205, 9973669472204895162, 14011001112246962877,
            12406186145184390807, 15849039046786891736, 10450023813501588000]
    assert [int(x) for x in orc.xoshiro_from_state([1, 2, 3, 4], 10)] == want


def test_splitmix64_reference_output(orc):
    # splitmix64.c reference: first output for state 1477776061723855037.
    assert orc.splitmix64(1477776061723855037) == 1985237415132408290


def test_seed_from_u64_uses_splitmix_stream(orc):
    # seed_from_u64(s): state = 4 consecutive SplitMix64 outputs from s.
    M = (1 << 64) - 1
    for seed in (0, 7, 42, 2**63 + 5):
        st, x = [], seed
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & M
            st.append(orc.splitmix64((x - 0x9E3779B97F4A7C15) & M))
        assert list(orc.rng_u64(seed, 6)) == list(orc.xoshiro_from_state(st, 6))


def test_gen_f64_range_and_moments(orc):
    u = orc.rng_f64(123, 200_000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 6 * math.sqrt(1 / 12 / len(u))
    raw = orc.rng_u64(123, 4)
    # 53 high bits scaled by 2^-53 (the top bit of the 53 must matter)
    assert np.all(u[:4] == (raw >> np.uint64(11)).astype(np.float64) / 2.0**53)


def _state_with_first_output(v, s0=0x243F6A8885A308D3, s1=0x13198A2E03707344, s2=0xA4093822299F31D0):
    """A xoshiro256++ state whose first output is v: the output is
    rotl(s0 + s3, 23) + s0, so s3 = rotr(v - s0, 23) - s0."""
    M = (1 << 64) - 1
    x = (v - s0) & M
    s3 = (((x >> 23) | (x << 41)) & M) - s0
    return [s0, s1, s2, s3 & M]


def test_gen_range_on_the_rejection_zone_edges(orc):
    """rand 0.8 UniformInt<u64>::sample_single(lo, hi) (uniform.rs,
    sample_single_inclusive(lo, hi - 1)): the widening product v * range
    splits into (hi64, lo64); lo64 <= zone accepts and returns lo + hi64,
    where zone = (range << leading_zeros(range)) - 1; otherwise the next
    output is drawn.  States crafted so the first output sits exactly on
    the zone's last accepted value and on the first rejected one catch an
    off-by-one in the zone, a strict comparison, the low word returned in
    place of the high one, or a missing redraw."""
    M = (1 << 64) - 1
    for lo, hi in ((1, 4), (1, 16), (1, 64), (5, 8)):
        r = hi - lo
        zone = ((r << (64 - r.bit_length())) & M) - 1
        inv = pow(r, -1, 1 << 64)
        for lo64, accepted in ((zone, True), (zone + 1, False), (0, True), (M, False)):
            v = (lo64 * inv) & M
            assert (v * r) & M == lo64
            st = _state_with_first_output(v)
            raw = [int(x) for x in orc.xoshiro_from_state(st, 8)]
            assert raw[0] == v
            got = int(orc.rng_range_from_state(st, lo, hi, 1)[0])
            if accepted:
                assert got == lo + ((v * r) >> 64), (lo, hi, lo64)
            else:
                nxt = next(w for w in raw[1:] if (w * r) & M <= zone)
                assert got == lo + ((nxt * r) >> 64), (lo, hi, lo64)


def test_gen_range_zone_rule_is_unbiased_by_brute_force():
    """The zone rule (accept lo <= (range << lz(range)) - 1) on 8-bit words,
    every range and every word: each output value is reached by the same
    number of accepted words, the property that fixes the rule."""
    for r in range(1, 256):
        zone = ((r << (8 - r.bit_length())) & 255) - 1
        counts = [0] * r
        for v in range(256):
            if (v * r) & 255 <= zone:
                counts[(v * r) >> 8] += 1
        assert min(counts) == max(counts) > 0, r


def test_gen_range_uniform(orc):
    for hi in (4, 16, 64):
        v = orc.rng_range(99, 1, hi, 60_000)
        assert v.min() >= 1 and v.max() < hi
        cnt = np.bincount(v.astype(np.int64), minlength=hi)[1:]
        p = 1.0 / (hi - 1)
        sig = math.sqrt(len(v) * p * (1 - p))
        assert np.all(np.abs(cnt - len(v) * p) < 6 * sig)


# ---------------------------------------------------------------- gates

def _dense(orc, name, params, arity):
    m, isd = orc.gate(name, params)
    return np.diag(m) if isd else m


def test_gate_library_matches_hand_coded(orc):
    rng = TestRng(0x51DECAFE)
    for name, arity, npar in NATIVE:
        for _ in range(3):
            params = [rng.angle() for _ in range(npar)]
            lib = _dense(orc, name, params, arity)
            mine = ref_gate(name, params)
            assert max_delta(lib.ravel(), mine.ravel()) <= 1e-14, name
            assert np.allclose(lib @ lib.conj().T, np.eye(1 << arity), atol=1e-13), name
    assert orc.gate("nope") is None
    assert orc.gate("rx", []) is None  # wrong param count


def test_diagonal_gates_are_tagged(orc):
    diag = {"id", "u0", "z", "s", "sdg", "t", "tdg", "rz", "u1", "p", "cz", "cp", "cu1", "crz", "rzz"}
    for name, arity, npar in NATIVE:
        _, isd = orc.gate(name, [0.3] * npar)
        assert isd == (name in diag), name


def test_kron_simulator_is_sound():
    # The numpy tensor-contraction simulator against explicit Kronecker
    # products (the independent library route).
    n = 4
    rng = np.random.default_rng(3)
    psi = rng.normal(size=16) + 1j * rng.normal(size=16)
    m = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    # qs = [1, 3] in argument order: local bit0 = q1, bit1 = q3
    got = kron_apply(psi, n, [1, 3], m)
    full = np.zeros((16, 16), complex)
    for i in range(16):
        for j in range(16):
            if (i & 0b0101) != (j & 0b0101):
                continue
            li = ((i >> 1) & 1) | (((i >> 3) & 1) << 1)
            lj = ((j >> 1) & 1) | (((j >> 3) & 1) << 1)
            full[i, j] = m[li, lj]
    assert np.allclose(got, full @ psi, atol=1e-13)


def test_refsim_matches_numpy(orc):
    n = 5
    rng = TestRng(77)
    for name, arity, npar in NATIVE:
        qs = []
        while len(qs) < arity:
            q = rng.below(n)
            if q not in qs:
                qs.append(q)
        params = [rng.angle() for _ in range(npar)]
        psi = np.random.default_rng(5).normal(size=32) + 1j * np.random.default_rng(6).normal(size=32)
        a = psi.copy()
        orc.refsim_apply(a, n, qs, ref_gate(name, params))
        b = kron_apply(psi, n, qs, ref_gate(name, params))
        assert max_delta(a, b) < 1e-13, (name, qs)


# -------------------------------------------------------------- circuits

def _check_levels(orc, prog, n, want, what, tol=1e-10):
    for fusion in range(7):
        sv = orc.statevector(prog, n, fusion=fusion)
        norm = math.sqrt(float(np.sum(np.abs(sv) ** 2)))
        assert abs(norm - 1) < 1e-12, f"{what} fusion={fusion}: norm {norm}"
        d = max_delta(sv, want)
        assert d < tol, f"{what} fusion={fusion}: max |d| = {d:.3e}"


def test_random_circuits_all_fusion_levels(orc):
    for n in range(2, 11):
        prog = orc.gen_random(n, 20, 1000 + n)
        _check_levels(orc, prog, n, ref_of_program(prog, n), f"random n={n}")


def test_mixed_circuits_all_fusion_levels(orc):
    for n in range(2, 11):
        prog = random_mixed(orc, n, 60, 9000 + n)
        _check_levels(orc, prog, n, ref_of_program(prog, n), f"mixed n={n}")


def test_diagonal_ladder_and_forced_flushes(orc):
    p = diag_ladder(orc)
    _check_levels(orc, p, 6, ref_of_program(p, 6), "diag ladder")
    p = disjoint_blocks(orc)
    _check_levels(orc, p, 8, ref_of_program(p, 8), "disjoint blocks")
    c = orc.Compiled(p, 8, fusion=5)
    assert c.stats["max_fused"] <= 5 and c.stats["output_ops"] >= 2


def test_matching_and_qft_generators(orc):
    n = 9
    p = orc.gen_matching(n, 10, 5, rzz_period=5)
    _check_levels(orc, p, n, ref_of_program(p, n), "matching+rzz")
    names = [r["name"].decode() for r in p]
    assert names.count("rzz") == 2 * (n - 1)
    # each layer's CX pairs are a matching: every qubit at most once
    per_layer = n + n // 2
    idx = 0
    for layer in range(10):
        cx = p[idx + n: idx + per_layer]
        qs = [int(q) for r in cx for q in r["q"][:2]]
        assert len(set(qs)) == len(qs) == 2 * (n // 2)
        idx += per_layer + (n - 1 if layer % 5 == 4 else 0)
    q = orc.gen_u3_qft(6, 3)
    _check_levels(orc, q, 6, ref_of_program(q, 6), "u3+qft")


def test_every_gate_every_argument_order(orc):
    orders = {1: [[0], [2], [3]], 2: [[0, 1], [1, 0], [3, 1], [2, 3], [3, 0]],
              3: [[0, 1, 2], [3, 0, 2], [2, 3, 0], [1, 3, 2]]}
    rng = TestRng(0xA5A50001)
    for name, arity, npar in NATIVE:
        for qs in orders[arity]:
            params = [rng.angle() for _ in range(npar)]
            prep = [[rng.angle() for _ in range(3)] for _ in range(4)]
            rows = [orc.instr("u3", [q], prep[q]) for q in range(4)] + [orc.instr(name, qs, params)]
            prog = orc.program(rows)
            want = ref_of_program(prog, 4)
            for fusion in (0, 5):
                d = max_delta(orc.statevector(prog, 4, fusion=fusion), want)
                assert d < 1e-12, (name, qs, fusion, d)


def test_harness_detects_planted_mistakes(orc):
    # A sign error (S vs S-dagger) and swapped CX arguments planted on the
    # oracle's input side must be visible against the independent simulator.
    good = orc.program([orc.instr("h", [0]), orc.instr("s", [0])])
    bad = orc.program([orc.instr("h", [0]), orc.instr("sdg", [0])])
    want = ref_of_program(good, 1)
    assert max_delta(orc.statevector(good, 1), want) < 1e-12
    assert max_delta(orc.statevector(bad, 1), want) > 0.5
    prep = [orc.instr("u3", [0], [0.9, 0.4, 1.3]), orc.instr("u3", [1], [0.5, 2.1, -0.7])]
    p1 = orc.program(prep + [orc.instr("cx", [0, 1])])
    p2 = orc.program(prep + [orc.instr("cx", [1, 0])])
    assert max_delta(orc.statevector(p1, 2), ref_of_program(p2, 2)) > 0.1


def test_f32_tracks_f64(orc):
    for n in (4, 8, 10):
        for prog in (orc.gen_random(n, 20, 1000 + n), random_mixed(orc, n, 60, 9000 + n)):
            want = ref_of_program(prog, n)
            for fusion in (0, 5):
                d = max_delta(orc.statevector(prog, n, fusion=fusion, prec=32), want)
                assert d < 2e-4


# -------------------------------------------------------------- compiler

def test_compiler_kats(orc):
    I = orc.instr
    c = orc.Compiled(orc.program([I("h", [0]), I("h", [0]), I("x", [1])]), 2, fusion=0)
    assert c.stats["cancelled"] == 2 and len(c) == 1
    c = orc.Compiled(orc.program([I("cx", [0, 1]), I("h", [2]), I("cx", [0, 1])]), 3, fusion=0)
    assert c.stats["cancelled"] == 2 and len(c) == 1
    c = orc.Compiled(orc.program([I("h", [0]), I("cx", [0, 1])] + [orc.measure(0, 0), orc.measure(1, 1)]), 2, fusion=5)
    assert len(c) == 1 and not c.stats["dynamic"] and c.final_measures == [(0, 0), (1, 1)]
    op = c.op(0)
    s = 1 / math.sqrt(2)
    assert op["kind"] == orc.OP_U and op["qubits"] == [0, 1]
    assert abs(op["mat"].ravel()[0] - s) < 1e-12 and abs(op["mat"].ravel()[12] - s) < 1e-12
    c = orc.Compiled(orc.program([I("t", [0]), I("cz", [0, 1]), I("s", [2]), I("rz", [1], [0.3])]), 3, fusion=5)
    assert len(c) == 1 and c.op(0)["kind"] == orc.OP_D and c.op(0)["qubits"] == [0, 1, 2]
    c = orc.Compiled(orc.program([I("h", [0]), orc.measure(0, 0), I("x", [1]), orc.measure(1, 1)]), 2, fusion=5)
    assert c.stats["dynamic"]
    c = orc.Compiled(orc.program([I("h", [0]), orc.barrier(), I("h", [0])]), 2, fusion=5)
    assert len(c) == 2
    with pytest.raises(orc.CompileError):
        orc.Compiled(orc.program([I("foo", [0])]), 2)
    with pytest.raises(orc.CompileError):
        orc.Compiled(orc.program([I("cx", [0, 0])]), 2)
    with pytest.raises(orc.CompileError):
        orc.Compiled(orc.program([I("h", [5])]), 2)


def test_compiler_cx_stays_permutation(orc):
    c = orc.Compiled(orc.gen_random(6, 3, 1), 6, fusion=1)
    kinds = [o["kind"] for o in c.ops()]
    assert kinds.count(orc.OP_CX) == 3 + 2 + 3


# ---------------------------------------------------------------- kernels

def test_state_kats(orc):
    h, _ = orc.gate("h")
    cx, _ = orc.gate("cx")
    st = orc.State(3)
    st.apply_unitary([1], h)
    assert abs(st.prob_one(1) - 0.5) < 1e-12 and st.prob_one(0) < 1e-12
    assert abs(st.norm_sqr() - 1) < 1e-12
    st = orc.State(2)
    st.apply_unitary([0], h)
    st.apply_unitary([0, 1], cx)
    a = st.statevector()
    assert abs(a[0] - 1 / math.sqrt(2)) < 1e-12 and abs(a[3] - 1 / math.sqrt(2)) < 1e-12
    assert st.measure(0, 0.0)  # u=0 < p1 -> outcome 1
    assert abs(st.norm_sqr() - 1) < 1e-12 and abs(abs(st.statevector()[3]) - 1) < 1e-12
    # CX as permutation vs dense, every ordering
    for c, t in [(0, 2), (2, 0), (1, 2), (2, 1)]:
        A, B = orc.State(3), orc.State(3)
        for s in (A, B):
            for q in range(3):
                s.apply_unitary([q], h)
            s.apply_unitary([1], np.diag(orc.gate("t")[0]))
        A.apply_cx(c, t)
        qs = sorted([c, t])
        m = ref_gate("cx") if [c, t] == qs else ref_gate("cx")[np.ix_([0, 2, 1, 3], [0, 2, 1, 3])]
        B.apply_unitary(qs, m)
        assert np.array_equal(A.statevector(), B.statevector())


def test_fixed_order_reductions(orc):
    for n in (1, 5, 13, 17):
        p = random_mixed(orc, n, 3 * n, 5 + n)
        st = orc.State(n)
        st.apply_compiled(orc.Compiled(p, n, fusion=2))
        sv = st.statevector()
        pr = np.abs(sv) ** 2
        assert abs(st.norm_sqr() - math.fsum(pr)) < 1e-14 * (1 << n)
        for q in range(n):
            mask = (np.arange(1 << n) >> q) & 1 == 1
            assert abs(st.prob_one(q) - math.fsum(pr[mask])) < 1e-14 * (1 << n)


# ---------------------------------------------------------------- closed forms

def _bitrev(j, n):
    return int(format(j, f"0{n}b")[::-1], 2)


def test_qft_is_bit_reversed_dft(orc):
    n = 5
    dim = 1 << n
    for x in (0, 1, 5, 19, 31):
        rows = [orc.instr("x", [q]) for q in range(n) if (x >> q) & 1]
        prog = orc.program(rows + [orc.gen_qft(n)])
        for fusion in (0, 1, 5):
            sv = orc.statevector(prog, n, fusion=fusion)
            for j in range(dim):
                want = complex(np.exp(2j * np.pi * x * j / dim)) / math.sqrt(dim)
                assert abs(sv[_bitrev(j, n)] - want) < 1e-12


# ---------------------------------------------------------------- sampler

def _brute_sample(sv, us):
    """Inverse CDF straight from the definition: index i with
    cum(i) > u, cum built serially in index order."""
    p = np.abs(sv) ** 2
    cum = np.cumsum(p)
    return np.searchsorted(cum, us, side="right")


def test_sampler_matches_inverse_cdf(orc):
    for n, seed in ((4, 1), (10, 2), (17, 3), (19, 4)):
        p = random_mixed(orc, n, 4 * n, 700 + n)
        st = orc.State(n)
        st.apply_compiled(orc.Compiled(p, n, fusion=3))
        shots = 3000
        got = st.sample(shots, seed)
        assert np.all(np.diff(got.astype(np.int64)) >= 0), "ascending output"
        total = float(np.sum(st.chunk_sums()))
        us = np.sort(orc.rng_f64(seed, shots) * total)
        want = _brute_sample(st.statevector(), us)
        same = np.mean(got.astype(np.int64) == want)
        if n < 16:
            assert same == 1.0
        else:  # chunk-restarted cumulative sums may round differently at a boundary
            assert same > 0.999


def test_sampler_statistics_and_determinism(orc):
    n = 12
    rows = [orc.instr("h", [0])] + [orc.instr("cx", [q, q + 1]) for q in range(n - 1)]
    st = orc.State(n)
    st.apply_compiled(orc.Compiled(orc.program(rows), n))
    s = st.sample(20000, 11)
    assert set(np.unique(s).tolist()) <= {0, (1 << n) - 1}
    ones = int(np.sum(s == (1 << n) - 1))
    assert abs(ones - 10000) < 6 * math.sqrt(20000 * 0.25)
    assert np.array_equal(s, st.sample(20000, 11))
    assert orc.State(3).sample(0, 1).size == 0


def test_run_counts_bell_and_f32_ghz(orc):
    rows = [orc.instr("h", [0]), orc.instr("cx", [0, 1]), orc.measure(0, 0), orc.measure(1, 1)]
    comp = orc.Compiled(orc.program(rows), 2)
    counts, _ = orc.run(comp, 4096, 1)
    assert set(counts) <= {0, 3} and sum(counts.values()) == 4096
    assert abs(counts.get(0, 0) / 4096 - 0.5) < 0.05
    n = 10
    rows = [orc.instr("h", [0])] + [orc.instr("cx", [q, q + 1]) for q in range(n - 1)]
    comp = orc.Compiled(orc.program(rows + [orc.measure_all(n)]), n)
    a, _ = orc.run(comp, 2000, 3, prec=64)
    b, _ = orc.run(comp, 2000, 3, prec=32)
    assert a == b and set(a) <= {0, (1 << n) - 1}


def test_teleportation_dynamic(orc):
    # exec.rs:686-738: |1> teleported from q0 to q2; cregs m0, m1, out.
    I = orc.instr
    rows = [I("x", [0]), I("h", [1]), I("cx", [1, 2]), I("cx", [0, 1]), I("h", [0]),
            orc.measure(0, 0), orc.measure(1, 1),
            I("x", [2], cond=(1, 1)), I("z", [2], cond=(0, 1)), orc.measure(2, 2)]
    comp = orc.Compiled(orc.program(rows), 3, cregs=[1, 1, 1])
    assert comp.stats["dynamic"]
    counts, _ = orc.run(comp, 512, 5)
    assert sum(counts.values()) == 512
    assert all((k >> 2) & 1 == 1 for k in counts)


# ---------------------------------------------------------------- noise

def _six_sigma(count, shots, p):
    sd = math.sqrt(shots * p * (1 - p))
    assert abs(count - shots * p) <= 6 * sd + 1e-9, (count, shots * p, sd)


def test_noise_closed_forms(orc):
    I = orc.instr
    shots = 4000
    # amplitude damping: X then k identity-like gates; P(1) = (1-g)^(k+1)
    g, k = 0.1, 4
    rows = [I("x", [0])] + [I("rz", [0], [0.0]) for _ in range(k)] + [orc.measure(0, 0)]
    comp = orc.Compiled(orc.program(rows), 1, fusion=0)
    counts, _ = orc.run(comp, shots, 7, noise={"amplitude_damping": g})
    _six_sigma(counts.get(1, 0), shots, (1 - g) ** (k + 1))
    # damping on |+>: P(1) = (1-g)/2
    rows = [I("h", [0]), orc.measure(0, 0)]
    comp = orc.Compiled(orc.program(rows), 1, fusion=0)
    counts, _ = orc.run(comp, shots, 8, noise={"amplitude_damping": g})
    _six_sigma(counts.get(1, 0), shots, (1 - g) / 2)
    # readout flips on |0>: P(record 1) = p01
    rows = [I("id", [0]), orc.measure(0, 0)]
    comp = orc.Compiled(orc.program(rows), 1, fusion=0)
    counts, _ = orc.run(comp, shots, 9, noise={"readout_flip_0to1": 0.2})
    _six_sigma(counts.get(1, 0), shots, 0.2)
    # 1q depolarizing after X: X/Y flip back to 0 w.p. 2p/3
    p = 0.3
    rows = [I("x", [0]), orc.measure(0, 0)]
    comp = orc.Compiled(orc.program(rows), 1, fusion=0)
    counts, _ = orc.run(comp, shots, 10, noise={"depolarizing_1q": p})
    _six_sigma(counts.get(0, 0), shots, 2 * p / 3)
    # 2q depolarizing after CX on |00>: each of 15 Paulis w.p. p/15; the
    # strings flipping qubit 0 (X or Y on bit 0: codes 1,2) are 8 of 15
    rows = [I("cx", [0, 1]), orc.measure(0, 0), orc.measure(1, 1)]
    comp = orc.Compiled(orc.program(rows), 2, fusion=0)
    counts, _ = orc.run(comp, shots, 11, noise={"depolarizing_2q": p})
    q0 = sum(c for k_, c in counts.items() if k_ & 1)
    _six_sigma(q0, shots, p * 8 / 15)


def test_noise_determinism_and_trivial_model(orc):
    comp = orc.Compiled(orc.program([orc.gen_random(4, 5, 3), orc.measure_all(4)]), 4, fusion=0)
    nz = {"depolarizing_1q": 0.01, "depolarizing_2q": 0.05, "amplitude_damping": 0.02,
          "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}
    a, _ = orc.run(comp, 300, 42, noise=nz)
    b, _ = orc.run(comp, 300, 42, noise=nz)
    assert a == b and sum(a.values()) == 300
    ideal, _ = orc.run(comp, 300, 42)
    triv, _ = orc.run(comp, 300, 42, noise={k: 0.0 for k in nz})
    assert ideal == triv


def test_shot_parallel_per_shot_path_equals_serial(orc):
    """run_per_shot's shot-parallel fold (src/exec.rs:507-528): dealing the
    shots to workers with a state each gives every shot the word the serial
    loop gives it, and zo_run's counts at any thread count are the same."""
    nz = {"depolarizing_1q": 0.02, "depolarizing_2q": 0.05, "bit_flip": 0.01, "phase_flip": 0.01,
          "amplitude_damping": 0.03, "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.03}
    I = orc.instr
    prog = orc.program([orc.gen_random(8, 5, 12), orc.reset(2), orc.measure(4, 0),
                        I("ry", [1], [0.4], cond=(0, 1)), orc.measure_all(8)])
    c = orc.Compiled(prog, 8, cregs=[8], fusion=0)
    for prec in (64, 32):
        serial = orc.run_shots(c, 300, 31, prec, nz, threads=1)
        for t in (2, 5, 8):
            assert np.array_equal(orc.run_shots(c, 300, 31, prec, nz, threads=t), serial)
        v, k = np.unique(serial, return_counts=True)
        want = {int(a): int(b) for a, b in zip(v, k)}
        orc.set_threads(6)
        try:
            assert orc.run(c, 300, 31, prec=prec, noise=nz)[0] == want
        finally:
            orc.set_threads(1)
        # shot s of a run seeded S is the single shot of a run seeded S ^ s
        assert int(orc.run_shots(c, 1, 31 ^ 17, prec, nz)[0]) == int(serial[17])


def test_trajectory_fixtures_are_the_oracles(orc):
    """tests/golden/trajectories_16q_d30_f{32,64}.npy hold the oracle's
    per-shot words for the noisy-trajectory configuration: recompute a
    spread of shots (each shot s as the single shot of seed 99 ^ s)."""
    import os
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_fixtures as mf

    c = mf.compiled()
    for prec in (32, 64):
        words = np.load(os.path.join(os.path.dirname(mf.__file__), f"trajectories_16q_d30_f{prec}.npy"))
        assert words.dtype == np.uint64 and len(words) == mf.SHOTS
        for s in (0, 1, 777, 2048, 4095):
            assert int(orc.run_shots(c, 1, mf.SEED ^ s, prec, mf.NOISE)[0]) == int(words[s]), (prec, s)
        # the readout flips make most words distinct; a broken writer would not
        assert len(np.unique(words)) > 1000
