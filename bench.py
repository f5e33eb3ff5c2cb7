"""Benchmark: random 30-qubit depth-40 circuit, GPU f64 vs f32-bulk, on one MI355X.

Workload (one "step"): zero-initialise the state and apply every compiled op
of a seeded random-matching circuit — U3 on every qubit, CX on a random
perfect matching, an RZZ ladder on all neighbour pairs every 5th layer —
first in f64, then in f32 (f32 sweeps, f64 reductions).  The state lives in
HBM for the whole run; nothing crosses PCIe inside the timed region.

`value` is the f64 circuit's algorithmic HBM throughput (GB/s): every sweep
counts 4*2^n*s bytes (read + write of re and im), a CX 2*2^n*s, the zero
init 2*2^n*s, divided by the measured f64 circuit time.  Per-op kernel
times come from HIP events recorded on the backend's own stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "HBM GB/s + s/circuit (random 30q d40, f64 vs f32-bulk); p99 latency mixed serve"
HBM_PEAK_GBS = 8000.0


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def effective_bytes(ops, n, s):
    """Bytes if every compiled op were its own sweep (4*2^n*s, CX 2*2^n*s),
    plus the zero init: the op stream's size, not the passes executed."""
    unit = (1 << n) * s
    return 2 * unit + sum((2 if o["kind"] == 2 else 4) * unit for o in ops)


TUNE = {}


def run_gpu(n, comp, prec, mode, steps, warmup, knobs=None):
    """Time `steps` full circuits (zero init + every op + finish) after
    `warmup` untimed ones; kernel durations come from HIP events the library
    records around each launch on its own stream."""
    import zeno_amd as z

    be = z.HipBackend(n, prec, mode)
    for k, v in dict(TUNE, **(knobs or {})).items():
        be.tune(k, v)
    step_times, recs = [], []
    for step in range(warmup + steps):
        timed = step >= warmup
        be.finish()
        if timed:
            be.profile_begin()
        t0 = time.perf_counter()
        be.reset_all()
        be.apply_compiled(comp)
        be.finish()
        dt = time.perf_counter() - t0
        if timed:
            step_times.append(dt)
            recs += be.profile_end()
        _log(f"  n={n} f{prec} mode={mode} step {step} {'timed' if timed else 'warmup'}: {dt:.4f} s")
    per_class = {}
    for name, ms, nbytes in recs:
        c = per_class.setdefault(name, {"launches": 0, "ms": 0.0, "bytes_per_launch": nbytes})
        c["launches"] += 1
        c["ms"] += ms
    circuit_bytes = sum(b for _, _, b in recs) / max(1, steps)
    return be, sum(step_times) / len(step_times), per_class, circuit_bytes


def cpu_baseline(args, prog):
    import oracle

    n = args.n
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    oracle.set_threads(cores)
    comp = oracle.Compiled(prog, n, fusion=1)  # the CPU path's own default
    m = min(args.cpu_ops, len(comp))
    t0 = time.perf_counter()
    st = oracle.State(n, 64)
    t_zero = time.perf_counter() - t0
    t1 = time.perf_counter()
    st.apply_compiled(comp, 0, m)
    t_ops = time.perf_counter() - t1
    ops = comp.ops()[:m]
    unit = (1 << n) * 8
    nbytes = sum((2 if o["kind"] == 2 else 4) * unit for o in ops)
    all_ops = comp.ops()
    per_op = t_ops / m
    est = t_zero + per_op * len(all_ops)
    del st
    return {
        "value": round(nbytes / t_ops / 1e9, 3),
        "unit": "GB/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {m} of {len(all_ops)} ops of the same {n}q circuit at the CPU path's default fusion 1, "
        f"f64, oracle/zeno_oracle.c on {cores} threads; est. {est:.1f} s/circuit "
        f"(zero {t_zero:.2f} s + {per_op:.3f} s/op)",
        "s_per_circuit_est": round(est, 3),
    }


NOISE_C4 = {"depolarizing_1q": 1e-3, "depolarizing_2q": 1e-2, "amplitude_damping": 5e-3,
            "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}


def bench_trajectories(args):
    """Batched noisy trajectories: brickwork random circuit 16q depth 30,
    4096 seeded trajectories, f32 states, one persistent launch; the CPU
    trajectory sampler timed on a bounded sample of the same shots."""
    import zeno_amd as z

    n, depth, shots = 16, 30, args.traj_shots
    prog = z.random_circuit(n, depth, 2024).to_program()[0]
    prog = z.Circuit(n).extend(prog).measure_all().to_program()[0]
    c = z.compile(prog, n, fusion=0)  # noisy runs compile with fusion 0 (src/lib.rs:228-231)
    nz = z.NoiseModel(**NOISE_C4)
    z.run_trajectories(c, 64, 1, 32, nz)
    _, secs = z.run_trajectories(c, shots, args.seed, 32, nz)
    out = {"workload": f"random_circuit {n}q depth {depth} (brickwork), fusion 0, {shots} trajectories, f32, "
                       f"noise {NOISE_C4}",
           "ops": len(c), "s": round(secs, 4), "trajectories_per_s": round(shots / secs, 1)}
    if not args.no_cpu:
        import oracle

        m = args.traj_cpu_shots
        oc = oracle.Compiled(prog, n, fusion=0)
        t0 = time.perf_counter()
        oracle.run(oc, m, args.seed, prec=32, noise=NOISE_C4)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(m / dt, 2), "unit": "trajectories/s", "cores": 1, "kind": "port",
                               "sample": f"first {m} of the {shots} trajectories, oracle/zeno_oracle.c per-shot "
                                         f"sampler, f32, 1 thread"}
    return out


def bench_serve(args):
    """Mixed-size serving: a Poisson stream of seeded random circuits (n ~
    U{10..28}, depth ~ U{8..40}, 1000 shots, f32 or f64 each) through the
    continuous-batching server, against the same stream served one at a
    time (one slot)."""
    import threading

    import numpy as np

    import zeno_amd as z

    rng = np.random.default_rng(args.seed)
    reqs = []
    for i in range(args.serve_requests):
        n = int(rng.integers(10, 29))
        depth = int(rng.integers(8, 41))
        prec = 32 if rng.integers(0, 2) else 64
        prog = z.Circuit(n).extend(z.random_circuit(n, depth, 1000 + i).to_program()[0]).measure_all().to_program()[0]
        reqs.append((z.compile(prog, n, fusion=1), prec, 1000, 5000 + i))
    results = {}
    for rate in [float(x) for x in args.serve_rates.split(",") if x]:
        gaps = rng.exponential(1.0 / rate, len(reqs))
        arrivals = np.cumsum(gaps) - gaps[0]
        row = {}
        for label, slots in (("served", args.serve_slots), ("one_at_a_time", 1)):
            with z.Server(args.serve_pool << 30, slots) as srv:
                ids = [None] * len(reqs)

                def feed():
                    t0 = time.perf_counter()
                    for k, (c, prec, shots, seed) in enumerate(reqs):
                        dt = arrivals[k] - (time.perf_counter() - t0)
                        if dt > 0:
                            time.sleep(dt)
                        ids[k] = srv.submit(c, prec, shots, seed)

                th = threading.Thread(target=feed)
                th.start()
                th.join()
                done = [srv.wait(i) for i in ids]
            lat = np.array([d.latency for d in done]) * 1e3
            span = max(d.t_end for d in done) - min(d.t_submit for d in done)
            row[label] = {"slots": slots, "circuits_per_s": round(len(done) / span, 2),
                          "p50_ms": round(float(np.percentile(lat, 50)), 2),
                          "p99_ms": round(float(np.percentile(lat, 99)), 2),
                          "mean_queue_ms": round(float(np.mean([d.queued for d in done])) * 1e3, 2)}
            _log(f"serve rate={rate} {label}: {row[label]}")
        results[str(rate)] = row
    return {"workload": f"{len(reqs)} requests, Poisson arrivals, n ~ U{{10..28}}, depth ~ U{{8..40}}, "
                        f"1000 shots, f32/f64 at random, fusion 1 (compiled before arrival), pool "
                        f"{args.serve_pool} GiB",
            "rates": results}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--fusion", type=int, default=1)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--mode", type=int, default=1, help="0 exact per-op sweeps, 1 tiled")
    ap.add_argument("--cpu-ops", type=int, default=8)
    ap.add_argument("--also", default="", help="extra fusion:mode pairs to time once, e.g. 5:0,5:1")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain device copy of 2^n f64 (re) + 2^n f64 (im), the known-size kernel "
                    "that scales rocprofv3 FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--tune", default="", help="tiled-kernel knobs, e.g. prefetch=1,win_grid=1024")
    ap.add_argument("--traj-shots", type=int, default=4096)
    ap.add_argument("--traj-cpu-shots", type=int, default=24)
    ap.add_argument("--serve-requests", type=int, default=80)
    ap.add_argument("--serve-rates", default="50", help="Poisson arrival rates (req/s), comma separated")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-pool", type=int, default=48, help="serving pool in GiB")
    ap.add_argument("--no-extra", action="store_true", help="skip the trajectory and serving legs")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the opt-in fma / mfma variants (within tolerance, not bit-identical)")
    args = ap.parse_args()
    for kv in [x for x in args.tune.split(",") if x]:
        k, v = kv.split("=")
        TUNE[k] = int(v)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(
            "bench.py measures one MI355X: the state-vector path is benchmarked on a single GPU; "
            "multi-GPU sharding by global qubits is a design note, not a run (use --gpus 1).",
            file=sys.stderr,
        )
        sys.exit(2)

    import zeno_amd as z

    if z.device_count() == 0:
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)

    n = args.n
    prog = z.matching_circuit(n, args.depth, args.seed, rzz_period=5).to_program()[0]
    t0 = time.perf_counter()
    comp = z.compile(prog, n, fusion=args.fusion)
    t_compile = time.perf_counter() - t0
    ops = comp.ops()
    _log(f"compiled {comp.stats['input_gates']} gates -> {len(ops)} ops at fusion {args.fusion} in {t_compile:.2f} s")

    if args.calibrate_copy:
        ms = z.calibration_copy(2 << n, reps=2)
        nb = (2 << n) * 8
        _log(f"calibration copy (k_copy): {2 * nb} bytes moved per copy ({nb} read, same written), {ms:.3f} ms")

    be64, s64, cls64, bytes64 = run_gpu(n, comp, 64, args.mode, args.steps, args.warmup)
    be32, s32, cls32, bytes32 = run_gpu(n, comp, 32, args.mode, args.steps, args.warmup)
    err = be32.max_abs_diff(be64)
    norm64 = be64.norm_sqr()
    be32.close()
    # sample_indices on the finished f64 state: 1000 shots, once warm
    be64.sample(1000, args.seed)
    t0 = time.perf_counter()
    be64.sample(1000, args.seed + 1)
    sample_ms = (time.perf_counter() - t0) * 1e3

    # Opt-in knobs that trade bit-identity with the CPU path for speed,
    # each timed once and checked against the exact f64 state above.
    variants = {}
    if not args.no_variants:
        comp5 = None
        for label, prec, fusion, knobs in (("f64_fma", 64, args.fusion, {"fma": 1}),
                                           ("f32_fma", 32, args.fusion, {"fma": 1}),
                                           ("f32_fusion5_mfma", 32, 5, {"mfma": 1}),
                                           ("f64_fusion5_mfma", 64, 5, {"mfma": 1}),
                                           ("f32_fusion5_valu", 32, 5, {})):
            if fusion == args.fusion:
                cv = comp
            else:
                comp5 = comp5 or z.compile(prog, n, fusion=fusion)
                cv = comp5
            _log(f"variant {label} ...")
            bv, tv, kv, byv = run_gpu(n, cv, prec, args.mode, 1, 1, knobs)
            variants[label] = {"precision": f"f{prec}", "fusion": fusion, "knobs": knobs, "ops": len(cv),
                               "s_per_circuit": round(tv, 4), "speedup_vs_f64_exact": round(s64 / tv, 3),
                               "max_abs_damp_vs_f64_exact": bv.max_abs_diff(be64),
                               "kernels": {k: {"launches": v["launches"], "avg_us": round(1e3 * v["ms"] / v["launches"], 2)}
                                           for k, v in kv.items()}}
            bv.close()
    be64.close()

    alternatives = []
    for spec in [x for x in args.also.split(",") if x]:
        fk, md = (int(v) for v in spec.split(":"))
        c2 = z.compile(prog, n, fusion=fk)
        for prec in (64, 32):
            b2, t2, k2, by2 = run_gpu(n, c2, prec, md, 1, 1)
            b2.close()
            alternatives.append({"fusion": fk, "mode": "tiled" if md == 1 else "exact", "precision": f"f{prec}",
                                 "ops": len(c2), "s_per_circuit": round(t2, 4), "hbm_gbps": round(by2 / t2 / 1e9, 2),
                                 "kernels": {k: {"launches": v["launches"], "avg_us": round(1e3 * v["ms"] / v["launches"], 2)}
                                             for k, v in k2.items()}})

    dom_cls, dom = max(cls64.items(), key=lambda kv: kv[1]["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    achieved = dom["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get(f"{dom_cls}@{n}q_f64")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu:
        _log("cpu baseline ...")
        cpu = cpu_baseline(args, prog)

    out = {
        "metric": METRIC,
        "value": round(bytes64 / s64 / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round((s64 + s32) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"random-matching circuit {n}q depth {args.depth} (U3 layer + CX on a seeded perfect "
            f"matching + RZZ ladder every 5th layer), seed {args.seed}, fusion K={args.fusion}, "
            f"{'tiled' if args.mode == 1 else 'exact per-op'} HIP sweeps",
            "n_qubits": n,
            "depth": args.depth,
            "fusion": args.fusion,
            "compiled_ops": len(ops),
            "tune": dict(TUNE),
            "input_gates": comp.stats["input_gates"],
        },
        "effective_gbps_f64": round(effective_bytes(ops, n, 8) / s64 / 1e9, 2),
        "s_per_circuit_f64": round(s64, 4),
        "s_per_circuit_f32": round(s32, 4),
        "hbm_gbps_f32": round(bytes32 / s32 / 1e9, 2),
        "f32_speedup_vs_f64": round(s64 / s32, 3),
        "f32_max_abs_damp_vs_f64": err,
        "f64_norm_sqr": norm64,
        "sample_1000_shots_ms_f64": round(sample_ms, 3),
        "kernels_f64": {k: {"launches": v["launches"], "avg_us": round(1e3 * v["ms"] / v["launches"], 2)} for k, v in cls64.items()},
        "kernels_f32": {k: {"launches": v["launches"], "avg_us": round(1e3 * v["ms"] / v["launches"], 2)} for k, v in cls32.items()},
        "alternatives": alternatives,
        "variants": variants,
        "roofline": {
            "bound": "hbm",
            "kernel": dom_cls,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "avg_us": round(avg_ms * 1e3, 2),
        },
        "cpu_baseline": cpu,
    }
    if not args.no_extra:
        _log("batched trajectories ...")
        out["trajectories"] = bench_trajectories(args)
        if args.serve_requests > 0:
            _log("serving ...")
            out["serve"] = bench_serve(args)
            first = next(iter(out["serve"]["rates"].values()))
            out["p99_latency_ms_mixed_serve"] = first["served"]["p99_ms"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
