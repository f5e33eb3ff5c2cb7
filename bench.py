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


def op_class(op, n_sorted=None):
    """Kernel class of a compiled op and its algorithmic bytes factor (in
    units of 2^n * s): sweeps 4, CX 2."""
    k = len(op["qubits"])
    if op["kind"] == 0:
        return ("k_1q" if k == 1 else f"k_kq<{k}>"), 4
    if op["kind"] == 1:
        return ("k_diag1" if k == 1 else "k_diagk"), 4
    return "k_cx", 2


def run_gpu(args, comp, ops, prec, torch):
    import zeno_amd as z

    n = args.n
    s = 8 if prec == 64 else 4
    unit = (1 << n) * s
    be = z.HipBackend(n, prec, args.mode)
    ext = torch.cuda.ExternalStream(be.stream())
    classes = [op_class(o) for o in ops]
    per_class = {}
    step_times = []
    total_bytes = 2 * unit + sum(f * unit for _, f in classes)  # zero init + sweeps
    for step in range(args.warmup + args.steps):
        timed = step >= args.warmup
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(len(ops) + 1)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        be.reset_all()
        for i in range(len(ops)):
            evs[i].record(ext)
            be.apply_compiled(comp, i, i + 1)
        evs[-1].record(ext)
        be.finish()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if timed:
            step_times.append(dt)
            for i, (cls, f) in enumerate(classes):
                ms = evs[i].elapsed_time(evs[i + 1])
                c = per_class.setdefault(cls, {"launches": 0, "ms": 0.0, "bytes_per_launch": f * unit})
                c["launches"] += 1
                c["ms"] += ms
        _log(f"  f{prec} step {step} {'timed' if timed else 'warmup'}: {dt:.3f} s")
    return be, step_times, per_class, total_bytes


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


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--fusion", type=int, default=5)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--mode", type=int, default=1, help="0 exact per-op sweeps, 1 tiled")
    ap.add_argument("--cpu-ops", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(
            "bench.py measures one MI355X: the state-vector path is benchmarked on a single GPU; "
            "multi-GPU sharding by global qubits is a design note, not a run (use --gpus 1).",
            file=sys.stderr,
        )
        sys.exit(2)

    import torch

    import zeno_amd as z

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)

    n = args.n
    prog = z.matching_circuit(n, args.depth, args.seed, rzz_period=5).to_program()[0]
    t0 = time.perf_counter()
    comp = z.compile(prog, n, fusion=args.fusion)
    t_compile = time.perf_counter() - t0
    ops = comp.ops()
    _log(f"compiled {comp.stats['input_gates']} gates -> {len(ops)} ops at fusion {args.fusion} in {t_compile:.2f} s")

    be64, t64, cls64, bytes64 = run_gpu(args, comp, ops, 64, torch)
    be32, t32, cls32, bytes32 = run_gpu(args, comp, ops, 32, torch)
    err = be32.max_abs_diff(be64)
    norm64 = be64.norm_sqr()
    be32.close()
    be64.close()

    s64 = sum(t64) / len(t64)
    s32 = sum(t32) / len(t32)
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
            "input_gates": comp.stats["input_gates"],
        },
        "s_per_circuit_f64": round(s64, 4),
        "s_per_circuit_f32": round(s32, 4),
        "hbm_gbps_f32": round(bytes32 / s32 / 1e9, 2),
        "f32_speedup_vs_f64": round(s64 / s32, 3),
        "f32_max_abs_damp_vs_f64": err,
        "f64_norm_sqr": norm64,
        "kernels_f64": {k: {"launches": v["launches"], "avg_us": round(1e3 * v["ms"] / v["launches"], 2)} for k, v in cls64.items()},
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
