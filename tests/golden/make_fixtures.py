"""Writes the per-shot fixtures of the noisy-trajectory configuration with
the CPU oracle (self-generated: the reference itself cannot be built or run
here, so these are the oracle's outputs, which the oracle's own tests pin).

16 qubits, depth 30, brickwork random circuit seed 2024 followed by
measure_all, fusion 0; noise p1 = 1e-3, p2 = 1e-2, amplitude damping 5e-3,
readout flips 2%; shot seed 99; 4096 shots; one file per precision holding
each shot's clbit word (uint64).

    python tests/golden/make_fixtures.py [--threads T]
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

NOISE = {"depolarizing_1q": 1e-3, "depolarizing_2q": 1e-2, "amplitude_damping": 5e-3,
         "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.02}
SHOTS, SEED = 4096, 99


def compiled():
    prog = oracle.program([oracle.gen_random(16, 30, 2024), oracle.measure_all(16)])
    return oracle.Compiled(prog, 16, fusion=0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    oracle.build()
    c = compiled()
    for prec in (32, 64):
        words = oracle.run_shots(c, SHOTS, SEED, prec, NOISE, args.threads)
        path = os.path.join(HERE, f"trajectories_16q_d30_f{prec}.npy")
        np.save(path, words.astype(np.uint64))
        print(f"wrote {path}: {len(words)} shots, {len(np.unique(words))} distinct words", flush=True)


if __name__ == "__main__":
    main()
