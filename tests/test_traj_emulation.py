"""The batched trajectory kernel's own source, run on the CPU (no GPU).

tests/emu builds zeno_amd/csrc/trajectories.hpp for the host: the 256 lanes
of a workgroup are cooperative contexts, __syncthreads a hand-over to a
scheduler that resumes every lane to its next barrier in a chosen order.
Built with AddressSanitizer and UndefinedBehaviorSanitizer, once with the
kernel's helpers inlined (as the library builds them) and once outlined,
every shot's clbit word must equal the oracle's per-shot path, with the
lanes resumed forward, backward and shuffled in every barrier phase: an
out-of-bounds or uninitialised access, undefined behaviour, or a read of
another lane's write that no barrier separates would show as a sanitizer
report or as an order-dependent word.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NOISE_KEYS = ["depolarizing_1q", "depolarizing_2q", "bit_flip", "phase_flip", "amplitude_damping",
              "readout_flip_0to1", "readout_flip_1to0"]


@pytest.fixture(scope="module")
def emulators(tmp_path_factory):
    d = tmp_path_factory.mktemp("traj_emu")
    flags = ["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=all", "-ffp-contract=off", "-fno-builtin-sin", "-fno-builtin-cos",
             "-I" + os.path.join(ROOT, "tests", "emu"), "-I" + os.path.join(ROOT, "zeno_amd", "csrc"),
             "-I" + os.path.join(ROOT, "include")]
    host = str(d / "host.o")
    subprocess.run(flags + ["-c", os.path.join(ROOT, "zeno_amd", "csrc", "host.cpp"), "-o", host], check=True)
    out = {}
    for name, extra in (("inlined", []), ("outlined", ["-DZENO_EMU_OUTLINE"])):
        exe = str(d / f"traj_emu_{name}")
        subprocess.run(flags + extra + [os.path.join(ROOT, "tests", "emu", "traj_emu.cpp"), host, "-o", exe],
                       check=True)
        out[name] = exe
    return out, d


def _run(exe, d, prog, n, cregs, fusion, prec, shots, seed, noise, order):
    pf, of = str(d / "prog.bin"), str(d / "out.bin")
    open(pf, "wb").write(prog.tobytes())
    subprocess.run([exe, pf, str(n), ",".join(map(str, cregs or [n])), str(fusion), str(prec), str(shots),
                    str(seed), ",".join(str(noise.get(k, 0.0)) for k in NOISE_KEYS), str(order), of], check=True)
    return np.fromfile(of, np.uint64)


def test_emulated_kernel_equals_the_oracle_per_shot(orc, emulators):
    exes, d = emulators
    I = orc.instr
    nz9 = {"depolarizing_1q": 0.01, "depolarizing_2q": 0.03, "bit_flip": 0.01, "phase_flip": 0.01,
           "amplitude_damping": 0.02, "readout_flip_0to1": 0.02, "readout_flip_1to0": 0.03}
    rows = [orc.gen_random(9, 6, 4), I("ccx", [0, 4, 8]), I("cswap", [1, 2, 7]), orc.reset(3),
            orc.measure(5, 5), I("rx", [3], [0.7], cond=(1, 1)), orc.barrier(), orc.measure_all(9)]
    nzh = {"depolarizing_1q": 0.05, "depolarizing_2q": 0.05, "bit_flip": 0.05, "phase_flip": 0.05,
           "amplitude_damping": 0.1, "readout_flip_0to1": 0.01, "readout_flip_1to0": 0.01}
    cases = [(orc.program(rows), 9, [5, 4], 3, 60, 123, nz9),  # every channel, reset, if, 3-qubit gates
             (orc.program([orc.gen_random(13, 6, 44), orc.measure_all(13)]), 13, None, 0, 16, 77, nzh)]
    for prog, n, cregs, fusion, shots, seed, noise in cases:
        for prec in (32, 64):
            want = orc.run_shots(orc.Compiled(prog, n, cregs=cregs, fusion=fusion), shots, seed, prec, noise, 4)
            for name, exe in exes.items():
                for order in (0, 1, 7):
                    got = _run(exe, d, prog, n, cregs, fusion, prec, shots, seed, noise, order)
                    assert np.array_equal(got, want), (n, prec, name, order, np.nonzero(got != want)[0][:8])
