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
    src = os.path.join(ROOT, "tests", "emu", "traj_emu.cpp")
    builds = {(name, prec): (str(d / f"traj_emu_{name}_{prec}.o"), extra + [f"-DEMU_PREC={prec}"])
              for name, extra in (("inlined", []), ("outlined", ["-DZENO_EMU_OUTLINE"])) for prec in (32, 64)}
    # the compilations are independent: run them side by side
    procs = [subprocess.Popen(flags + ["-c", os.path.join(ROOT, "zeno_amd", "csrc", "host.cpp"), "-o", host])]
    procs += [subprocess.Popen(flags + extra + ["-c", src, "-o", obj]) for obj, extra in builds.values()]
    assert all(p.wait() == 0 for p in procs)
    out = {}
    for (name, prec), (obj, _) in builds.items():
        exe = str(d / f"traj_emu_{name}_{prec}")
        subprocess.run(flags + [obj, host, "-o", exe], check=True)
        out.setdefault(name, {})[prec] = exe
    return out, d


def _start(exes, d, prog, n, cregs, fusion, prec, shots, seed, noise, order):
    exe = exes[prec]
    tag = f"{os.path.basename(exe)}_{n}_{order}_{len(noise)}"
    pf, of = str(d / f"prog_{tag}.bin"), str(d / f"out_{tag}.bin")
    open(pf, "wb").write(prog.tobytes())
    p = subprocess.Popen([exe, pf, str(n), ",".join(map(str, cregs or [n])), str(fusion), str(prec), str(shots),
                          str(seed), ",".join(str(noise.get(k, 0.0)) for k in NOISE_KEYS), str(order), of])
    return p, of


def _run_all(jobs, width=4):
    """Runs emulator jobs (argument tuples of _start, the first being the
    builds of one kind by precision), `width` at a time; returns each job's
    words in order."""
    out, running = [None] * len(jobs), []
    for i, job in enumerate(jobs):
        running.append((i, *_start(*job)))
        if len(running) == width or i == len(jobs) - 1:
            for j, p, of in running:
                assert p.wait() == 0, jobs[j][1:]
                out[j] = np.fromfile(of, np.uint64)
            running = []
    return out



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
            jobs = [(exe, d, prog, n, cregs, fusion, prec, shots, seed, noise, order)
                    for exe in exes.values() for order in (0, 1, 7)]
            for job, got in zip(jobs, _run_all(jobs)):
                assert np.array_equal(got, want), (n, prec, job[0], job[-1], np.nonzero(got != want)[0][:8])


def test_emulated_wide_kernel_equals_the_oracle_per_shot(orc, emulators):
    """The wide instance: a dynamic circuit compiled at fusion 5 (dense
    blocks on 4 and 5 qubits, fused diagonals) with mid-circuit measure,
    reset and a conditional, noiseless and noisy."""
    exes, d = emulators
    I = orc.instr
    rows = [orc.gen_random(10, 5, 9), orc.measure(2, 0), orc.reset(6), I("rx", [4], [0.9], cond=(0, 1)),
            orc.gen_random(10, 3, 10), orc.measure_all(10)]
    prog = orc.program(rows)
    comp = orc.Compiled(prog, 10, cregs=[10], fusion=5)
    widths = [len(o["qubits"]) for o in comp.ops() if o["kind"] in (0, 1)]
    assert max(widths) >= 4, widths  # the wide instance runs
    nz = {"depolarizing_1q": 0.02, "depolarizing_2q": 0.05, "amplitude_damping": 0.05, "readout_flip_0to1": 0.02}
    for noise in ({}, nz):
        for prec in (32, 64):
            want = orc.run_shots(comp, 12, 31, prec, noise, 4)
            jobs = [(exe, d, prog, 10, [10], 5, prec, 12, 31, noise, order)
                    for exe in exes.values() for order in (0, 7)]
            for job, got in zip(jobs, _run_all(jobs)):
                assert np.array_equal(got, want), (prec, job[0], job[-1], bool(noise))


def test_emulated_kernel_with_block_sums_in_global_memory(orc, emulators):
    """22 qubits: every probability sum has 512 blocks, more than the LDS
    holds, so the block totals go to the workgroup's slot and are reduced
    level by level there.  Damping after every gate and all 22 final
    measurements exercise it."""
    exes, d = emulators
    I = orc.instr
    n = 22
    prog = orc.program([I("h", [0]), I("h", [21]), I("cx", [0, 11]), I("ry", [20], [1.1]), orc.measure_all(n)])
    nz = {"amplitude_damping": 0.3, "bit_flip": 0.05}
    jobs = [(exes["inlined"], d, prog, n, None, 0, prec, 2, 5, nz, 0) for prec in (32, 64)]
    for job, got in zip(jobs, _run_all(jobs)):
        want = orc.run_shots(orc.Compiled(prog, n, fusion=0), 2, 5, job[6], nz, 2)
        assert np.array_equal(got, want), job[6]
