"""The jump-ahead maps behind the sampler's device draws (CPU only).

From 32768 shots the draws are generated on the device, each thread starting
a block of the xoshiro256++ stream at a state reached by GF(2) jump maps
(zeno_amd/csrc/draws.hpp).  tests/emu/draws_check.cpp compares the jumped
states and each block's first output with plain sequential stepping; the GPU
tests then compare the resulting sample indices with the oracle's at 1M shots.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_jump_maps_equal_sequential_stepping(tmp_path):
    exe = str(tmp_path / "draws_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "zeno_amd", "csrc"),
                    os.path.join(ROOT, "tests", "emu", "draws_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
