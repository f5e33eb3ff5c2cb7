"""The exact tiled dense kernel's own source, run on the CPU (no GPU).

tests/emu builds k_kq_tiled (zeno_amd/csrc/kq_tiles.hpp) for the host, with
the library's own tile description, column swizzle and matrix packing.
Each workgroup's 512 lanes are cooperative contexts, and __syncthreads hands
control to a scheduler that resumes every lane to its next barrier:
forward, backward or shuffled.  Built with AddressSanitizer and
UndefinedBehaviorSanitizer, the result must equal the oracle's
kern_kq_scalar bit for bit.  Cases cover ops on low, high, mixed and spread
state bits, op bits out of order (a permuted state layout), and grids where
a workgroup walks several tiles, some unevenly.  Any of these faults would
show as a sanitizer report, a miscompare or an order-dependent result: an
out-of-bounds LDS or state access, a swizzle that is not a bijection, or a
missing barrier.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = ["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-ffp-contract=off", "-I" + os.path.join(ROOT, "tests", "emu"),
         "-I" + os.path.join(ROOT, "zeno_amd", "csrc"), "-I" + os.path.join(ROOT, "include")]


@pytest.fixture(scope="module")
def kq_emu(tmp_path_factory):
    d = tmp_path_factory.mktemp("kq_emu")
    src = os.path.join(ROOT, "tests", "emu", "kq_emu.cpp")
    exes = {prec: str(d / f"kq_emu_{prec}") for prec in (32, 64)}
    procs = [subprocess.Popen(FLAGS + [f"-DEMU_PREC={prec}", src, "-o", exe]) for prec, exe in exes.items()]
    assert all(p.wait() == 0 for p in procs)
    return exes, d


def _unitary(rng, k):
    q, r = np.linalg.qr(rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _phys_index(n, phys):
    """Physical index of every logical index: bit q moves to bit phys[q]."""
    x = np.arange(1 << n, dtype=np.int64)
    p = np.zeros_like(x)
    for q, b in enumerate(phys):
        p |= ((x >> q) & 1) << b
    return p


def _cases():
    rng = np.random.default_rng(21)
    out = []
    for prec in (64, 32):
        for k in (2, 3, 4, 5, 6):
            out += [(prec, 12, list(range(k)), None, 1, 0),
                    (prec, 13, list(range(13 - k, 13)), None, 2, 1),
                    (prec, 14, [0, 1, 2, 11, 12, 13][:k], None, 3, 2),
                    (prec, 14, sorted(rng.choice(14, k, replace=False).tolist()), rng.permutation(14).tolist(), 1,
                     3 + k),
                    (prec, 13, list(range(k)), rng.permutation(13).tolist(), 2, 9 + k)]
    return out


def test_tiled_kernel_source_equals_the_oracle(orc, kq_emu):
    exes, d = kq_emu
    rng = np.random.default_rng(5)
    jobs = []
    for i, (prec, n, qs, phys, grid, order) in enumerate(_cases()):
        dt = np.float64 if prec == 64 else np.float32
        v = rng.normal(size=(2, 1 << n))
        v /= np.sqrt(np.sum(v**2))
        re, im = v[0].astype(dt), v[1].astype(dt)
        m = _unitary(rng, len(qs))
        o = orc.State(n, prec)
        o.set_raw(re, im)
        o.apply_unitary(qs, m)
        want = o.raw()
        phys = phys if phys is not None else list(range(n))
        p = _phys_index(n, phys)
        pre, pim = np.empty_like(re), np.empty_like(im)
        pre[p], pim[p] = re, im
        sf, mf, of = (str(d / f"{w}_{i}.bin") for w in ("state", "mat", "out"))
        np.concatenate([pre, pim]).tofile(sf)
        np.ascontiguousarray(m, np.complex128).reshape(-1).view(np.float64).tofile(mf)
        pq = [phys[q] for q in qs]
        proc = subprocess.Popen([exes[prec], str(n), str(len(qs)), ",".join(map(str, pq)), str(grid), str(order),
                                 sf, mf, of])
        jobs.append((proc, of, p, dt, want, f"prec={prec} n={n} qs={qs} phys={phys} grid={grid} order={order}"))
        if len(jobs) == 4 or i == len(_cases()) - 1:
            for proc, of, p, dt, want, what in jobs:
                assert proc.wait() == 0, what
                got = np.fromfile(of, dt)
                half = got.size // 2
                gre, gim = got[:half][p], got[half:][p]
                bits = np.uint64 if dt == np.float64 else np.uint32
                assert np.array_equal(gre.view(bits), want[0].view(bits)), what
                assert np.array_equal(gim.view(bits), want[1].view(bits)), what
            jobs = []
