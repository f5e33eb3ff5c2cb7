"""zb_sample_noalloc (the sampler without device allocation) against zb_sample
through the C ABI: equal indices on consecutive calls that share the fixed
scratch region (the second with a smaller state and fewer shots, so stale
contents of the first would show), on both routing paths (device sort from
32768 shots, host routing below), and a clear error past the region's size.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import zeno_amd as z
from zeno_amd import _lib

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _prepared(orc, n, prec, seed):
    be = z.HipBackend(n, prec)
    be.apply_compiled(z.compile(orc.gen_random(n, 6, seed), n, fusion=1))
    be.finish()
    return be


def _both(be, shots, seed):
    L = _lib.lib()
    want = np.zeros(max(shots, 1), np.uint64)
    got = np.full(max(shots, 1), 0xDEAD, np.uint64)
    assert L.zb_sample(be.handle, shots, seed, _lib.ptr(want)) == 0, _lib.last_error()
    rc = L.zb_sample_noalloc(be.handle, shots, seed, _lib.ptr(got))
    assert rc == 0, _lib.last_error()
    return want[:shots], got[:shots]


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_sampler_equals_zb_sample(orc, prec):
    cases = [(20, 100_000, 11), (13, 3_000, 12), (22, 40_000, 13), (9, 500, 14), (17, 0, 15), (16, 1, 16)]
    for n, shots, seed in cases:  # consecutive calls share the one scratch region
        _say(f"noalloc prec={prec} n={n} shots={shots} seed={seed}")
        with _prepared(orc, n, prec, seed) as be:
            want, got = _both(be, shots, seed)
            assert np.array_equal(want, got), f"n={n} shots={shots}"
            if shots:
                assert np.all(np.diff(got.astype(np.int64)) >= 0)  # ascending, as zb_sample
    # the oracle agrees on a small case (same state bits, same indices)
    with _prepared(orc, 12, prec, 3) as be:
        o = orc.State(12, prec)
        o.set_raw(*be.raw())
        want = o.sample(2000, 99)
        _, got = _both(be, 2000, 99)
        assert np.array_equal(want, got)


def test_noalloc_sampler_rejects_oversize_requests():
    with z.HipBackend(10, 64) as be:
        L = _lib.lib()
        out = np.zeros(1, np.uint64)
        rc = L.zb_sample_noalloc(be.handle, 1 << 23, 1, _lib.ptr(out))  # 64 MiB of draws alone
        assert rc == _lib.ZB_ERR_ARG
        assert "need" in _lib.last_error() and "scratch" in _lib.last_error()
        # the handle still works afterwards, with and without the fixed region
        got = np.zeros(100, np.uint64)
        assert L.zb_sample_noalloc(be.handle, 100, 5, _lib.ptr(got)) == 0
        assert np.array_equal(got, be.sample(100, 5))
