"""The allocation-free sort entry point (atr_sort_pairs_noalloc) against
atr_sort_pairs, through the C ABI: two calls in a row on the same GPU's
scratch array, the second smaller (stale contents of the first must not leak
into it), plus the identity-value form and the size limit."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import atrbvh_amd as ab

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20


def _keys(rng, n, key_bits):
    dt = np.uint64 if key_bits == 64 else np.uint32
    hi = (1 << 63) if key_bits == 64 else (1 << 32) - 1
    k = rng.integers(0, hi, n, dtype=np.uint64).astype(dt)
    k[rng.integers(0, n, n // 4)] = k[0]  # many equal keys: stability matters
    return k


def _both(ctx, key_bits, keys, vals):
    L = ab.lib()
    n = keys.size
    dt = keys.dtype
    a_k, b_k = ctx.to_device(keys), ctx.to_device(keys)
    a_v = ctx.to_device(vals) if vals is not None else None
    b_v = ctx.to_device(vals) if vals is not None else None
    ab.check(L.atr_sort_pairs(ctx.h, key_bits, a_k.ptr, a_v.ptr if a_v else None, n))
    ab.check(L.atr_sort_pairs_noalloc(ctx.h, key_bits, b_k.ptr, b_v.ptr if b_v else None, n))
    ctx.sync()
    assert np.array_equal(a_k.download(dt, n), b_k.download(dt, n))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(b_k.download(dt, n), keys[order])
    if vals is not None:
        got = b_v.download(np.uint32, n)
        assert np.array_equal(a_v.download(np.uint32, n), got)
        assert np.array_equal(got, vals[order])


def test_no_alloc_sort_matches_sort_pairs(ctx):
    rng = np.random.default_rng(41)
    # first call: the largest size the scratch array holds, 64-bit keys
    n1 = MAX_KEYS
    _both(ctx, 64, _keys(rng, n1, 64), rng.permutation(n1).astype(np.uint32))
    # second call, smaller, 32-bit keys: the scratch still holds the first
    # call's histograms, look-back words and keys
    n2 = 100_003
    _both(ctx, 32, _keys(rng, n2, 32), rng.permutation(n2).astype(np.uint32))
    # identity values (NULL) and a tiny sort
    _both(ctx, 64, _keys(rng, 4097, 64), None)
    _both(ctx, 32, _keys(rng, 2, 32), rng.permutation(2).astype(np.uint32))


def test_no_alloc_sort_rejects_sizes_beyond_its_scratch(ctx):
    L = ab.lib()
    d = ctx.buffer(8)
    st = L.atr_sort_pairs_noalloc(ctx.h, 64, d.ptr, None, MAX_KEYS + 1)
    assert st == 1
    assert b"at most 1048576" in L.atr_last_error()
    assert L.atr_sort_pairs_noalloc(ctx.h, 16, d.ptr, None, 10) == 1
    ctx.sync()
