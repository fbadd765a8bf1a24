"""GPU parity of the reference tracer's layout and traversal: the Compact2 /
Compact layout written on the device byte for byte equal to the oracle's
restatement of CudaBVH::createCompact on the same collapsed tree, and the
`trace` ABI (atr_trace_compact) returning exactly the oracle's Woop traversal
hits (id and t bits) and, ties and rounding of the different hit test
excepted, the same hits as the Moller-Trumbore traversal atr_trace."""
from __future__ import annotations

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import (CAMERA_BASIS, CAMERA_EYE, C1_N, C1_SEED, bits, compare_hits,
                     compare_hits_two_tests)

pytestmark = pytest.mark.gpu


def _pipeline(ctx, v, bits_=30, opt="atrbvh"):
    s = ctx.scene(v)
    mn, mx = s.bounds()
    t = ab.LBVHBuilder(bits_ == 63).BuildTree(s)
    ref = oracle.lbvh(v, mn, mx, bits_)
    if opt == "atrbvh":
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
        oracle.atrbvh(ref, 9, 2)
    elif opt == "trbvh":
        ab.TRBVHOptimizer(7, 3).Optimize(t)
        oracle.trbvh(ref, 7, 3)
    c, _ = ab.BVHTreeCollapser().Collapse(t)
    rc, _ = oracle.collapse(ref)
    return s, t, c, rc


@pytest.mark.parametrize("scene", ["c1", "overlap", "dup", "gmm63"])
def test_compact_layout_byte_equal(ctx, scene):
    if scene == "c1":
        v, b, opt = ab.gen_soup(C1_SEED, C1_N), 30, "atrbvh"
    elif scene == "overlap":
        v, b, opt = ab.gen_soup(8, 30000, 0.2), 30, "trbvh"
    elif scene == "dup":
        base = ab.gen_soup(3, 40)
        v, b, opt = np.ascontiguousarray(np.concatenate([base] * 60)), 30, "atrbvh"
    else:
        v, b, opt = ab.gen_gmm(9, 300_000), 63, "atrbvh"
    s, t, c, rc = _pipeline(ctx, v, b, opt)
    for div in (16, 1):
        lay = c.compact(div)
        nodes, woop, tri = lay.download(ctx)
        rn, rw, rt = oracle.compact(rc, v, div)
        assert nodes.shape == rn.shape and woop.shape == rw.shape
        assert np.array_equal(nodes, rn), f"{scene}: node records differ (div {div})"
        assert np.array_equal(bits(woop), bits(rw)), f"{scene}: Woop rows differ"
        assert np.array_equal(tri, rt)
    c.close()
    t.close()
    s.close()


def test_trace_compact_c1(ctx):
    v = ab.gen_soup(C1_SEED, C1_N)
    s, t, c, rc = _pipeline(ctx, v)
    lay = c.compact(16)
    nodes, woop, tri = lay.download(ctx)
    g = 256
    r = g * g
    d_rays = ctx.buffer(r * 32)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    d_mt = ctx.buffer(r * 16)
    c.trace(d_rays, r, d_mt)
    d_sec = ctx.buffer(r * 32)
    ctx.rays_diffuse(s, d_rays, d_mt, r, 99, np.inf, d_sec)
    for name, d_in in (("primary", d_rays), ("diffuse", d_sec)):
        rays = d_in.download(np.float32, (r, 8))
        pad = np.full((r, 4), 0x5A5A5A5A, np.int32)
        d_res = ctx.to_device(pad)
        ctx.trace_compact(r, d_in, d_res, lay)
        got = d_res.download(np.int32, (r, 4))
        assert np.all(got[:, 2:] == 0x5A5A5A5A), "padding must stay untouched"
        want = oracle.trace_compact(nodes, woop, tri, rays)
        assert np.array_equal(got[:, 0], want[:, 0]), name
        assert np.array_equal(got[:, 1], want[:, 1]), name
        c.trace(d_in, r, d_mt)
        mt = d_mt.download(np.int32, (r, 4))
        assert (mt[:, 0] >= 0).sum() > 1000
        compare_hits_two_tests(got, mt, v, rays)
        d_any = ctx.to_device(pad)
        ctx.trace_compact(r, d_in, d_any, lay, any_hit=True)
        anyh = d_any.download(np.int32, (r, 4))
        assert np.array_equal(anyh[:, 0] >= 0, want[:, 0] >= 0)
    c.close()
    t.close()
    s.close()


@pytest.mark.parametrize("count", [1, 63, 65, 1000])
def test_trace_compact_every_ray_once_at_any_count(ctx, count):
    """The reference-ABI traversal on batches of any size: each ray traced
    once with the oracle's result, nothing written past the batch."""
    v = ab.gen_soup(C1_SEED, C1_N)
    s, t, c, rc = _pipeline(ctx, v)
    lay = c.compact(16)
    nodes, woop, tri = lay.download(ctx)
    g = 256
    d_rays = ctx.buffer(g * g * 32)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    rays = np.ascontiguousarray(d_rays.download(np.float32, (g * g, 8))[::5][:count])
    assert rays.shape == (count, 8)  # the ray buffer holds the whole batch
    d_in = ctx.to_device(rays)
    d_res = ctx.to_device(np.full((count + 64, 4), 0x5A5A5A5A, np.int32))
    ctx.trace_compact(count, d_in, d_res, lay)
    got = d_res.download(np.int32, (count + 64, 4))
    want = oracle.trace_compact(nodes, woop, tri, rays)
    assert np.array_equal(got[:count, 0], want[:, 0])
    assert np.array_equal(got[:count, 1], want[:, 1])
    assert np.all(got[count:] == 0x5A5A5A5A) and np.all(got[:count, 2:] == 0x5A5A5A5A)
    c.close()
    t.close()
    s.close()


def test_compact_full_size_gmm(ctx):
    """The 2,097,152-triangle GMM scene: layout byte-equal, every 16th of the
    2048 x 2048 primary rays equal to the oracle's Woop traversal."""
    v = ab.gen_gmm(2, 2_097_152)
    s, t, c, rc = _pipeline(ctx, v, 63)
    lay = c.compact(16)
    nodes, woop, tri = lay.download(ctx)
    rn, rw, rt = oracle.compact(rc, v, 16)
    assert np.array_equal(nodes, rn)
    assert np.array_equal(bits(woop), bits(rw)) and np.array_equal(tri, rt)
    g = 2048
    r = g * g
    d_rays = ctx.buffer(r * 32)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    d_res = ctx.buffer(r * 16)
    ctx.trace_compact(r, d_rays, d_res, lay)
    got = d_res.download(np.int32, (r, 4))
    rays = d_rays.download(np.float32, (r, 8))
    sub = slice(0, r, 16)
    want = oracle.trace_compact(nodes, woop, tri, rays[sub])
    assert np.array_equal(got[sub, 0], want[:, 0]) and np.array_equal(got[sub, 1], want[:, 1])
    ctx.sync()
    c.close()
    t.close()
    s.close()
