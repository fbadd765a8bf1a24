"""GPU parity of the triangle-splitting LBVH (atr_lbvh_build_boxes: the build
over primitive boxes that each name a triangle, BVHRT-Core/src/LBVH.cu:99-161,
320-428, 572-669) against the oracle, bit for bit through ATRBVH and the
collapse with duplicate removal, and its traversal against the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import (CAMERA_BASIS, CAMERA_EYE, assert_tree_equal, bits, compare_hits,
                     split_fragments, tree_arrays, triangle_boxes)

pytestmark = pytest.mark.gpu


def _box_build(ctx, s, bmin, bmax, tri, bits_):
    d_mn = ctx.to_device(bmin)
    d_mx = ctx.to_device(bmax)
    d_tri = ctx.to_device(tri)
    t = ab.LBVHBuilder(bits_ == 63).BuildTreeFromBoxes(s, d_mn, d_mx, d_tri, bmin.shape[0])
    ctx.sync()
    return t, (d_mn, d_mx, d_tri)


@pytest.mark.parametrize("bits_", [30, 63])
def test_own_boxes_give_the_plain_build(ctx, bits_):
    v = ab.gen_soup(31, 50000, 0.01)
    s = ctx.scene(v)
    t, keep = _box_build(ctx, s, *triangle_boxes(v), bits_)
    plain = ab.LBVHBuilder(bits_ == 63).BuildTree(s)
    assert_tree_equal(t.download(), plain.download(), "own boxes")
    t.close()
    plain.close()
    s.close()


@pytest.mark.parametrize("bits_,every,levels", [(30, 3, 1), (63, 2, 2), (30, 1, 3)])
def test_split_build_atrbvh_collapse_bit_exact(ctx, bits_, every, levels):
    v = ab.gen_soup(32 + every, 12000, 0.03)
    s = ctx.scene(v)
    mn, mx = s.bounds()
    bmin, bmax, tri = split_fragments(v, every=every, levels=levels)
    m = bmin.shape[0]
    t, keep = _box_build(ctx, s, bmin, bmax, tri, bits_)
    ref = oracle.lbvh_boxes(bmin, bmax, tri, mn, mx, bits_)
    assert_tree_equal(t.download(), tree_arrays(ref), "lbvh")
    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
    oracle.atrbvh(ref, 9, 2)
    assert_tree_equal(t.download(), tree_arrays(ref), "atrbvh")
    c, sah = ab.BVHTreeCollapser().Collapse(t)
    rc, rsah = oracle.collapse(ref)
    assert bits(np.float32(sah)) == bits(np.float32(rsah))
    dev = c.download()
    for k in ("parent", "left", "right", "data", "count"):
        assert np.array_equal(dev[k], getattr(rc, k)), k
    assert dev["used"] == rc.used
    assert np.array_equal(dev["tri_index"][: rc.used], rc.tri_index[: rc.used])
    g = 128
    r = g * g
    d_rays = ctx.buffer(r * 32)
    d_res = ctx.buffer(r * 16)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    c.trace(d_rays, r, d_res)
    rays = d_rays.download(np.float32, (r, 8))
    got = d_res.download(np.int32, (r, 4))
    want = oracle.trace(rc, v, rays)
    assert (want[:, 0] >= 0).sum() > 300
    compare_hits(got, want)
    sub = slice(0, r, 7)
    compare_hits(want[sub], oracle.trace_brute(v, rays[sub]))
    # the reference tracer's layout of the same collapsed tree
    lay = c.compact(16)
    nodes, woop, tri_rows = lay.download(ctx)
    rn, rw, rt = oracle.compact(rc, v, 16)
    assert np.array_equal(nodes, rn) and np.array_equal(tri_rows, rt)
    assert np.array_equal(bits(woop), bits(rw))
    assert m > 12000
    c.close()
    t.close()
    s.close()


def test_box_tree_refuses_a_vertex_refit(ctx):
    v = ab.gen_soup(35, 3000, 0.03)
    s = ctx.scene(v)
    t, keep = _box_build(ctx, s, *split_fragments(v), 30)
    with pytest.raises(ab.AtrError):
        t.refit()
    t.close()
    s.close()
