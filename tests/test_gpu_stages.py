"""GPU parity of the stage calls (the reference's low-level Device* functions,
LBVH.h:80-155, TRBVH.h:42-46, AgglomerativeTreelet.h:39-75,
TreeCollapser.h:19-62): run one at a time -- encode, sort, radix tree + boxes,
one treelet pass per threshold, costs, lists -- they give the composite calls'
trees and collapsed trees bit for bit, and the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import assert_tree_equal, bits, split_fragments, tree_arrays

pytestmark = pytest.mark.gpu


def _sorted_keys(ctx, dv, n, mn, mx, bits_):
    kb = 8 if bits_ == 63 else 4
    dk = ctx.buffer(n * kb)
    di = ctx.buffer(n * 4)
    ctx.morton_encode(dv, n, mn, mx, bits_, dk, di)
    ctx.sort_pairs(kb * 8, dk, di, n)
    return dk, di


def _collapsed_equal(a: dict, b: dict):
    for k in ("parent", "left", "right", "data", "count"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("bbmin", "bbmax"):
        assert np.array_equal(bits(a[k]), bits(b[k])), k
    assert a["used"] == b["used"]
    assert np.array_equal(a["tri_index"][: a["used"]], b["tri_index"][: b["used"]])


@pytest.mark.parametrize("kind,bits_,n", [("atrbvh", 30, 65536), ("atrbvh", 63, 30000),
                                          ("trbvh", 30, 20000), ("trbvh", 63, 9)])
def test_stages_match_composite_and_oracle(ctx, kind, bits_, n):
    v = ab.gen_soup(41 + n, n)
    s = ctx.scene(v)
    mn, mx = s.bounds()
    dv = ctx.to_device(v)
    dk, di = _sorted_keys(ctx, dv, n, mn, mx, bits_)
    staged = ab.DeviceBuildTree(s, dk, di, n, bits_)
    whole = ab.LBVHBuilder(bits_ == 63).BuildTree(s)
    ref = oracle.lbvh(v, mn, mx, bits_)
    assert_tree_equal(staged.download(), whole.download(), "radix tree stage")
    assert_tree_equal(staged.download(), tree_arrays(ref), "radix tree stage vs oracle")

    treelet, iters = (9, 2) if kind == "atrbvh" else (7, 3)
    gamma = treelet
    for _ in range(iters):
        if kind == "atrbvh":
            ab.DeviceAgglomerativeTreeletOptimizer(staged, treelet, gamma)
        else:
            ab.DeviceTreeletReestructureOptimizer(staged, treelet, gamma)
        gamma *= 2
    if kind == "atrbvh":
        ab.AgglomerativeTreeletOptimizer(treelet, iters).Optimize(whole)
        oracle.atrbvh(ref, treelet, iters)
    else:
        ab.TRBVHOptimizer(treelet, iters).Optimize(whole)
        oracle.trbvh(ref, treelet, iters)
    assert_tree_equal(staged.download(), whole.download(), "treelet passes")
    assert_tree_equal(staged.download(), tree_arrays(ref), "treelet passes vs oracle")

    c = ab.DeviceCalculateCosts(staged)
    with pytest.raises(ab.AtrError):  # unfinished until its lists are emitted
        c.stats()
    with pytest.raises(ab.AtrError):  # the lists belong to the tree the costs came from
        ab.DeviceCollapseTree(whole, c)
    sah = ab.DeviceCollapseTree(staged, c)
    with pytest.raises(ab.AtrError):  # and are emitted once
        ab.DeviceCollapseTree(staged, c)
    cw, sah_w = ab.BVHTreeCollapser().Collapse(whole)
    rc, rsah = oracle.collapse(ref)
    assert bits(np.float32(sah)) == bits(np.float32(sah_w)) == bits(np.float32(rsah))
    got = c.download()
    _collapsed_equal(got, cw.download())
    assert got["used"] == rc.used
    for k in ("parent", "left", "right", "data", "count"):
        assert np.array_equal(got[k], getattr(rc, k)), k
    c.close()
    cw.close()
    staged.close()
    whole.close()
    s.close()


@pytest.mark.parametrize("bits_", [30, 63])
def test_box_stage_matches_box_build(ctx, bits_):
    """The triangle-splitting radix-tree stage over keys that the caller made:
    a box (lo, hi) has the key of the degenerate triangle (lo, hi, lo), so
    atr_morton_encode over such triangles gives the box keys."""
    v = ab.gen_soup(53, 9000, 0.03)
    s = ctx.scene(v)
    mn, mx = s.bounds()
    bmin, bmax, tri = split_fragments(v, every=2, levels=2)
    m = bmin.shape[0]
    deg = np.stack([bmin, bmax, bmin], axis=1).reshape(3 * m, 4).astype(np.float32)
    dk, di = _sorted_keys(ctx, ctx.to_device(deg), m, mn, mx, bits_)
    d_mn, d_mx, d_tri = ctx.to_device(bmin), ctx.to_device(bmax), ctx.to_device(tri)
    staged = ab.DeviceBuildTreeTS(s, dk, di, d_mn, d_mx, d_tri, m, bits_)
    whole = ab.LBVHBuilder(bits_ == 63).BuildTreeFromBoxes(s, d_mn, d_mx, d_tri, m)
    assert_tree_equal(staged.download(), whole.download(), "box radix tree stage")
    ref = oracle.lbvh_boxes(bmin, bmax, tri, mn, mx, bits_)
    assert_tree_equal(staged.download(), tree_arrays(ref), "box radix tree stage vs oracle")
    # collapse drops repeated triangles from its lists as the composite does
    c = ab.DeviceCalculateCosts(staged)
    ab.DeviceCollapseTree(staged, c, want_sah=False)
    cw, _ = ab.BVHTreeCollapser().Collapse(whole, want_sah=False)
    _collapsed_equal(c.download(), cw.download())
    c.close()
    cw.close()
    staged.close()
    whole.close()
    s.close()


def test_stage_argument_checks(ctx):
    v = ab.gen_soup(5, 100)
    s = ctx.scene(v)
    mn, mx = s.bounds()
    dk, di = _sorted_keys(ctx, ctx.to_device(v), 100, mn, mx, 30)
    with pytest.raises(ab.AtrError):  # count differs from the scene's
        ab.DeviceBuildTree(s, dk, di, 99, 30)
    with pytest.raises(ab.AtrError):
        ab.DeviceBuildTree(s, dk, di, 100, 31)
    t = ab.DeviceBuildTree(s, dk, di, 100, 30)
    with pytest.raises(ab.AtrError):  # gamma below the treelet size
        ab.DeviceAgglomerativeTreeletOptimizer(t, 9, 8)
    with pytest.raises(ab.AtrError):
        ab.DeviceTreeletReestructureOptimizer(t, 5, 10)
    c, _ = ab.BVHTreeCollapser().Collapse(t, want_sah=False)
    with pytest.raises(ab.AtrError):  # a finished collapsed tree takes no lists
        ab.DeviceCollapseTree(t, c)
    c.close()
    t.close()
    s.close()
