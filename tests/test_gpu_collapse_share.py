"""A collapsed tree reads its tree's parent / left / right / bbmin / bbmax
arrays in place until the tree changes or is destroyed (the collapse would copy
them unchanged).  Whatever happens to the tree afterwards -- another collapse,
treelet passes, a refit, an upload, destruction, in any order -- every
collapsed tree keeps the values it was made from, and traces as before."""
from __future__ import annotations

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import CAMERA_BASIS, CAMERA_EYE, bits

pytestmark = pytest.mark.gpu

KEYS = ("parent", "left", "right", "data", "count", "bbmin", "bbmax")


def _same(a: dict, b: dict):
    for k in KEYS:
        assert np.array_equal(bits(a[k]) if a[k].dtype == np.float32 else a[k],
                              bits(b[k]) if b[k].dtype == np.float32 else b[k]), k
    assert a["used"] == b["used"]
    assert np.array_equal(a["tri_index"][: a["used"]], b["tri_index"][: b["used"]])


def _collapsed_oracle(v, mn, mx, iters):
    ref = oracle.lbvh(v, mn, mx, 30)
    if iters:
        oracle.atrbvh(ref, 9, iters)
    rc, _ = oracle.collapse(ref)
    return rc


def _trace(ctx, c, g=64):
    r = g * g
    d_rays = ctx.buffer(r * 32)
    d_res = ctx.buffer(r * 16)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    c.trace(d_rays, r, d_res)
    return d_res.download(np.int32, (r, 4))


@pytest.mark.parametrize("order", ["tree_first", "collapsed_first"])
def test_collapsed_tree_outlives_tree_changes(ctx, order):
    n = 30000
    v = ab.gen_soup(77, n)
    s = ctx.scene(v)
    mn, mx = s.bounds()
    t = ab.LBVHBuilder(False).BuildTree(s)
    c0, _ = ab.BVHTreeCollapser().Collapse(t)          # reads t's arrays in place
    a0 = c0.download()
    hits0 = _trace(ctx, c0)
    rc0 = _collapsed_oracle(v, mn, mx, 0)
    for k in ("parent", "left", "right", "data", "count"):
        assert np.array_equal(a0[k], getattr(rc0, k)), k

    c1, _ = ab.BVHTreeCollapser().Collapse(t)          # a second one gets a copy
    _same(c1.download(), a0)

    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)  # t changes: c0 keeps its values
    assert not np.array_equal(t.download()["left"], a0["left"])
    _same(c0.download(), a0)
    _same(c1.download(), a0)
    # c1's traversal layout is built now, from its own copy
    assert np.array_equal(_trace(ctx, c1), hits0)

    c2, _ = ab.BVHTreeCollapser().Collapse(t)          # reads the optimised arrays in place
    a2 = c2.download()
    rc2 = _collapsed_oracle(v, mn, mx, 2)
    for k in ("parent", "left", "right", "data", "count"):
        assert np.array_equal(a2[k], getattr(rc2, k)), k
    hits2 = _trace(ctx, c2)

    t.refit(s)                                          # same vertices: same boxes, but a write
    _same(c2.download(), a2)
    c3, _ = ab.BVHTreeCollapser().Collapse(t)
    if order == "tree_first":
        t.close()                                       # c3 takes the arrays over
        _same(c3.download(), a2)
        assert np.array_equal(_trace(ctx, c3), hits2)
        c3.close()
    else:
        c3.close()
        _same(t.download() | {"count": a2["count"], "used": a2["used"],
                              "tri_index": a2["tri_index"], "data": a2["data"]}, a2)
        t.close()
    _same(c2.download(), a2)
    _same(c0.download(), a0)
    for c in (c0, c1, c2):
        c.close()
    s.close()


def test_stage_collapse_shares_and_upload_unshares(ctx):
    n = 5000
    v = ab.gen_soup(78, n)
    s = ctx.scene(v)
    t = ab.LBVHBuilder(True).BuildTree(s)
    c = ab.DeviceCalculateCosts(t)                      # the stage calls share as well
    ab.DeviceCollapseTree(t, c, want_sah=False)
    a = c.download()
    s2 = ctx.scene(ab.gen_soup(79, n))                  # another tree of the same size
    other = ab.LBVHBuilder(True).BuildTree(s2)
    od = other.download()
    other.close()
    assert not np.array_equal(od["left"], a["left"])
    t.upload(od)                                        # t takes the other topology
    assert np.array_equal(t.download()["left"], od["left"])
    _same(c.download(), a)
    t.close()
    _same(c.download(), a)
    c.close()
    s2.close()
    s.close()
