"""Pins of the oracle's triangle-splitting LBVH (the build over primitive boxes
that each name a triangle: generateMortonCodesTSKernel, buildTreeTSKernel and
calculateNodeBoundingBoxesTSKernel, BVHRT-Core/src/LBVH.cu:99-161, 320-428,
572-669) and of the collapse's duplicate removal (TreeCollapser.cu:112-125),
by what the mathematics fixes:

- A box (lo, hi) is the box of the degenerate triangle (lo, hi, lo), whose
  box centroid is the box's, so the build over boxes equals the build over
  those triangles with each leaf's data index mapped through the box's
  triangle.  With each triangle's own box this is the plain build.
- After collapse every leaf list holds distinct triangles, every triangle is
  reachable, and a traversal of the collapsed tree finds the same closest hits
  as brute-force Moller-Trumbore over the triangles (the fragments' boxes
  cover their triangles)."""
from __future__ import annotations

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import (assert_tree_equal, camera_rays_host, compare_hits, split_fragments,
                     tree_arrays, triangle_boxes)


def _degenerate(bmin, bmax):
    m = bmin.shape[0]
    v = np.zeros((m, 3, 4), np.float32)
    v[:, 0, :3] = bmin[:, :3]
    v[:, 1, :3] = bmax[:, :3]
    v[:, 2, :3] = bmin[:, :3]
    return v.reshape(-1, 4)


@pytest.mark.parametrize("bits_", [30, 63])
def test_boxes_reduce_to_degenerate_triangles(bits_):
    v = ab.gen_soup(21, 20000, 0.02)
    mn, mx = oracle.scene_bbox(v)
    bmin, bmax, tri = split_fragments(v, every=3, levels=2)
    assert bmin.shape[0] > 30000
    t = oracle.lbvh_boxes(bmin, bmax, tri, mn, mx, bits_)
    ref = oracle.lbvh(_degenerate(bmin, bmax), mn, mx, bits_)
    m = bmin.shape[0]
    leaves = np.arange(m - 1, 2 * m - 1)
    ref.data[leaves] = tri[ref.data[leaves]]
    assert_tree_equal(tree_arrays(t), tree_arrays(ref), "boxes")
    # the triangles' own boxes give the plain build
    tb = triangle_boxes(v)
    assert_tree_equal(tree_arrays(oracle.lbvh_boxes(*tb, mn, mx, bits_)),
                      tree_arrays(oracle.lbvh(v, mn, mx, bits_)), "own boxes")


def _lists(c):
    out = []
    stack = [c.root]
    while stack:
        node = stack.pop()
        for ch in (c.left[node], c.right[node]):
            if c.data[ch] >= 0:
                if ch < c.n - 1:
                    out.append(list(c.tri_index[c.data[ch]: c.data[ch] + c.count[ch]]))
                else:
                    out.append([int(c.data[ch])])
            else:
                stack.append(int(ch))
    return out


def test_collapse_of_a_box_tree_lists_each_triangle_once_per_leaf():
    v = ab.gen_soup(22, 6000, 0.05)
    mn, mx = oracle.scene_bbox(v)
    bmin, bmax, tri = split_fragments(v, every=1, levels=2)
    t = oracle.lbvh_boxes(bmin, bmax, tri, mn, mx, 30)
    oracle.atrbvh(t, 9, 2)
    c, _ = oracle.collapse(t)
    lists = _lists(c)
    assert all(len(set(x)) == len(x) for x in lists)
    assert sorted(set(int(x) for l_ in lists for x in l_)) == list(range(6000))
    # the scene exercises the removal: some collapsed subtree holds two
    # fragments of one triangle
    dup = 0
    m = t.n

    def leaves_under(x):
        st, out = [x], []
        while st:
            y = st.pop()
            if y >= m - 1:
                out.append(int(t.data[y]))
            else:
                st += [int(t.left[y]), int(t.right[y])]
        return out
    for node in range(m - 1):
        for ch in (c.left[node], c.right[node]):
            if c.data[ch] >= 0 and ch < m - 1:
                under = leaves_under(int(ch))
                dup += len(under) - len(set(under))
                assert c.count[ch] == len(set(under))
    assert dup > 0


def test_box_tree_trace_equals_brute_force():
    v = ab.gen_soup(23, 4096, 0.03)
    mn, mx = oracle.scene_bbox(v)
    bmin, bmax, tri = split_fragments(v, every=2, levels=2)
    t = oracle.lbvh_boxes(bmin, bmax, tri, mn, mx, 63)
    oracle.atrbvh(t, 9, 2)
    c, _ = oracle.collapse(t)
    rays = camera_rays_host(96)
    got = oracle.trace(c, v, rays)
    want = oracle.trace_brute(v, rays)
    assert (want[:, 0] >= 0).sum() > 300
    compare_hits(got, want)
