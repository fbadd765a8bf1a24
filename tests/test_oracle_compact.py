"""Pins of the oracle's restatement of the reference tracer's layout and hit
test (CudaBVH::createCompact / woopifyTri, gpu-ray-traversal/src/rt/cuda/
CudaBVH.cpp:469-600; kepler_dynamic_fetch `trace`, rt/kernels/
kepler_dynamic_fetch.cu:61-320), by what the mathematics fixes:

- Woop rows are rows of the inverse of [v0-v2 | v1-v2 | n | v2]: applied to
  v0, v1, v2 they give the barycentric coordinates (1,0), (0,1), (0,0) and a
  zero plane distance, and they equal the float64 inverse to fp32 accuracy.
- The layout decodes back into the collapsed tree: the same internal nodes,
  the same child boxes bit for bit, the same triangles in every leaf in list
  order, each triangle once; node slots follow the last-in-first-out
  discovery order (a node's internal children get consecutive slots, left
  first, after every slot handed out before it was processed).
- The Woop traversal finds the same closest triangle as brute-force
  Moller-Trumbore over all triangles (ties excepted), t within 1e-5 relative
  plus 1e-6 absolute (the Woop test's t = (w - o.n) / (d.n) carries rounding
  of the scene's unit-size coordinates, which dominates for hits at tiny t)."""
from __future__ import annotations

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import C1_N, C1_SEED, camera_rays_host, compare_hits, compare_hits_two_tests


def _collapsed(seed=1, n=65536, sigma=0.005):
    v = ab.gen_soup(seed, n, sigma)
    mn, mx = oracle.scene_bbox(v)
    t = oracle.lbvh(v, mn, mx, 30)
    oracle.atrbvh(t, 9, 2)
    c, _ = oracle.collapse(t)
    return v, c


def test_woopify_inverts_the_triangle_frame():
    rng = np.random.default_rng(4)
    for _ in range(200):
        p = rng.normal(size=(3, 3)).astype(np.float32)
        w = oracle.woopify(p[0], p[1], p[2]).astype(np.float64)
        tri = p.astype(np.float64)
        a, b = tri[0] - tri[2], tri[1] - tri[2]
        M = np.eye(4)
        M[:3, 0], M[:3, 1], M[:3, 2], M[:3, 3] = a, b, np.cross(a, b), tri[2]
        inv = np.linalg.inv(M)
        want = np.stack([np.r_[inv[2, :3], -inv[2, 3]], inv[0], inv[1]])
        scale = np.abs(want).max()
        assert np.allclose(w, want, rtol=0, atol=2e-5 * scale)
        for k, (u, v) in enumerate(((1, 0), (0, 1), (0, 0))):
            q = tri[k]
            assert abs(w[1, :3] @ q + w[1, 3] - u) < 1e-4
            assert abs(w[2, :3] @ q + w[2, 3] - v) < 1e-4
            assert abs(-(w[0, :3] @ q) + w[0, 3]) < 1e-4 * max(1.0, scale)


@pytest.mark.parametrize("sigma", [0.005, 0.2])
def test_compact_layout_decodes_to_the_collapsed_tree(sigma):
    v, c = _collapsed(5, 20000, sigma)
    nodes, woop, tri = oracle.compact(c, v, 16)
    n = c.n
    bits = lambda x: np.float32(x).view(np.int32)  # noqa: E731
    seen_tri = np.zeros(n, np.int32)
    next_slot = 1
    stack = [(0, 0)]  # processing order reconstructed from the layout: LIFO over slots
    visited = 0
    while stack:
        node, slot = stack.pop()
        visited += 1
        rec = nodes[slot]
        kids = (c.left[node], c.right[node])
        for k, ch in enumerate(kids):
            box = rec[[0, 1, 2, 3]] if k == 0 else rec[[4, 5, 6, 7]]
            zz = rec[[8, 9]] if k == 0 else rec[[10, 11]]
            assert np.array_equal(box, [bits(c.bbmin[ch, 0]), bits(c.bbmax[ch, 0]),
                                        bits(c.bbmin[ch, 1]), bits(c.bbmax[ch, 1])])
            assert np.array_equal(zz, [bits(c.bbmin[ch, 2]), bits(c.bbmax[ch, 2])])
            ptr = rec[12 + k]
            if c.data[ch] < 0:
                assert ptr == 4 * next_slot, "internal children take the next slots in order"
                stack.append((ch, next_slot))
                next_slot += 1
            else:
                assert ptr < 0
                row = ~ptr
                tris = (list(c.tri_index[c.data[ch]: c.data[ch] + c.count[ch]])
                        if c.count[ch] > 1 else [c.data[ch]])
                for t in tris:
                    assert tri[row] == t and tri[row + 1] == 0 and tri[row + 2] == 0
                    seen_tri[t] += 1
                    row += 3
                assert np.all(woop[row].view(np.uint32) == 0x80000000) and tri[row] == 0
        assert rec[14] == 0 and rec[15] == 0
    assert visited == nodes.shape[0] == next_slot
    assert np.all(seen_tri == 1), "every triangle in exactly one leaf"
    # Compact (byte offsets) differs from Compact2 only in the pointer unit
    n1, w1, t1 = oracle.compact(c, v, 1)
    assert np.array_equal(w1, woop) and np.array_equal(t1, tri)
    inner = n1[:, 12:14] >= 0
    assert np.array_equal(n1[:, 12:14][inner], 16 * nodes[:, 12:14][inner])


def test_woop_traversal_equals_brute_force():
    v, c = _collapsed(1, 4096, 0.02)
    nodes, woop, tri = oracle.compact(c, v, 16)
    rays = camera_rays_host(64)
    rng = np.random.default_rng(3)
    extra = np.zeros((2048, 8), np.float32)
    extra[:, 0:3] = rng.uniform(-0.5, 1.5, (2048, 3))
    d = rng.normal(size=(2048, 3))
    extra[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    extra[:, 7] = np.inf
    rays = np.concatenate([rays, extra])
    got = oracle.trace_compact(nodes, woop, tri, rays)
    want = oracle.trace_brute(v, rays)
    assert (want[:, 0] >= 0).sum() > 500
    compare_hits(got, want, rtol=1e-5, atol=1e-6)
    anyh = oracle.trace_compact(nodes, woop, tri, rays, any_hit=True)
    assert np.array_equal(anyh[:, 0] >= 0, want[:, 0] >= 0)


def test_woop_and_moller_trumbore_agree_off_edges_c1():
    """The 65,536-triangle soup and its 256 x 256 primary rays: the Woop
    traversal and the Moller-Trumbore traversal of the same collapsed tree
    give the same hits except on rays that graze a triangle edge (one such
    ray here: a hit for the Woop test, a miss for Moller-Trumbore)."""
    v, c = _collapsed(C1_SEED, C1_N)
    nodes, woop, tri = oracle.compact(c, v, 16)
    rays = camera_rays_host(256)
    got = oracle.trace_compact(nodes, woop, tri, rays)
    want = oracle.trace(c, v, rays)
    assert (want[:, 0] >= 0).sum() > 5000
    assert compare_hits_two_tests(got, want, v, rays) <= 3
