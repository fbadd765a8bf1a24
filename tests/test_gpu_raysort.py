"""GPU parity of the ray coherence sort (atr_rays_morton_sort: RayBuffer::
mortonSort, gpu-ray-traversal/src/rt/ray/RayBuffer.cpp:103-163, its kernels in
rt/ray/RayBufferKernels.cu:64-196) against the oracle: keys bit for bit, the
stable order, the reordered rays and slot maps exactly, and traversal results
of the sorted rays, mapped back, identical to those of the unsorted rays."""
from __future__ import annotations

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import CAMERA_BASIS, CAMERA_EYE, bits

pytestmark = pytest.mark.gpu


def _rays(r, seed, tmax=None, dup=False):
    rng = np.random.default_rng(seed)
    rays = np.zeros((r, 8), np.float32)
    rays[:, 0:3] = rng.uniform(-1.0, 2.0, (r, 3))
    d = rng.normal(size=(r, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = rng.uniform(0.0, 3.0, r) if tmax is None else tmax
    if dup:  # many equal keys: the order among them is the slot order
        rays[1::3] = rays[0]
    return rays


def _sort(ctx, rays, slot_in=None):
    r = rays.shape[0]
    d_in = ctx.to_device(rays)
    d_out = ctx.buffer(r * 32)
    d_s2i = ctx.buffer(r * 4)
    d_i2s = ctx.buffer(r * 4)
    d_sin = ctx.to_device(slot_in) if slot_in is not None else None
    ctx.rays_morton_sort(d_in, r, d_out, d_s2i, d_i2s, d_sin)
    return (d_out.download(np.float32, (r, 8)), d_s2i.download(np.int32, r),
            d_i2s.download(np.int32, r))


@pytest.mark.parametrize("r,tmax,dup", [(1, None, False), (63, None, False),
                                        (4097, np.inf, False), (100003, None, True),
                                        (1 << 20, None, False)])
def test_ray_keys_and_sort_match_the_oracle(ctx, r, tmax, dup):
    rays = _rays(r, r, tmax, dup)
    d_in = ctx.to_device(rays)
    d_keys = ctx.buffer(3 * r * 8)
    d_box = ctx.buffer(32)
    ctx.rays_morton_keys(d_in, r, d_keys, d_box)
    box = oracle.ray_box(rays)
    assert np.array_equal(bits(d_box.download(np.float32, 6)), bits(box))
    assert np.array_equal(d_keys.download(np.uint64, (3, r)), oracle.ray_keys(rays, box))
    out, s2i, i2s = _sort(ctx, rays)
    order = oracle.ray_sort_order(rays)
    assert np.array_equal(s2i, order.astype(np.int32))
    assert np.array_equal(bits(out), bits(rays[order]))
    assert np.array_equal(i2s[s2i], np.arange(r, dtype=np.int32))


def test_ray_sort_composes_slot_maps(ctx):
    r = 50000
    rays = _rays(r, 7)
    perm = np.random.default_rng(8).permutation(r).astype(np.int32)
    out, s2i, i2s = _sort(ctx, rays, perm)
    order = oracle.ray_sort_order(rays)
    assert np.array_equal(s2i, perm[order])
    assert np.array_equal(i2s[s2i], np.arange(r, dtype=np.int32))


def test_sorted_diffuse_rays_trace_to_the_same_hits(ctx):
    v = ab.gen_gmm(2, 262144)
    s = ctx.scene(v)
    t = ab.LBVHBuilder(True).BuildTree(s)
    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
    c, _ = ab.BVHTreeCollapser().Collapse(t)
    g = 512
    r = g * g
    d_rays = ctx.buffer(r * 32)
    d_hit = ctx.buffer(r * 16)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    c.trace(d_rays, r, d_hit)
    d_sec = ctx.buffer(r * 32)
    ctx.rays_diffuse(s, d_rays, d_hit, r, 77, np.inf, d_sec)
    d_plain = ctx.buffer(r * 16)
    c.trace(d_sec, r, d_plain)
    d_sorted = ctx.buffer(r * 32)
    d_s2i = ctx.buffer(r * 4)
    ctx.rays_morton_sort(d_sec, r, d_sorted, d_s2i)
    d_res_sorted = ctx.buffer(r * 16)
    c.trace(d_sorted, r, d_res_sorted)
    d_back = ctx.buffer(r * 16)
    ctx.rays_unsort_results(d_res_sorted, d_s2i, r, d_back)
    plain = d_plain.download(np.int32, (r, 4))
    back = d_back.download(np.int32, (r, 4))
    assert (plain[:, 0] >= 0).sum() > 1000
    assert np.array_equal(plain[:, :2], back[:, :2])
    sec = d_sec.download(np.float32, (r, 8))
    assert np.array_equal(d_s2i.download(np.int32, r), oracle.ray_sort_order(sec).astype(np.int32))
    c.close()
    t.close()
    s.close()
