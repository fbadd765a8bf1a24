"""GPU checks of the workload configurations at their stated sizes:

- 2,097,152-triangle GMM scene: TRBVH (7, 3) bit-exact against the oracle,
  2048 x 2048 primary rays and their diffuse secondaries traced and checked
  against the oracle on a strided 1/16 of the rays, and the diffuse ray
  generator against the oracle's restatement of rayGenAOKernel;
- 4,096-mesh batch (64..65,536 triangles each, about 39M triangles, 42-bit
  forest keys, 6 radix passes), tuned launch shape, graph replay: 64 meshes
  spanning the size range bit-exact against the oracle, every mesh's SAH
  equal to the plain per-mesh path;
- 64-mesh batched differentiable form: the batched hit adjoint against the
  FP64 adjoint on four meshes, then step -> refit -> device SAH -> selective
  rebuild;
- the sort's bounded wait forced to give up: every later kernel stays in
  bounds and the error is reported once.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import (CAMERA_BASIS, CAMERA_EYE, assert_tree_equal, bits, check_tree_invariants,
                     compare_hits, tree_arrays)

pytestmark = pytest.mark.gpu

C2_N = 2_097_152


@pytest.fixture(scope="module")
def c2_scene():
    return ab.gen_gmm(2, C2_N)


def test_c2_trbvh_full_size_bit_exact(ctx, c2_scene):
    v = c2_scene
    s = ctx.scene(v)
    mn, mx = s.bounds()
    t = ab.LBVHBuilder(True).BuildTree(s)
    ab.TRBVHOptimizer(7, 3).Optimize(t)
    dev = t.download()
    ref = oracle.lbvh(v, mn, mx, 63)
    oracle.trbvh(ref, 7, 3)
    assert_tree_equal(dev, tree_arrays(ref), "c2.trbvh")
    c, sah = ab.BVHTreeCollapser().Collapse(t)
    rc, rsah = oracle.collapse(ref)
    assert bits(np.float32(sah)) == bits(np.float32(rsah))
    cd = c.download()
    for k in ("data", "count"):
        assert np.array_equal(cd[k], getattr(rc, k)), k
    c.close()
    t.close()
    s.close()


def test_c2_traces_full_size(ctx, c2_scene):
    """4,194,304 primary rays and one diffuse ray per primary hit on the
    ATRBVH (9, 2) tree of the 2M scene: every 16th ray against the oracle's
    traversal of the oracle's tree; the device diffuse rays against the
    oracle's rayGenAOKernel restatement (origins and tmax exact, directions
    within 2e-6: the two sides' cosf/sinf may differ in the last bits)."""
    v = c2_scene
    s = ctx.scene(v)
    mn, mx = s.bounds()
    t = ab.LBVHBuilder(True).BuildTree(s)
    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
    c, _ = ab.BVHTreeCollapser().Collapse(t)
    g = 2048
    r = g * g
    d_rays = ctx.buffer(r * 32)
    d_res = ctx.buffer(r * 16)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    c.trace(d_rays, r, d_res)
    d_sec = ctx.buffer(r * 32)
    ctx.rays_diffuse(s, d_rays, d_res, r, 7, np.inf, d_sec)
    d_res2 = ctx.buffer(r * 16)
    c.trace(d_sec, r, d_res2)
    ctx.sync()
    rays = d_rays.download(np.float32, (r, 8))
    hits = d_res.download(np.int32, (r, 4))
    sec = d_sec.download(np.float32, (r, 8))
    hits2 = d_res2.download(np.int32, (r, 4))
    c.close()
    t.close()
    s.close()
    ref = oracle.lbvh(v, mn, mx, 63)
    oracle.atrbvh(ref, 9, 2)
    rc, _ = oracle.collapse(ref)
    sub = slice(0, r, 16)
    assert (hits[:, 0] >= 0).sum() > 100_000
    compare_hits(hits[sub], oracle.trace(rc, v, rays[sub]))
    want = oracle.rays_diffuse(v, rays, hits, 7)
    assert np.array_equal(bits(sec[:, 0:4]), bits(want[:, 0:4])), "diffuse origins"
    assert np.array_equal(bits(sec[:, 7]), bits(want[:, 7])), "diffuse tmax"
    assert np.max(np.abs(sec[:, 4:7] - want[:, 4:7])) <= 2e-6, "diffuse directions"
    live = np.nonzero(sec[:, 7] > 0)[0][::16]
    compare_hits(hits2[live], oracle.trace(rc, v, sec[live]))


# ------------------------------------------------------------------ batched (4,096 meshes)
def test_c3_full_batch_against_oracle_and_plain_path(ctx):
    from atrbvh_amd.autotune import tune_launch

    off = ab.gen_batch_sizes(3, 4096, 64, 65536)
    v = ab.gen_soup_batch(3000, off)
    sizes = np.diff(off).astype(np.int64)
    assert sizes.min() >= 64 and sizes.max() <= 65536 and 35_000_000 < off[-1] < 45_000_000
    b = ab.Batch(ctx, v, off)
    # the benchmarked shape: 42-bit forest keys -> 6 radix passes
    ctx.profile(True)
    ctx.profile_reset()
    b.build(9, 2, graph=False)
    ctx.sync()
    recs = ctx.profile_records()
    ctx.profile(False)
    assert recs["onesweep64"]["launches"] == 6, recs.get("onesweep64")
    eager = b.sah()
    tuned = tune_launch(ctx, b, 9, 2, reps=1)
    assert b.config() == {k: tuned["best"][k] for k in ("encode_items", "sort_items",
                                                        "treelet_block")}
    b.build(9, 2, graph=True)
    sah = b.sah()
    assert np.array_equal(bits(sah), bits(eager))
    order = np.argsort(sizes, kind="stable")
    pick = np.unique(np.concatenate([order[np.linspace(0, 4095, 62).astype(int)],
                                     [order[0], order[-1]]]))
    assert pick.size >= 62 and order[-1] in pick
    for m in pick:
        vm = v[3 * int(off[m]): 3 * int(off[m + 1])]
        mn, mx = oracle.scene_bbox(vm)
        ref = oracle.lbvh(vm, mn, mx, 30)
        oracle.atrbvh(ref, 9, 2)
        tm = b.extract(int(m))
        assert_tree_equal(tm.download(), tree_arrays(ref), f"mesh{m}")
        tm.close()
        assert bits(sah[m]) == bits(oracle.sah(ref)), m
    base = b.device_vertices()
    plain = np.zeros(4096, np.float32)
    for m in range(4096):
        s = ab.Scene.from_device(ctx, base + 48 * int(off[m]), int(sizes[m]))
        t = ab.LBVHBuilder(False).BuildTree(s)
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
        plain[m] = t.SAH()
        t.close()
        s.close()
    assert np.array_equal(bits(plain), bits(sah))
    b.close()


# ------------------------------------------------------------------ batched differentiable
def _centroid_rays(v, off, per, seed):
    rng = np.random.default_rng(seed)
    meshes = off.size - 1
    out = np.zeros((meshes * per, 8), np.float32)
    for m in range(meshes):
        vm = v[3 * int(off[m]): 3 * int(off[m + 1]), :3].reshape(-1, 3, 3)
        cen = vm[rng.integers(0, vm.shape[0], per)].mean(axis=1)
        r = out[m * per:(m + 1) * per]
        r[:, 0] = cen[:, 0]
        r[:, 1] = cen[:, 1]
        r[:, 2] = vm[:, :, 2].min() - 0.1
        r[:, 6] = 1.0
        r[:, 7] = np.inf
    return out


def _tree_of(b, m):
    t = b.extract(m)
    out = t.download()
    t.close()
    return out


def _download_vertices(ctx, b, n):
    out = np.empty((3 * n, 4), np.float32)
    ab.check(ab.lib().atr_memcpy_d2h(ctx.h, out.ctypes.data, b.device_vertices(), out.nbytes))
    return out


def test_c5_batched_adjoint_refit_and_rebuild(ctx):
    from adjoint_ref import rel_err, tuv_adjoint_fp64

    meshes, per = 64, 512 * 512
    off = ab.gen_batch_sizes(5, meshes, 1024, 524288)
    n = int(off[-1])
    v0 = ab.gen_soup_batch(5000, off)
    vp = v0.copy()
    vp[:, :3] += np.random.default_rng(5).normal(0.0, 0.002, (vp.shape[0], 3)).astype(np.float32)
    rays = _centroid_rays(v0, off, per, 6)
    R = rays.shape[0]
    d_rays = ctx.to_device(rays)
    bt = ab.Batch(ctx, v0, off)
    bt.build(9, 2, graph=False)
    d_target = ctx.buffer(R * 16)
    bt.trace(d_rays, per, d_target)
    ctx.sync()
    bt.close()
    b = ab.Batch(ctx, vp, off)
    b.build(9, 2, graph=True)
    d_hits = ctx.buffer(R * 16)
    b.trace(d_rays, per, d_hits)
    hits = d_hits.download(np.int32, (R, 4))
    # batched adjoint of random upstream gradients against the FP64 adjoint
    g_tuv = np.random.default_rng(9).normal(size=(R, 3)).astype(np.float32)
    d_g = ctx.to_device(g_tuv)
    d_grad = ctx.to_device(np.zeros((3 * n, 3), np.float32))
    b.hit_adjoint(d_rays, d_hits, R, d_g, d_grad)
    grad = d_grad.download(np.float32, (3 * n, 3))
    sizes = np.diff(off)
    check = [0, int(np.argmin(sizes)), int(np.argmax(sizes)), meshes // 2]
    for m in check:
        lo, hi = int(off[m]), int(off[m + 1])
        rs = slice(m * per, (m + 1) * per)
        h = hits[rs, 0].copy()
        assert (h >= 0).mean() > 0.5
        h[h >= 0] -= lo
        verts = vp[3 * lo: 3 * hi, :3].astype(np.float64)
        index = np.arange(3 * (hi - lo)).reshape(-1, 3)
        ref = tuv_adjoint_fp64(verts, index, rays[rs], h, g_tuv[rs].astype(np.float64))
        e = rel_err(grad[3 * lo: 3 * hi], ref)
        assert e < 1e-4, (m, e)
    # a few optimisation steps: same-triangle depth loss, adjoint, capped soup
    # step, batched refit, device SAH, selective rebuild of meshes over 1.10x
    # their post-build SAH (and of meshes 0 and 63 at the last step)
    d_sah = ctx.buffer(4 * meshes)
    b.sah_device(d_sah)
    built = d_sah.download(np.float32, meshes).copy()
    d_loss = ctx.to_device(np.zeros(3, np.float64))
    d_grad.upload(np.zeros((3 * n, 3), np.float32))
    g_up = ctx.buffer(R * 12)
    for k in range(3):
        b.trace(d_rays, per, d_hits)
        ctx.depth_loss(d_hits, d_target, R, g_up, d_loss, offset=k, same_triangle=True)
        b.hit_adjoint(d_rays, d_hits, R, g_up, d_grad)
        b.step(d_grad, 0.002, 0.0005)
        before = [_tree_of(b, m) for m in (0, 31, 63)]
        b.refit()
        b.sah_device(d_sah)
        fast = d_sah.download(np.float32, meshes)
        exact = b.sah()
        # the reference-order fp32 running sum carries ~1e-5 relative rounding
        # on the largest meshes; the device sum accumulates in double
        assert np.all(np.abs(fast - exact) <= 1e-4 * exact), k
        for m, bt_ in zip((0, 31, 63), before):  # refit keeps every topology
            a = _tree_of(b, m)
            assert np.array_equal(a["left"], bt_["left"]) and np.array_equal(a["data"], bt_["data"])
        flagged = set(np.nonzero(fast > np.float32(1.10) * built)[0].tolist())
        if k == 2:
            flagged |= {0, 63}
        if flagged:
            fl = sorted(flagged)
            b.rebuild(fl, 9, 2)
            cur = _download_vertices(ctx, b, n)
            fresh = ab.Batch(ctx, cur, off)
            fresh.build(9, 2, graph=False)
            for m in fl:
                assert_tree_equal(_tree_of(b, m), _tree_of(fresh, m), f"rebuilt{m}")
            fresh.close()
            b.sah_device(d_sah)
            built[fl] = d_sah.download(np.float32, meshes)[fl]
    losses = d_loss.download(np.float64, 3)
    assert losses[2] < losses[0], losses
    b.close()


# ------------------------------------------------------------------ sort give-up path
def test_sort_give_up_keeps_later_kernels_in_bounds():
    """A context whose bounded waits give up at once (spin limit 0): the sort
    may abandon its look-back, the build then ignores the sort's output and
    makes the input-order tree (well formed), ATRBVH, collapse and traversal
    run on it, atr_sync reports the give-up once, and the next sync is clean."""
    os.environ["ATRBVH_SPIN_LIMIT"] = "0"
    try:
        c2 = ab.Context(0)
    finally:
        del os.environ["ATRBVH_SPIN_LIMIT"]
    n = 1 << 20
    v = ab.gen_gmm(12, n)
    s = c2.scene(v)
    mn, mx = s.bounds()
    t = ab.LBVHBuilder(True).BuildTree(s)
    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
    c, _ = ab.BVHTreeCollapser().Collapse(t, want_sah=False)
    g = 256
    d_rays = c2.buffer(g * g * 32)
    d_res = c2.buffer(g * g * 16)
    c2.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    c.trace(d_rays, g * g, d_res)
    dev = t.download()
    check_tree_invariants(dev, n)
    gave_up = False
    try:
        c2.sync()
    except ab.AtrError as e:
        assert e.status == 5 and "gave up" in str(e)
        gave_up = True
    c2.sync()  # reported once, then clear
    ref_keys = np.zeros(n, np.uint64) if gave_up else None
    if gave_up:
        # the input-order tree: the radix tree of n equal keys, refit and optimised
        ref = oracle.karras(ref_keys, np.arange(n, dtype=np.uint32))
        oracle.refit(v, ref)
    else:
        ref = oracle.lbvh(v, mn, mx, 63)
    oracle.atrbvh(ref, 9, 2)
    assert_tree_equal(dev, tree_arrays(ref), "give-up" if gave_up else "no give-up")
    print(f"[give-up] sort gave up: {gave_up}", flush=True)
    c.close()
    t.close()
    s.close()
    c2.close()
