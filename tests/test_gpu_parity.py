"""GPU parity: every stage of the HIP path against the CPU oracle on the same
seeded inputs (bit-exact for keys, permutations, topology, boxes and areas;
hit ids exact except ties with t within 1e-6 relative)."""
from __future__ import annotations

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import (C1_N, C1_SEED, CAMERA_BASIS, CAMERA_EYE, assert_tree_equal, bits,
                     check_tree_invariants, collapsed_leaf_sets, compare_hits, tree_arrays)

pytestmark = pytest.mark.gpu


def _scene_pair(ctx, v):
    s = ctx.scene(v)
    mn, mx = s.bounds()
    omn, omx = oracle.scene_bbox(v)
    assert np.array_equal(bits(mn), bits(omn)) and np.array_equal(bits(mx), bits(omx))
    return s, mn, mx


# ------------------------------------------------------------------ encode
@pytest.mark.parametrize("bits_", [30, 63])
@pytest.mark.parametrize("n", [1, 3, 4095, 4096, 4097, 70001])
def test_morton_encode(ctx, bits_, n):
    v = ab.gen_soup(7 + n, n)
    mn, mx = oracle.scene_bbox(v)
    dv = ctx.to_device(v)
    kb = 8 if bits_ == 63 else 4
    dk = ctx.buffer(n * kb)
    di = ctx.buffer(n * 4)
    ctx.morton_encode(dv, n, mn, mx, bits_, dk, di)
    keys = dk.download(np.uint64 if bits_ == 63 else np.uint32, n)
    idx = di.download(np.uint32, n)
    rk, ri = oracle.encode(v, mn, mx, bits_)
    assert np.array_equal(keys, rk)
    assert np.array_equal(idx, ri)


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("key_bits", [32, 64])
@pytest.mark.parametrize("n", [2, 17, 4096, 4097, 100003, 1 << 20])
def test_sort_pairs_stable(ctx, key_bits, n):
    rng = np.random.default_rng(n + key_bits)
    dt = np.uint64 if key_bits == 64 else np.uint32
    hi = (1 << 63) if key_bits == 64 else (1 << 30)
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(dt)
    keys[rng.integers(0, n, n // 3)] = keys[0]  # many duplicates
    vals = rng.permutation(n).astype(np.uint32)
    dk = ctx.to_device(keys)
    dvv = ctx.to_device(vals)
    ctx.sort_pairs(key_bits, dk, dvv, n)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(dk.download(dt, n), keys[order])
    assert np.array_equal(dvv.download(np.uint32, n), vals[order])


def test_sort_identity_values(ctx):
    n = 50000
    keys = np.random.default_rng(5).integers(0, 1000, n).astype(np.uint32)
    dk = ctx.to_device(keys)
    ctx.sort_pairs(32, dk, None, n)
    assert np.array_equal(dk.download(np.uint32, n), np.sort(keys, kind="stable"))


# ------------------------------------------------------------------ LBVH
@pytest.mark.parametrize("bits_", [30, 63])
@pytest.mark.parametrize("n", [2, 3, 5, 64, 1000, 4097, C1_N])
def test_lbvh_bit_exact(ctx, bits_, n):
    v = ab.gen_soup(C1_SEED if n == C1_N else 100 + n, n)
    s, mn, mx = _scene_pair(ctx, v)
    t = ab.LBVHBuilder(bits_ == 63).BuildTree(s)
    dev = t.download()
    ref = oracle.lbvh(v, mn, mx, bits_)
    assert_tree_equal(dev, tree_arrays(ref), f"lbvh{bits_}")
    check_tree_invariants(dev, n)
    assert t.SAH() == pytest.approx(oracle.sah(ref), rel=0, abs=0)
    t.close()
    s.close()


def test_lbvh_duplicate_keys(ctx):
    """Identical triangles (equal Morton codes) split by index."""
    base = ab.gen_soup(3, 40)
    v = np.concatenate([base] * 50 + [ab.gen_soup(4, 37)])
    n = v.shape[0] // 3
    for bits_ in (30, 63):
        s, mn, mx = _scene_pair(ctx, v)
        t = ab.LBVHBuilder(bits_ == 63).BuildTree(s)
        ref = oracle.lbvh(v, mn, mx, bits_)
        assert_tree_equal(t.download(), tree_arrays(ref), "dup")
        t.close()
        s.close()


# ------------------------------------------------------------------ treelets
@pytest.mark.parametrize("treelet,iters", [(9, 2), (7, 3), (3, 1), (16, 1), (32, 1)])
def test_atrbvh_bit_exact(ctx, treelet, iters):
    n = C1_N if (treelet, iters) == (9, 2) else 20000
    v = ab.gen_soup(C1_SEED if n == C1_N else 55, n)
    s, mn, mx = _scene_pair(ctx, v)
    t = ab.LBVHBuilder(False).BuildTree(s)
    ab.AgglomerativeTreeletOptimizer(treelet, iters).Optimize(t)
    dev = t.download()
    ref = oracle.lbvh(v, mn, mx, 30)
    oracle.atrbvh(ref, treelet, iters)
    assert_tree_equal(dev, tree_arrays(ref), f"atrbvh{treelet}_{iters}")
    check_tree_invariants(dev, n)
    t.close()
    s.close()


def _tie_scene(kind: str) -> np.ndarray:
    """Scenes full of exact ties: `dup` repeats one 40-triangle soup 60 times
    (equal Morton codes, boxes, areas and union areas); `grid` puts one
    triangle shape on a 16 x 16 x 12 lattice (equal areas, and equal union
    areas for every pair at the same lattice offset)."""
    if kind == "dup":
        base = ab.gen_soup(3, 40)
        return np.ascontiguousarray(np.concatenate([base] * 60 + [ab.gen_soup(4, 37)]))
    g = np.stack(np.meshgrid(np.arange(16), np.arange(16), np.arange(12), indexing="ij"), -1)
    p = g.reshape(-1, 3).astype(np.float32) * np.float32(0.0625)
    shape = np.array([[0.0, 0.0, 0.0], [0.03125, 0.0, 0.0], [0.0, 0.03125, 0.015625]],
                     np.float32)
    v = np.zeros((p.shape[0] * 3, 4), np.float32)
    v[:, :3] = (p[:, None, :] + shape[None, :, :]).reshape(-1, 3)
    return v


@pytest.mark.parametrize("kind", ["dup", "grid"])
@pytest.mark.parametrize("opt", ["atrbvh9", "atrbvh5", "atrbvh16", "trbvh"])
def test_treelets_and_collapse_under_ties(ctx, kind, opt):
    """Treelet restructuring and collapse where many distances, areas and
    keys are exactly equal, so every tie rule of the reference's reductions
    (findLeafToExpand, findMinimum, reduceOptimal) decides the tree."""
    v = _tie_scene(kind)
    n = v.shape[0] // 3
    s, mn, mx = _scene_pair(ctx, v)
    t = ab.LBVHBuilder(False).BuildTree(s)
    ref = oracle.lbvh(v, mn, mx, 30)
    if opt == "trbvh":
        ab.TRBVHOptimizer(7, 3).Optimize(t)
        oracle.trbvh(ref, 7, 3)
    else:
        size = int(opt[6:])
        ab.AgglomerativeTreeletOptimizer(size, 2).Optimize(t)
        oracle.atrbvh(ref, size, 2)
    dev = t.download()
    assert_tree_equal(dev, tree_arrays(ref), f"{kind}-{opt}")
    check_tree_invariants(dev, n)
    c, sah = ab.BVHTreeCollapser().Collapse(t)
    rc, rsah = oracle.collapse(ref)
    cd = c.download()
    assert bits(np.float32(sah)) == bits(np.float32(rsah))
    for k in ("parent", "left", "right", "data", "count"):
        assert np.array_equal(cd[k], getattr(rc, k)), k
    assert cd["used"] == rc.used


@pytest.mark.parametrize("kind", ["dup", "grid"])
def test_63bit_build_and_atrbvh_under_ties(ctx, kind):
    """The 63-bit Morton path (the benchmark's) on the same tie scenes: equal
    64-bit keys fall back to the index prefix in the split search, the sort
    must keep equal keys in index order, and ATRBVH (9, 2) + collapse must
    then match the oracle bit for bit."""
    v = _tie_scene(kind)
    n = v.shape[0] // 3
    s, mn, mx = _scene_pair(ctx, v)
    t = ab.LBVHBuilder(True).BuildTree(s)
    ref = oracle.lbvh(v, mn, mx, 63)
    assert_tree_equal(t.download(), tree_arrays(ref), f"{kind}-lbvh63")
    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
    oracle.atrbvh(ref, 9, 2)
    dev = t.download()
    assert_tree_equal(dev, tree_arrays(ref), f"{kind}-lbvh63-atrbvh9")
    check_tree_invariants(dev, n)
    c, sah = ab.BVHTreeCollapser().Collapse(t)
    rc, rsah = oracle.collapse(ref)
    cd = c.download()
    assert bits(np.float32(sah)) == bits(np.float32(rsah))
    for k in ("parent", "left", "right", "data", "count"):
        assert np.array_equal(cd[k], getattr(rc, k)), k
    assert cd["used"] == rc.used
    c.close()
    t.close()
    s.close()
    assert np.array_equal(cd["tri_index"][: rc.used], rc.tri_index[: rc.used])
    c.close()
    t.close()
    s.close()


@pytest.mark.parametrize("bits_,n", [(30, C1_N), (63, 30000), (30, 9), (30, 300)])
def test_trbvh_bit_exact(ctx, bits_, n):
    v = ab.gen_soup(C1_SEED if n == C1_N else 77 + n, n)
    s, mn, mx = _scene_pair(ctx, v)
    t = ab.LBVHBuilder(bits_ == 63).BuildTree(s)
    ab.TRBVHOptimizer(7, 3).Optimize(t)
    dev = t.download()
    ref = oracle.lbvh(v, mn, mx, bits_)
    oracle.trbvh(ref, 7, 3)
    assert_tree_equal(dev, tree_arrays(ref), "trbvh")
    check_tree_invariants(dev, n)
    t.close()
    s.close()


# ------------------------------------------------------------------ collapse + trace
def _c1_pipeline(ctx):
    v = ab.gen_soup(C1_SEED, C1_N)
    s, mn, mx = _scene_pair(ctx, v)
    t = ab.LBVHBuilder(False).BuildTree(s)
    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
    c, sah = ab.BVHTreeCollapser().Collapse(t)
    ref = oracle.lbvh(v, mn, mx, 30)
    oracle.atrbvh(ref, 9, 2)
    rc, rsah = oracle.collapse(ref)
    return v, s, t, c, sah, ref, rc, rsah


def test_collapse_bit_exact(ctx):
    v, s, t, c, sah, ref, rc, rsah = _c1_pipeline(ctx)
    dev = c.download()
    assert bits(np.float32(sah)) == bits(np.float32(rsah))
    for k in ("parent", "left", "right", "data", "count"):
        assert np.array_equal(dev[k], getattr(rc, k)), k
    assert dev["used"] == rc.used
    assert np.array_equal(dev["tri_index"][: rc.used], rc.tri_index[: rc.used])
    for k in ("bbmin", "bbmax"):
        assert np.array_equal(bits(dev[k][:, :3]), bits(getattr(rc, k)[:, :3]))
    sets = collapsed_leaf_sets(dev, C1_N)
    tris = sorted(x for s_ in sets.values() for x in s_)
    assert tris == list(range(C1_N))
    c.close()
    t.close()
    s.close()


def test_trace_primary_c1(ctx):
    v, s, t, c, sah, ref, rc, rsah = _c1_pipeline(ctx)
    g = 256
    r = g * g
    d_rays = ctx.buffer(r * 32)
    d_res = ctx.buffer(r * 16)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    c.trace(d_rays, r, d_res)
    rays = d_rays.download(np.float32, (r, 8))
    dev = d_res.download(np.int32, (r, 4))
    rr = oracle.trace(rc, v, rays)
    compare_hits(dev, rr)
    # brute force on a subset pins the oracle traversal itself
    sub = slice(0, r, 97)
    compare_hits(rr[sub], oracle.trace_brute(v, rays[sub]))
    # diffuse secondary rays from the primary hits, against the oracle's
    # restatement of rayGenAOKernel (directions within 2e-6: cosf/sinf)
    d_sec = ctx.buffer(r * 32)
    ctx.rays_diffuse(s, d_rays, d_res, r, 12345, np.inf, d_sec)
    d_res2 = ctx.buffer(r * 16)
    c.trace(d_sec, r, d_res2)
    sec = d_sec.download(np.float32, (r, 8))
    want = oracle.rays_diffuse(v, rays, dev, 12345)
    assert np.array_equal(bits(sec[:, 0:4]), bits(want[:, 0:4]))
    assert np.array_equal(bits(sec[:, 7]), bits(want[:, 7]))
    assert np.max(np.abs(sec[:, 4:7] - want[:, 4:7])) <= 2e-6
    compare_hits(d_res2.download(np.int32, (r, 4)), oracle.trace(rc, v, sec))
    c.close()
    t.close()
    s.close()


def test_trace_any_hit(ctx):
    v, s, t, c, sah, ref, rc, rsah = _c1_pipeline(ctx)
    g = 128
    r = g * g
    d_rays = ctx.buffer(r * 32)
    d_res = ctx.buffer(r * 16)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    c.trace(d_rays, r, d_res, any_hit=True)
    dev = d_res.download(np.int32, (r, 4))
    rays = d_rays.download(np.float32, (r, 8))
    closest = oracle.trace(rc, v, rays)
    assert np.array_equal(dev[:, 0] >= 0, closest[:, 0] >= 0)
    c.close()
    t.close()
    s.close()


# ------------------------------------------------------------------ full size
def test_c2_full_pipeline_bit_exact(ctx):
    """2,097,152-triangle GMM scene: every kernel runs with thousands of
    workgroups, so nodes that cross workgroups (cross-XCD hand-over) are
    exercised at the real shape."""
    n = 2_097_152
    v = ab.gen_gmm(2, n)
    s, mn, mx = _scene_pair(ctx, v)
    t = ab.LBVHBuilder(True).BuildTree(s)
    lb = t.download()
    ref = oracle.lbvh(v, mn, mx, 63)
    assert_tree_equal(lb, tree_arrays(ref), "c2.lbvh")
    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
    dev = t.download()
    oracle.atrbvh(ref, 9, 2)
    assert_tree_equal(dev, tree_arrays(ref), "c2.atrbvh")
    c, sah = ab.BVHTreeCollapser().Collapse(t)
    rc, rsah = oracle.collapse(ref)
    assert bits(np.float32(sah)) == bits(np.float32(rsah))
    cd = c.download()
    for k in ("data", "count"):
        assert np.array_equal(cd[k], getattr(rc, k)), k
    assert np.array_equal(cd["tri_index"][: rc.used], rc.tri_index[: rc.used])
    c.close()
    t.close()
    s.close()


def test_trbvh_large_bit_exact(ctx):
    n = 300_000
    v = ab.gen_gmm(9, n)
    s, mn, mx = _scene_pair(ctx, v)
    t = ab.LBVHBuilder(True).BuildTree(s)
    ab.TRBVHOptimizer(7, 3).Optimize(t)
    ref = oracle.lbvh(v, mn, mx, 63)
    oracle.trbvh(ref, 7, 3)
    assert_tree_equal(t.download(), tree_arrays(ref), "trbvh300k")
    t.close()
    s.close()


def test_repeat_builds_identical(ctx):
    v = ab.gen_gmm(3, 500_000)
    s = ctx.scene(v)
    outs = []
    for _ in range(3):
        t = ab.LBVHBuilder(True).BuildTree(s)
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
        outs.append(t.download())
        t.close()
    for o in outs[1:]:
        assert_tree_equal(o, outs[0], "repeat")
    s.close()


def test_refit_matches_rebuild_boxes(ctx):
    """Refit over a fixed topology after moving vertices equals a recursive
    union of the new leaf boxes (checked against the oracle refit)."""
    n = 200_000
    v = ab.gen_soup(21, n)
    s, mn, mx = _scene_pair(ctx, v)
    t = ab.LBVHBuilder(False).BuildTree(s)
    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
    topo = t.download()
    v2 = v.copy()
    v2[:, :3] += np.float32(0.01) * np.sin(np.arange(v2.shape[0], dtype=np.float32))[:, None]
    s2 = ctx.scene(v2, mn, mx)
    t.refit(s2)
    dev = t.download()
    ref = oracle.Tree(n)
    for k in ("parent", "left", "right", "data"):
        getattr(ref, k)[...] = topo[k]
    oracle.refit(v2, ref)
    assert_tree_equal(dev, tree_arrays(ref), "refit")
    t.close()
    s.close()
    s2.close()


# ------------------------------------------------------------------ C++ facade
@pytest.mark.parametrize("opt,morton64", [("atrbvh", False), ("trbvh", True), ("none", False)])
def test_cpp_facade_matches_oracle(tmp_path, opt, morton64):
    """The Renderer::getCudaBVH call sequence through include/bvhrt_compat.hpp
    (a separate host process on the C ABI) gives the oracle's trees."""
    import os

    from helpers import COMPAT_BIN, run_compat_driver

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if not os.path.exists(os.path.join(root, COMPAT_BIN)):
        pytest.fail(f"{COMPAT_BIN} is not built (the build step compiles it)")
    n = 30000
    v = ab.gen_soup(91, n)
    rc, out, res = run_compat_driver(root, v, morton64, opt, str(tmp_path))
    assert rc == 0, out
    mn, mx = oracle.scene_bbox(v)
    ref = oracle.lbvh(v, mn, mx, 63 if morton64 else 30)
    if opt == "atrbvh":
        oracle.atrbvh(ref, 9, 2)
    elif opt == "trbvh":
        oracle.trbvh(ref, 7, 3)
    for k in ("parent", "left", "right", "data"):
        assert np.array_equal(res[k], getattr(ref, k)), k
    assert bits(res["sah_tree"]) == bits(oracle.sah(ref))
    c, csah = oracle.collapse(ref)
    assert bits(res["sah_collapsed"]) == bits(csah)
    assert res["used"] == c.used
    assert np.array_equal(res["c_data"], c.data) and np.array_equal(res["c_count"], c.count)
    assert np.array_equal(res["tri_index"], c.tri_index[: c.used])


@pytest.mark.parametrize("count", [1, 5, 63, 64, 65, 511, 1000, 70001])
def test_trace_every_ray_once_at_any_count(ctx, count):
    """The persistent traversal hands rays out through interleaved counters
    and wave-local chunks: every ray of a batch of any size is traced exactly
    once, with the result it gets in a full batch (and from the oracle), and
    nothing is written past the batch."""
    v, s, t, c, sah, ref, rc, rsah = _c1_pipeline(ctx)
    g = 512
    r = g * g
    d_rays = ctx.buffer(r * 32)
    ctx.rays_pinhole(CAMERA_EYE, CAMERA_BASIS, 1.0, g, g, np.inf, d_rays)
    rays = d_rays.download(np.float32, (r, 8))
    # take every 3rd ray so that hits and misses mix in small batches
    sel = np.ascontiguousarray(rays[::3][:count])
    assert sel.shape == (count, 8)   # the ray buffer holds the whole batch
    d_sel = ctx.to_device(sel)
    d_out = ctx.to_device(np.full((count + 64, 4), 0x7F7F7F7F, np.int32))
    c.trace(d_sel, count, d_out)
    got = d_out.download(np.int32, (count + 64, 4))
    compare_hits(got[:count], oracle.trace(rc, v, sel))
    assert np.all(got[count:] == 0x7F7F7F7F)
    c.close()
    t.close()
    s.close()
