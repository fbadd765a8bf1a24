"""GPU parity of batched builds: every mesh of a forest built in one launch
sequence (eager or replayed from a captured HIP graph) is bit-identical to the
oracle's LBVH (30-bit) + ATRBVH of that mesh alone, and to the plain per-mesh
path of the library."""
from __future__ import annotations

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import assert_tree_equal, bits, tree_arrays

pytestmark = pytest.mark.gpu


def _uneven_batch(seed: int, meshes: int, lo: int, hi: int, extra_small=True):
    off = ab.gen_batch_sizes(seed, meshes, lo, hi)
    if extra_small:  # the smallest legal meshes and sizes around tile edges
        sizes = np.diff(off).astype(np.int64)
        sizes[:6] = [2, 3, 255, 256, 257, 4097][: min(6, meshes)]
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    v = ab.gen_soup_batch(1000 * seed, off)
    return v, off


def _oracle_mesh(v, off, m, treelet=9, iters=2):
    vm = v[3 * int(off[m]): 3 * int(off[m + 1])]
    mn, mx = oracle.scene_bbox(vm)
    t = oracle.lbvh(vm, mn, mx, 30)
    oracle.atrbvh(t, treelet, iters)
    return t


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_oracle_per_mesh(ctx, graph):
    v, off = _uneven_batch(3, 40, 2, 6000)
    b = ab.Batch(ctx, v, off)
    b.build(9, 2, graph=graph)
    sah = b.sah()
    for m in range(b.meshes):
        t = b.extract(m)
        dev = t.download()
        t.close()
        ref = _oracle_mesh(v, off, m)
        assert_tree_equal(dev, tree_arrays(ref), f"mesh{m}")
        assert bits(sah[m]) == bits(oracle.sah(ref)), m
    b.close()


def test_batch_graph_replays_are_identical(ctx):
    """The captured sequence replayed several times, with eager library work
    in between, gives the eager result every time."""
    v, off = _uneven_batch(4, 64, 16, 20000)
    b = ab.Batch(ctx, v, off)
    b.build(9, 2, graph=False)
    want = b.sah()
    want_trees = [b.extract(m).download() for m in (0, 17, 63)]
    other = ctx.scene(ab.gen_soup(99, 5000))
    for r in range(4):
        b.build(9, 2, graph=True)
        t = ab.LBVHBuilder(False).BuildTree(other)  # unrelated eager work between replays
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
        t.close()
        got = b.sah()
        assert np.array_equal(bits(got), bits(want)), r
        for m, w in zip((0, 17, 63), want_trees):
            tt = b.extract(m)
            assert_tree_equal(tt.download(), w, f"replay{r}.mesh{m}")
            tt.close()
    other.close()
    b.close()


def test_batch_given_boxes_and_other_treelet(ctx):
    v, off = _uneven_batch(5, 12, 100, 3000, extra_small=False)
    boxes = []
    for m in range(12):
        vm = v[3 * int(off[m]): 3 * int(off[m + 1])]
        mn, mx = oracle.scene_bbox(vm)
        boxes.append(np.concatenate([mn, mx]))
    b = ab.Batch(ctx, v, off, np.array(boxes, np.float32))
    b.build(7, 3, graph=True)
    for m in range(12):
        ref = _oracle_mesh(v, off, m, 7, 3)
        t = b.extract(m)
        assert_tree_equal(t.download(), tree_arrays(ref), f"mesh{m}")
        t.close()
    b.close()


def test_batch_equals_plain_loop_at_scale(ctx):
    """1,024 meshes of 64..65,536 triangles (about 10M triangles): per-mesh SAH
    of the batch equals the plain per-mesh library path bit for bit."""
    off = ab.gen_batch_sizes(3, 1024, 64, 65536)
    v = ab.gen_soup_batch(3000, off)
    b = ab.Batch(ctx, v, off)
    b.build(9, 2, graph=True)
    sah = b.sah()
    for m in range(0, 1024, 7):
        s = ctx.scene(v[3 * int(off[m]): 3 * int(off[m + 1])])
        t = ab.LBVHBuilder(False).BuildTree(s)
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
        assert bits(np.float32(t.SAH())) == bits(sah[m]), m
        t.close()
        s.close()
    b.close()


def _batch_rays(off, v, per_side, seed=11):
    """per_side^2 rays per mesh along +z, each aimed at the centroid of a
    random triangle of its mesh (so most rays hit: a soup of small blobs
    covers little of its own box)."""
    rng = np.random.default_rng(seed)
    rays = []
    for m in range(off.size - 1):
        vm = v[3 * int(off[m]): 3 * int(off[m + 1]), :3].reshape(-1, 3, 3)
        pick = rng.integers(0, vm.shape[0], per_side * per_side)
        cen = vm[pick].mean(axis=1)
        r = np.zeros((per_side * per_side, 8), np.float32)
        r[:, 0] = cen[:, 0]
        r[:, 1] = cen[:, 1]
        r[:, 2] = vm[:, :, 2].min() - 0.1
        r[:, 6] = 1.0
        r[:, 7] = np.inf
        rays.append(r)
    return np.concatenate(rays)


def test_batch_trace_matches_per_mesh_collapsed_trace(ctx):
    v, off = _uneven_batch(6, 10, 500, 20000, extra_small=False)
    b = ab.Batch(ctx, v, off)
    b.build(9, 2, graph=True)
    per = 64
    rays = _batch_rays(off, v, per)
    d_rays = ctx.to_device(rays)
    d_hits = ctx.buffer(rays.shape[0] * 16)
    b.trace(d_rays, per * per, d_hits)
    hits = d_hits.download(np.int32, (rays.shape[0], 4))
    ctx.sync()
    from helpers import compare_hits
    for m in range(10):
        s = ctx.scene(v[3 * int(off[m]): 3 * int(off[m + 1])])
        t = b.extract(m)
        ab.lib().atr_tree_bind_scene(t.h, s.h)
        c, _ = ab.BVHTreeCollapser().Collapse(t)
        rm = rays[m * per * per:(m + 1) * per * per]
        dr = ctx.to_device(rm)
        dh = ctx.buffer(rm.shape[0] * 16)
        c.trace(dr, rm.shape[0], dh)
        ref = dh.download(np.int32, (rm.shape[0], 4))
        got = hits[m * per * per:(m + 1) * per * per].copy()
        got[:, 0] = np.where(got[:, 0] >= 0, got[:, 0] - int(off[m]), -1)
        assert (ref[:, 0] >= 0).mean() > 0.5
        compare_hits(got, ref)
        c.close()
        t.close()
        s.close()
    b.close()


def test_batch_refit_step_and_selective_rebuild(ctx):
    """Soup step + batched refit give the boxes of a recursive union over the
    moved vertices; rebuilding a subset of meshes gives exactly the trees a
    fresh batch of the moved vertices gives for them and leaves the others."""
    v, off = _uneven_batch(7, 8, 300, 9000, extra_small=False)
    b = ab.Batch(ctx, v, off)
    b.build(9, 2, graph=False)
    before = [b.extract(m).download() for m in range(8)]
    n = int(off[-1])
    g = np.random.default_rng(2).normal(size=(3 * n, 3)).astype(np.float32)
    d_g = ctx.to_device(g)
    b.step(d_g, 0.001)
    b.refit()
    moved = v.copy()
    moved[:, :3] -= np.float32(0.001) * g
    assert not np.any(d_g.download(np.float32, (3 * n, 3)))
    for m in (0, 5):
        t = b.extract(m).download()
        assert np.array_equal(t["parent"], before[m]["parent"])  # topology kept
        vm = moved[3 * int(off[m]): 3 * int(off[m + 1])]
        tri = vm[:, :3].reshape(-1, 3, 3)
        N = int(off[m + 1] - off[m])
        lo = np.zeros((2 * N - 1, 3), np.float32)
        hi = np.zeros((2 * N - 1, 3), np.float32)
        leaves = np.arange(N - 1, 2 * N - 1)
        lo[leaves] = tri[t["data"][leaves]].min(axis=1)
        hi[leaves] = tri[t["data"][leaves]].max(axis=1)
        order, stack = [], [0]
        while stack:
            x = stack.pop()
            if x < N - 1:
                order.append(x)
                stack += [t["left"][x], t["right"][x]]
        for x in reversed(order):
            lo[x] = np.minimum(lo[t["left"][x]], lo[t["right"][x]])
            hi[x] = np.maximum(hi[t["left"][x]], hi[t["right"][x]])
        assert np.array_equal(bits(t["bbmin"][:, :3]), bits(lo))
        assert np.array_equal(bits(t["bbmax"][:, :3]), bits(hi))
    b.rebuild([1, 6])
    fresh = ab.Batch(ctx, moved, off)
    fresh.build(9, 2, graph=False)
    for m in range(8):
        t = b.extract(m).download()
        if m in (1, 6):
            assert_tree_equal(t, fresh.extract(m).download(), f"rebuilt{m}")
        else:
            assert np.array_equal(t["parent"], before[m]["parent"])
            assert np.array_equal(t["left"], before[m]["left"])
    fresh.close()
    b.close()


def test_autotuner_launch_shapes_keep_trees(ctx):
    """Every launch shape gives identical trees (the tuner checks per-mesh SAH;
    here whole trees of a few meshes are compared too)."""
    from atrbvh_amd.autotune import tune_launch, tune_treelet

    v, off = _uneven_batch(8, 30, 64, 30000)
    b = ab.Batch(ctx, v, off)
    b.build(9, 2, graph=False)
    ref = [b.extract(m).download() for m in (0, 5, 29)]
    res = tune_launch(ctx, b, reps=1)
    assert len(res["table"]) == 12 and b.config() == {k: res["best"][k] for k in
                                                      ("encode_items", "sort_items",
                                                       "treelet_block")}
    for enc, srt, blk in ((8, 8, 64), (16, 8, 256), (8, 16, 128)):
        b.set_config(enc, srt, blk)
        b.build(9, 2, graph=True)
        for m, r in zip((0, 5, 29), ref):
            assert_tree_equal(b.extract(m).download(), r, f"shape{enc}/{srt}/{blk}.mesh{m}")
    tr = tune_treelet(ctx, b, sah_tol=0.05, reps=1)
    assert tr["best"]["mean_sah"] <= tr["sah_limit"]
    b.close()


def test_batch_rebuild_keeps_given_boxes(ctx):
    """A batch built on caller boxes rebuilds its meshes in those same frames:
    after a step that moves vertices, a selective rebuild gives exactly the
    trees of a fresh batch of the moved vertices on the same boxes (computed
    boxes would change the Morton quantisation and the trees)."""
    v, off = _uneven_batch(9, 6, 200, 5000, extra_small=False)
    boxes = []
    for m in range(6):
        vm = v[3 * int(off[m]): 3 * int(off[m + 1])]
        mn, mx = oracle.scene_bbox(vm)
        boxes.append(np.concatenate([mn - 0.25, mx + 0.25]))
    boxes = np.array(boxes, np.float32)
    b = ab.Batch(ctx, v, off, boxes)
    b.build(9, 2, graph=False)
    n = int(off[-1])
    g = np.random.default_rng(4).normal(size=(3 * n, 3)).astype(np.float32)
    d_g = ctx.to_device(g)
    b.step(d_g, 0.01)
    moved = v.copy()
    moved[:, :3] -= np.float32(0.01) * g
    b.rebuild([2, 4])
    fresh = ab.Batch(ctx, moved, off, boxes)
    fresh.build(9, 2, graph=False)
    for m in (2, 4):
        t = b.extract(m)
        f = fresh.extract(m)
        assert_tree_equal(t.download(), f.download(), f"given-box rebuild {m}")
        t.close()
        f.close()
    vm = moved[3 * int(off[2]): 3 * int(off[3])]
    ref = oracle.lbvh(vm, boxes[2, :3], boxes[2, 3:], 30)
    oracle.atrbvh(ref, 9, 2)
    t = b.extract(2)
    assert_tree_equal(t.download(), tree_arrays(ref), "given-box rebuild vs oracle")
    t.close()
    fresh.close()
    b.close()
