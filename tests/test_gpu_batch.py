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
