"""Pins of the treelet passes' composition in the oracle (formTreelet ->
restructure -> rewrite, and the bottom-up pass around it) by what the
mathematics fixes rather than by restating the code:

- ATRBVH with one treelet: on an n-triangle scene with treelet size n and one
  iteration (gamma = n), only the root roots a treelet and that treelet holds
  every triangle.  The result must be greedy minimum-union-area clustering of
  the n leaves, each merge k (k = 1..n-1) landing in the internal node that
  formTreelet expanded (n-1-k)-th (largest stored area first), applied only
  when its SAH beats the tree's (Treelet.cuh:57-114,
  AgglomerativeTreelet.cu:473-546, :190-212).
- TRBVH with one treelet: 7 triangles, TRBVH (7, 1): the rewritten tree's cost
  under the DP's own objective (split cost or collapse cost with the treelet's
  triangle count, TRBVH.cu:127-128,447-448) equals the minimum over all 10,395
  binary trees of the 7 leaves, and where that optimum is unique the topology
  is it (TRBVH.cu:222-327).
- Treelet counts: a pass with threshold gamma processes exactly the nodes with
  at least gamma triangles (closed form on its input tree), and the
  optimisers' iterations are passes at gamma = n, 2n, 4n
  (TRBVHOptimizer.cpp:91,105, AgglomerativeTreeletOptimizer.cpp:73-91).

tests/test_oracle_mutations.py rebuilds the oracle with planted mistakes and
shows that these tests fail on each of them.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle

F32 = np.float32
CI, CT = F32(1.2), F32(1.0)


def _area(mn, mx) -> np.float32:
    s = (mx - mn).astype(np.float32)
    return F32(F32(2.0) * F32(F32(F32(s[0] * s[1]) + F32(s[0] * s[2])) + F32(s[1] * s[2])))


def _scene(seed: int, n: int, spread: bool = True) -> np.ndarray:
    """n small triangles at random centres (spread) or n triangles of about
    the same size around one point (overlapping: their boxes nearly coincide,
    so splitting costs more than keeping four or more of them together),
    with random shapes, so areas and union areas are distinct."""
    rng = np.random.default_rng(seed)
    v = np.zeros((3 * n, 4), np.float32)
    if spread:
        c = rng.random((n, 3))
        size = rng.uniform(0.01, 0.08, n)
        v[:, :3] = (c[:, None, :] + size[:, None, None] * rng.normal(size=(n, 3, 3))).reshape(-1, 3)
    else:
        base = np.array([[0.0, 0.0, 0.0], [0.2, 0.0, 0.05], [0.05, 0.2, 0.2]])
        v[:, :3] = (0.5 + base[None] + 0.01 * rng.normal(size=(n, 3, 3))).reshape(-1, 3)
    return v


def _lbvh(v):
    mn, mx = oracle.scene_bbox(v)
    return oracle.lbvh(v, mn, mx, 30)


def _bottom_up_sah(t, ci=CI, ct=CT, collapse_tris=None):
    """nodesSah of the bottom-up pass: leaves ct * area, internal nodes
    ci * area + left + right (TRBVH.cu:363-409; AgglomerativeTreelet.cu:467)."""
    n = t.n
    sah = np.zeros(2 * n - 1, np.float32)

    def rec(x):
        if x >= n - 1:
            sah[x] = F32(ct * t.area[x])
            return
        rec(t.left[x])
        rec(t.right[x])
        sah[x] = F32(F32(F32(ci * t.area[x]) + sah[t.left[x]]) + sah[t.right[x]])

    rec(0)
    return sah


def _expansion_order(t, n):
    """formTreelet from the root: repeatedly expand the treelet leaf that is
    an internal node with the largest stored area (distinct areas here)."""
    internal = [0]
    frontier = [int(t.left[0]), int(t.right[0])]
    while len(internal) < n - 1:
        cand = [x for x in frontier if x < t.n - 1]
        areas = [t.area[x] for x in cand]
        assert len(set(areas)) == len(areas), "tied areas: pick another seed"
        x = cand[int(np.argmax(areas))]
        internal.append(x)
        frontier.remove(x)
        frontier += [int(t.left[x]), int(t.right[x])]
    return internal


def _children_sets(t):
    return {p: frozenset((int(t.left[p]), int(t.right[p]))) for p in range(t.n - 1)}


# ------------------------------------------------------------------ ATRBVH, one treelet
@pytest.mark.parametrize("n", [3, 4, 5, 7, 9, 12, 16, 24, 32])
def test_atrbvh_single_treelet_is_greedy_clustering(n):
    checked = rewritten = 0
    for seed in range(6):
        v = _scene(100 * n + seed, n)
        t0 = _lbvh(v)
        order = _expansion_order(t0, n)
        leaf_sah = {x: F32(CT * t0.area[x]) for x in range(n - 1, 2 * n - 1)}
        boxes = {x: (t0.bbmin[x, :3].copy(), t0.bbmax[x, :3].copy()) for x in leaf_sah}
        clusters = dict(leaf_sah)
        kids = {}
        for k in range(1, n):
            ids = sorted(clusters)
            pairs = [(_area(np.minimum(boxes[a][0], boxes[b][0]),
                            np.maximum(boxes[a][1], boxes[b][1])), a, b)
                     for a, b in itertools.combinations(ids, 2)]
            pairs.sort(key=lambda p: p[0])
            assert len(pairs) == 1 or pairs[0][0] != pairs[1][0], "tied union areas"
            ua, a, b = pairs[0]
            node = order[n - 1 - k]
            boxes[node] = (np.minimum(boxes[a][0], boxes[b][0]), np.maximum(boxes[a][1], boxes[b][1]))
            clusters[node] = F32(F32(clusters[a] + clusters[b]) + F32(CI * ua))
            del clusters[a], clusters[b]
            kids[node] = frozenset((a, b))
        assert list(clusters) == [0]
        new_cost = clusters[0]
        old_cost = _bottom_up_sah(t0)[0]
        t = t0.copy()
        assert oracle.treelet_pass(t, "atrbvh", n, n) == 1  # only the root has n triangles
        got = _children_sets(t)
        if new_cost < old_cost:
            assert got == kids, (n, seed)
            for p in kids:
                assert np.array_equal(t.bbmin[p, :3], boxes[p][0])
                assert np.array_equal(t.bbmax[p, :3], boxes[p][1])
                assert t.area[p] == _area(*boxes[p])
                for c in kids[p]:
                    assert t.parent[c] == p
            rewritten += 1
        else:
            assert got == _children_sets(t0)
        checked += 1
    assert rewritten >= 2, "the rewrite branch was not exercised"


# ------------------------------------------------------------------ TRBVH, one treelet
def _all_trees(items):
    if len(items) == 1:
        yield items[0]
        return
    first, rest = items[0], items[1:]
    for k in range(len(rest)):
        for comb in itertools.combinations(rest, k):
            other = [x for x in rest if x not in comb]
            for lt in _all_trees([first, *comb]):
                for rt in _all_trees(other):
                    yield (lt, rt)


def _leafset(t):
    return frozenset([t]) if isinstance(t, int) else _leafset(t[0]) | _leafset(t[1])


@pytest.mark.parametrize("spread", [True, False])
def test_trbvh_single_treelet_is_exhaustive_optimum(spread):
    """Overlapping triangles make the collapse branch (ct * area * 7) win for
    the larger subsets; spread ones never collapse and have a unique optimal
    topology."""
    n = 7
    rewritten = unique_checked = 0
    collapsing = set()
    for seed in range(8):
        v = _scene(7000 + seed + (0 if spread else 500), n, spread)
        t0 = _lbvh(v)
        leaves = list(range(n - 1, 2 * n - 1))
        lo = {x: t0.bbmin[x, :3] for x in leaves}
        hi = {x: t0.bbmax[x, :3] for x in leaves}
        nsah = {x: F32(CT * t0.area[x]) for x in leaves}
        area_of = {}

        def sarea(s):
            if s not in area_of:
                area_of[s] = _area(np.min([lo[x] for x in s], axis=0),
                                   np.max([hi[x] for x in s], axis=0))
            return area_of[s]

        memo = {}

        def cost(tr):
            if isinstance(tr, int):
                return nsah[tr]
            if tr in memo:
                return memo[tr]
            a = sarea(_leafset(tr))
            split = F32(F32(CI * a) + F32(cost(tr[0]) + cost(tr[1])))
            collapse = F32(F32(CT * a) * F32(n))
            c = min(split, collapse)
            memo[tr] = c
            return c

        trees = list(_all_trees(leaves))
        assert len(trees) == 10395
        costs = np.array([cost(tr) for tr in trees], np.float32)
        best = costs.min()
        old = _bottom_up_sah(t0)[0]
        t = t0.copy()
        assert oracle.treelet_pass(t, "trbvh", 7, 7) == 1

        def as_tuple(x):
            return x if x >= n - 1 else (as_tuple(int(t.left[x])), as_tuple(int(t.right[x])))

        got = as_tuple(0)
        if best < old:
            rewritten += 1
            assert cost(got) == best, (seed, cost(got), best)
            winners = [tr for tr, c in zip(trees, costs) if c == best]

            def shape(tr):
                return _leafset(tr) if isinstance(tr, int) else frozenset(
                    [frozenset([shape(tr[0]), shape(tr[1])])])

            def collapses(tr):  # nodes of the tree where the collapse cost is the smaller
                if isinstance(tr, int):
                    return 0
                a = sarea(_leafset(tr))
                split = F32(F32(CI * a) + F32(cost(tr[0]) + cost(tr[1])))
                return int(F32(F32(CT * a) * F32(n)) < split) + collapses(tr[0]) + collapses(tr[1])

            if collapses(got):
                collapsing.add(seed)
            if len(winners) == 1:
                unique_checked += 1
                assert shape(got) == shape(winners[0])
        else:
            assert _children_sets(t) == _children_sets(t0)
    assert rewritten >= 3
    if spread:
        assert unique_checked >= 3 and not collapsing
    else:
        assert collapsing, "the collapse branch was never the cheaper one"


# ------------------------------------------------------------------ treelet counts
def _subtree_sizes(t):
    n = t.n
    size = np.zeros(2 * n - 1, np.int64)

    def rec(x):
        if x >= n - 1:
            size[x] = 1
            return 1
        size[x] = rec(t.left[x]) + rec(t.right[x])
        return size[x]

    import sys

    sys.setrecursionlimit(100000)
    rec(0)
    return size


@pytest.mark.parametrize("kind,treelet", [("atrbvh", 9), ("atrbvh", 5), ("trbvh", 7)])
def test_treelet_counts_and_gamma_schedule(kind, treelet):
    v = ab.gen_soup(31, 6000)
    t = _lbvh(v)
    base = t.copy()
    counts = []
    for it in range(3):
        gamma = treelet * 2 ** it
        want = int(np.count_nonzero(_subtree_sizes(t) >= gamma))
        got = oracle.treelet_pass(t, kind, treelet, gamma)
        assert got == want, (it, got, want)
        counts.append(got)
    assert counts[0] > counts[1] > counts[2] > 0
    opt = base.copy()
    (oracle.atrbvh if kind == "atrbvh" else oracle.trbvh)(opt, treelet, 3)
    for k in ("parent", "left", "right", "data", "area"):
        assert np.array_equal(getattr(opt, k), getattr(t, k)), k
