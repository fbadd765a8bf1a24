"""CPU checks that pin the oracle to the reference's behaviour by means other
than its own formulas: closed forms, brute force, exhaustive enumeration, the
reference's own scheduler (compiled from its source into oracle/_ref, output
committed in tests/golden/) and structural invariants."""
from __future__ import annotations

import itertools
import json
import os

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle
from helpers import camera_rays_host, compare_hits

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F32 = np.float32


# ------------------------------------------------------------------ Morton
def _interleave(x, y, z, bits):
    """Bit interleave by definition: bit b of x goes to 3b+2, y to 3b+1, z to 3b."""
    out = 0
    for b in range(bits):
        out |= ((x >> b) & 1) << (3 * b + 2)
        out |= ((y >> b) & 1) << (3 * b + 1)
        out |= ((z >> b) & 1) << (3 * b)
    return out


def test_expand_bits10_exhaustive():
    for v in range(1024):
        assert oracle.lib().orc_expand_bits10(v) == _interleave(v, 0, 0, 10) >> 2


def test_expand_bits21_sampled():
    rng = np.random.default_rng(0)
    for v in list(range(64)) + [2097151, 1 << 20, (1 << 21) - 2] + rng.integers(0, 1 << 21, 300).tolist():
        assert oracle.lib().orc_expand_bits21(int(v)) == _interleave(int(v), 0, 0, 21) >> 2


@pytest.mark.parametrize("bits,scale", [(30, 1024), (63, 2097152)])
def test_morton_quantisation(bits, scale):
    """q = trunc(clamp(x*scale, 0, scale-1)); x most significant."""
    rng = np.random.default_rng(bits)
    pts = np.concatenate([rng.random((200, 3), dtype=np.float32),
                          np.array([[0, 0, 0], [1, 1, 1], [-0.5, 2.0, 0.999999],
                                    [np.nextafter(F32(1), F32(0))] * 3], np.float32)])
    f = oracle.lib().orc_morton30 if bits == 30 else oracle.lib().orc_morton63
    for p in pts:
        q = [int(min(max(F32(c) * F32(scale), F32(0)), F32(scale - 1))) for c in p]
        assert f(*[float(c) for c in p]) == _interleave(q[0], q[1], q[2], bits // 3)


def test_morton_degenerate_axis_is_zero():
    """A flat scene (bbmax == bbmin on an axis) gives 0/0 = NaN, which the
    clamp maps to quantised coordinate 0 (fmaxf ignores NaN)."""
    assert oracle.lib().orc_morton30(float("nan"), 0.5, 0.5) == _interleave(0, 512, 512, 10)


def test_encode_centroid_and_normalisation():
    v = ab.gen_soup(3, 500)
    mn, mx = oracle.scene_bbox(v)
    keys, idx = oracle.encode(v, mn, mx, 30)
    t = v.reshape(-1, 3, 4)[:, :, :3]
    c = (t.max(axis=1) + t.min(axis=1)) * F32(0.5)
    nrm = (c - mn) / (mx - mn)
    q = np.clip(nrm * F32(1024), 0, 1023).astype(np.uint32)
    ref = [_interleave(int(a), int(b), int(cc), 10) for a, b, cc in q]
    assert np.array_equal(keys, np.array(ref, np.uint32))
    assert np.array_equal(idx, np.arange(500, dtype=np.uint32))


def test_scene_bbox_quirk():
    """SceneConverter starts the running maximum at FLT_MIN (+1.18e-38): an
    all-negative scene keeps max = FLT_MIN (SceneConverter.cpp:32-34)."""
    v = ab.gen_soup(2, 10) - F32(5)
    mn, mx = oracle.scene_bbox(v)
    assert np.all(mx == np.finfo(np.float32).tiny)
    assert np.array_equal(mn, v.reshape(-1, 4)[:, :3].min(axis=0))


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("dt", [np.uint32, np.uint64])
def test_sort_is_stable_sort(dt):
    rng = np.random.default_rng(7)
    n = 20000
    hi = np.iinfo(dt).max
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(dt)
    keys[rng.integers(0, n, 5000)] = keys[3]
    vals = np.arange(n, dtype=np.uint32)
    k, vv = oracle.sort_pairs(keys, vals)
    o = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[o]) and np.array_equal(vv, vals[o])


# ------------------------------------------------------------------ Karras
def _radix_tree(keys, bits):
    """Top-down radix tree over (key, index): split at the highest differing
    bit; returns Karras-numbered parent/left/right arrays."""
    n = len(keys)
    aug = [(int(k) << 32) | i for i, k in enumerate(keys)]
    total = bits + 32
    left = np.full(2 * n - 1, -1, np.int32)
    right = np.full(2 * n - 1, -1, np.int32)
    parent = np.full(2 * n - 1, -1, np.int32)

    def build(a, b, idx):
        # split: highest bit where aug[a] and aug[b] differ; last index with that bit 0
        x = aug[a] ^ aug[b]
        hb = x.bit_length() - 1
        g = a
        while g + 1 <= b and not (aug[g + 1] >> hb) & 1:
            g += 1
        li = g + n - 1 if g == a else g
        ri = g + 1 + n - 1 if g + 1 == b else g + 1
        left[idx], right[idx] = li, ri
        parent[li] = parent[ri] = idx
        if g != a:
            build(a, g, li)
        if g + 1 != b:
            build(g + 1, b, ri)

    assert total <= 96
    build(0, n - 1, 0)
    return parent, left, right


@pytest.mark.parametrize("n,bits", [(2, 30), (3, 30), (17, 30), (600, 30), (600, 63), (3000, 63)])
def test_karras_matches_radix_tree(n, bits):
    rng = np.random.default_rng(n)
    if bits == 30:
        keys = np.sort(rng.integers(0, 1 << 12, n).astype(np.uint32))  # many duplicates
    else:
        keys = np.sort(rng.integers(0, 1 << 40, n).astype(np.uint64))
        keys[n // 3: n // 3 + 40] = keys[n // 3]
    idx = rng.permutation(n).astype(np.uint32)
    t = oracle.karras(keys, idx)
    p, l, r = _radix_tree(keys, bits)
    internal = np.arange(n - 1)
    assert np.array_equal(t.left[internal], l[internal])
    assert np.array_equal(t.right[internal], r[internal])
    assert np.array_equal(t.parent, p)
    assert np.array_equal(t.data[n - 1:], idx)
    assert np.all(t.data[: n - 1] == -1)


# ------------------------------------------------------------------ refit + SAH
def _box_area(mn, mx):
    s = (mx - mn).astype(np.float32)
    return F32(2) * (s[..., 0] * s[..., 1] + s[..., 0] * s[..., 2] + s[..., 1] * s[..., 2])


def test_refit_is_recursive_union():
    v = ab.gen_soup(8, 3000)
    mn, mx = oracle.scene_bbox(v)
    t = oracle.lbvh(v, mn, mx, 30)
    n = 3000
    tri = v.reshape(-1, 3, 4)[:, :, :3]
    lo = np.zeros((2 * n - 1, 3), np.float32)
    hi = np.zeros((2 * n - 1, 3), np.float32)

    def rec(x):
        if x >= n - 1:
            lo[x] = tri[t.data[x]].min(axis=0)
            hi[x] = tri[t.data[x]].max(axis=0)
        else:
            rec(t.left[x])
            rec(t.right[x])
            lo[x] = np.minimum(lo[t.left[x]], lo[t.right[x]])
            hi[x] = np.maximum(hi[t.left[x]], hi[t.right[x]])

    import sys
    sys.setrecursionlimit(100000)
    rec(0)
    assert np.array_equal(t.bbmin[:, :3], lo) and np.array_equal(t.bbmax[:, :3], hi)
    assert np.array_equal(t.area.view(np.uint32), _box_area(lo, hi).view(np.uint32))


def test_sah_summation_order():
    v = ab.gen_soup(9, 2000)
    mn, mx = oracle.scene_bbox(v)
    t = oracle.lbvh(v, mn, mx, 30)
    n = 2000
    total = t.area[0]
    si = F32(total)
    for i in range(1, n - 1):
        si = F32(si + t.area[i])
    sl = F32(0)
    for i in range(n - 1, 2 * n - 1):
        sl = F32(sl + t.area[i])
    want = F32(F32(F32(1.2) * si) + F32(F32(1.0) * sl)) / F32(total)
    assert np.float32(oracle.sah(t)) == want


# ------------------------------------------------------------------ schedules
def test_trbvh_schedule_matches_reference_fixture():
    with open(os.path.join(GOLDEN, "trbvh_schedule_ref.json")) as f:
        ref = json.load(f)
    for key, rounds in ref.items():
        n, w = (int(x) for x in key.split("_"))
        assert oracle.trbvh_schedule(n, w).tolist() == rounds, key
        assert ab.trbvh_schedule(n, w).tolist() == rounds, key


def test_trbvh_schedule_matches_live_reference_build():
    live = oracle.reference_trbvh_schedule(7, 32)
    if live is None:
        pytest.skip("oracle/_ref not built here")
    assert np.array_equal(live, oracle.trbvh_schedule(7, 32))


@pytest.mark.parametrize("n", [5, 6, 7])
def test_trbvh_schedule_dependency_order(n):
    s = oracle.trbvh_schedule(n, 32)
    seen = {}
    for r, row in enumerate(s):
        for x in row:
            if x:
                assert x not in seen
                seen[int(x)] = r
    want = {m for m in range(1, 1 << n) if 2 <= bin(m).count("1") <= n - 2}
    assert set(seen) == want
    for m, r in seen.items():
        sub = m
        while sub:
            sub = (sub - 1) & m
            if sub and bin(sub).count("1") >= 2:
                assert seen[sub] < r


@pytest.mark.parametrize("n", [3, 9, 12, 20, 32])
def test_atrbvh_schedule_covers_pairs_once(n):
    s = oracle.atrbvh_schedule(n, 32)
    pairs = []
    for w in s.astype(np.int64) & 0xFFFFFFFF:
        for half in ((w >> 16) & 0xFFFF, w & 0xFFFF):
            a, b = half >> 8, half & 0xFF
            if a or b:
                pairs.append((int(a), int(b)))
    assert sorted(pairs) == sorted((i, j) for i in range(n) for j in range(i))


def test_lower_trm_row_col_exhaustive():
    p = 0
    for r in range(1, 32):
        for c in range(r):
            assert oracle.lower_trm_row(p) == r and oracle.lower_trm_col(p) == c
            p += 1


# ------------------------------------------------------------------ TRBVH DP
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


def _mask(t):
    return (1 << t) if isinstance(t, int) else _mask(t[0]) | _mask(t[1])


def test_trbvh_dp_equals_exhaustive_topologies():
    """For 7 leaves the subset DP must equal the best of all 10,395 binary
    trees evaluated with the same fp32 cost formula (fminf of the split cost
    and the collapse cost, the latter using the treelet's triangle count)."""
    rng = np.random.default_rng(11)
    collapsed_somewhere = False
    # leaf-cost scales from "splitting always wins" to "collapsing wins often"
    for trial, scale in enumerate((1.0, 4.0, 12.0, 40.0)):
        lo = rng.random((7, 3), dtype=np.float32)
        hi = lo + rng.random((7, 3), dtype=np.float32) * F32(0.3)
        area = np.zeros(128, np.float32)
        for s in range(128):
            if s == 0:
                continue
            m = [i for i in range(7) if s >> i & 1]
            area[s] = _box_area(lo[m].min(axis=0), hi[m].max(axis=0))
        leaf = (area[[1 << i for i in range(7)]] * F32(scale)).astype(np.float32)
        T = 7 + trial * 3
        ci, ct = F32(1.2), F32(1.0)
        dp, costs, masks = oracle.trbvh_dp(leaf, area, T)
        for s_ in range(128):
            if bin(s_).count("1") >= 2 and costs[s_] == F32(F32(ct * area[s_]) * F32(T)):
                collapsed_somewhere = True

        def cost(t):
            if isinstance(t, int):
                return leaf[t]
            s = _mask(t)
            a = area[s]
            split = F32(F32(ci * a) + F32(cost(t[0]) + cost(t[1])))
            return min(split, F32(F32(ct * a) * F32(T)))

        trees = list(_all_trees(list(range(7))))
        assert len(trees) == 10395
        best = min(cost(t) for t in trees)
        assert np.float32(dp) == best
        # the chosen partition at the root realises the optimum
        p = int(masks[127]) & 0xFF
        assert 0 < p < 127 and not (p & 1)
    assert collapsed_somewhere, "the collapse branch of the DP was never exercised"


# ------------------------------------------------------------------ ATRBVH clustering
def test_atrbvh_cluster_merges_closest_pair():
    """Each step merges a pair of current clusters whose union area is the
    minimum over all current pairs (greedy agglomerative clustering)."""
    rng = np.random.default_rng(12)
    for n in (3, 5, 9, 16, 32):
        lo = rng.random((n, 3), dtype=np.float32)
        hi = lo + rng.random((n, 3), dtype=np.float32) * F32(0.2)
        boxes = np.concatenate([lo, hi], axis=1)
        sah = rng.random(n, dtype=np.float32)
        merges, areas = oracle.atrbvh_cluster(boxes, sah)
        slots = [[i] for i in range(n)]  # member leaves per slot
        ssah = list(sah)
        for step, (r, c) in enumerate(merges):
            L = n - 1 - step
            assert 0 <= c < r <= L
            cur = slots[: L + 1]

            def ua(a, b):
                m = a + b
                return _box_area(lo[m].min(axis=0), hi[m].max(axis=0))

            best = min(ua(cur[i], cur[j]) for i in range(L + 1) for j in range(i))
            assert ua(cur[r], cur[c]) == best
            merged = cur[r] + cur[c]
            s = F32(F32(ssah[r]) + F32(ssah[c]))
            want = F32(s + F32(F32(1.2) * ua(cur[r], cur[c])))
            assert areas[L] == want
            slots[c] = merged
            ssah[c] = want
            slots[r] = slots[L]
            ssah[r] = ssah[L]


# ------------------------------------------------------------------ whole optimisers
def _check_invariants(t, n):
    from helpers import check_tree_invariants, tree_arrays
    check_tree_invariants(tree_arrays(t), n)


@pytest.mark.parametrize("kind", ["atrbvh", "trbvh"])
def test_optimisers_keep_invariants_and_lower_sah(kind):
    v = ab.gen_soup(13, 20000)
    mn, mx = oracle.scene_bbox(v)
    t = oracle.lbvh(v, mn, mx, 30)
    s0 = oracle.sah(t)
    if kind == "atrbvh":
        oracle.atrbvh(t, 9, 2)
    else:
        oracle.trbvh(t, 7, 3)
    _check_invariants(t, 20000)
    assert oracle.sah(t) < s0
    # refit of the optimised topology reproduces its boxes and areas exactly
    t2 = t.copy()
    oracle.refit(v, t2)
    assert np.array_equal(t2.bbmin[:, :3], t.bbmin[:, :3])
    assert np.array_equal(t2.area.view(np.uint32), t.area.view(np.uint32))


def test_c1_regression_fixture():
    with open(os.path.join(GOLDEN, "c1_summary.json")) as f:
        g = json.load(f)
    import hashlib

    def dig(a):
        return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()

    v = ab.gen_soup(1, 65536)
    mn, mx = oracle.scene_bbox(v)
    assert [np.float32(x).tobytes().hex() for x in mn] == g["bbmin"]
    lb = oracle.lbvh(v, mn, mx, 30)
    at = lb.copy()
    oracle.atrbvh(at, 9, 2)
    tr = lb.copy()
    oracle.trbvh(tr, 7, 3)
    for name, t in (("lbvh", lb), ("atrbvh", at), ("trbvh", tr)):
        for k, d in g[name].items():
            assert dig(getattr(t, k)) == d, f"{name}.{k}"
    c, csah = oracle.collapse(at)
    assert np.float32(csah).tobytes().hex() == g["sah_collapsed"]
    assert c.used == g["collapsed_used"]


# ------------------------------------------------------------------ collapse
@pytest.mark.parametrize("sigma", [0.005, 0.2])
def test_collapse_recursive_semantics(sigma):
    """sigma 0.2 gives large overlapping triangles, where most subtrees
    collapse, both children of a parent are emitted and lists are long."""
    v = ab.gen_soup(14, 4000, sigma)
    mn, mx = oracle.scene_bbox(v)
    t = oracle.lbvh(v, mn, mx, 30)
    oracle.atrbvh(t, 9, 2)
    c, sah = oracle.collapse(t)
    n = 4000
    ci, ct = F32(1.2), F32(1.0)
    cost = {}
    cnt = {}
    flag = {}
    deliver = {}
    import sys
    sys.setrecursionlimit(100000)

    def rec(x):
        a = _box_area(t.bbmin[x, :3], t.bbmax[x, :3])
        if x >= n - 1:
            cost[x], cnt[x], flag[x], deliver[x] = F32(ct * a), 1, True, True
            return
        l, r = t.left[x], t.right[x]
        rec(l)
        rec(r)
        cnt[x] = cnt[l] + cnt[r]
        cc = F32(F32(F32(ci * a) + cost[l]) + cost[r])
        leafc = F32(F32(ct * a) * F32(cnt[x]))
        flag[x] = bool(leafc < cc)
        cost[x] = leafc if flag[x] else cc
        deliver[x] = deliver[l] and deliver[r] and flag[x]

    rec(0)
    assert np.float32(sah) == F32(cost[0] / t.area[0])
    emitted = set()
    expected_offset = 0
    for p in range(n - 1):
        l, r = t.left[p], t.right[p]
        if deliver[l] and deliver[r] and not flag[p]:
            for ch in (l, r):
                if cnt[ch] > 1:
                    emitted.add(int(ch))
                    # lists are laid out by emitting parent, left before right
                    assert c.data[ch] == expected_offset and c.count[ch] == cnt[ch]
                    expected_offset += cnt[ch]
                    got = c.tri_index[c.data[ch]: c.data[ch] + cnt[ch]].tolist()

                    def rl(x):  # right-first leaf order
                        if x >= n - 1:
                            return [int(t.data[x])]
                        return rl(t.right[x]) + rl(t.left[x])

                    assert got == rl(ch)
    for x in range(n - 1):
        if x not in emitted:
            assert c.data[x] == -1
    assert c.used == sum(cnt[x] for x in emitted)
    if sigma > 0.1:
        both = sum(1 for p in range(n - 1)
                   if int(t.left[p]) in emitted and int(t.right[p]) in emitted)
        assert both > 0 and max(cnt[x] for x in emitted) >= 4


# ------------------------------------------------------------------ traversal
def test_ray_triangle_closed_forms():
    v0, v1, v2 = (np.array(p, np.float32) for p in ([0, 0, 1], [1, 0, 1], [0, 1, 1]))
    o = np.zeros(3, np.float32)
    d = np.array([0.25, 0.25, 1.0], np.float32)
    hit, t, u, v = oracle.ray_triangle(v0, v1, v2, o, d)
    assert hit and t == pytest.approx(1.0) and u == pytest.approx(0.25) and v == pytest.approx(0.25)
    # two-sided: the same triangle seen from behind
    hit, t, _, _ = oracle.ray_triangle(v0, v2, v1, o, d)
    assert hit and t == pytest.approx(1.0)
    # parallel ray (det = 0) misses; outside barycentric range misses; tmax clips
    assert not oracle.ray_triangle(v0, v1, v2, o, np.array([1, 0, 0], np.float32))[0]
    assert not oracle.ray_triangle(v0, v1, v2, o, np.array([0.8, 0.8, 1], np.float32))[0]
    assert not oracle.ray_triangle(v0, v1, v2, o, d, 0.0, 0.5)[0]


def test_trace_equals_brute_force():
    v = ab.gen_soup(15, 3000, 0.02)
    mn, mx = oracle.scene_bbox(v)
    t = oracle.lbvh(v, mn, mx, 30)
    oracle.atrbvh(t, 9, 2)
    c, _ = oracle.collapse(t)
    rays = camera_rays_host(48)
    rng = np.random.default_rng(3)
    extra = np.zeros((500, 8), np.float32)
    extra[:, 0:3] = rng.random((500, 3))
    d = rng.normal(size=(500, 3))
    extra[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    extra[:, 7] = np.inf
    rays = np.concatenate([rays, extra])
    compare_hits(oracle.trace(c, v, rays), oracle.trace_brute(v, rays))


# ------------------------------------------------------------------ tie order of the emulated reductions
def _bitrev(x: int, bits: int) -> int:
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def test_find_minimum_tie_order_closed_form():
    """findMinimum (AgglomerativeTreelet.cu:65-99): lane l keeps the first
    strict minimum of entries l, l+32, ...; the 32-lane halving tree keeps its
    own value on ties, so among lanes holding the minimum the one with the
    smallest 5-bit bit-reversed index wins.  Predicted from that rule alone
    for planted ties."""
    rng = np.random.default_rng(21)
    for _ in range(3000):
        L = int(rng.integers(1, 32))
        size = L * (L + 1) // 2
        D = (1.0 + rng.random(size)).astype(np.float32)
        k = int(rng.integers(1, min(size, 8) + 1))
        S = sorted(rng.choice(size, k, replace=False).tolist())
        D[S] = np.float32(0.5)
        first_per_lane = {}
        for p in S:
            first_per_lane.setdefault(p % 32, p)
        lane = min(first_per_lane, key=lambda l: _bitrev(l, 5))
        assert oracle.find_minimum(D, L) == first_per_lane[lane], (L, S)


def test_find_leaf_to_expand_tie_order():
    """findLeafToExpand (Treelet.cuh:23-40) over m candidates, strict '>':
    a unique maximum always wins; for power-of-two m the halving tree gives
    ties to the smallest log2(m)-bit bit-reversed index; for any m lane 0 wins
    every tie it takes part in and a tie is always won by a tied lane."""
    rng = np.random.default_rng(22)
    for m in range(1, 33):
        for _ in range(300):
            a = rng.random(m).astype(np.float32)
            assert oracle.find_leaf_to_expand(a) == int(np.argmax(a))
            k = int(rng.integers(2, m + 1)) if m > 1 else 1
            S = sorted(rng.choice(m, k, replace=False).tolist())
            a[S] = np.float32(2.0)
            w = oracle.find_leaf_to_expand(a)
            assert w in S
            if 0 in S:
                assert w == 0
            if m & (m - 1) == 0:
                b = m.bit_length() - 1
                assert w == min(S, key=lambda i: _bitrev(i, b))


def _submasks_ascending(mask: int):
    return [p for p in range(1, mask + 1) if p & mask == p]


@pytest.mark.parametrize("pattern", ["popcount", "random_int"])
def test_trbvh_dp_tie_order_closed_form(pattern):
    """TRBVH DP partition choice under ties (TRBVH.cu:100-220).  For every
    subset S the candidates are the submasks P of S without S's lowest element
    (cost[P] + cost[S^P]).  Sizes 2..n-2: sequential scan in ascending P keeps
    the first strict minimum.  Size n-1: base b (the b-th such P in ascending
    order) is scanned by lane (b-1) mod 4, ties between lanes go to the
    2-bit bit-reversed order 0, 2, 1, 3.  Size n: P = 2(l+1) + 64j is scanned
    by lane l, ties go to the smallest 5-bit bit-reversed lane.  Inputs are
    small integers so that many candidates tie exactly."""
    rng = np.random.default_rng(23 if pattern == "popcount" else 24)
    n = 7
    full = (1 << n) - 1
    for trial in range(40):
        leaf = np.ones(n, np.float32) if pattern == "popcount" else \
            rng.integers(1, 4, n).astype(np.float32)
        if pattern == "popcount":
            area = np.array([bin(s).count("1") for s in range(1 << n)], np.float32)
        else:
            area = rng.integers(1, 6, 1 << n).astype(np.float32)
        T = int(rng.integers(7, 40))
        _, costs, masks = oracle.trbvh_dp(leaf, area, T, ci=1.0 + trial % 3, ct=1.0)
        ties_seen = 0
        for s in range(1, 1 << n):
            size = bin(s).count("1")
            if size < 2:
                continue
            low = s & -s
            cands = _submasks_ascending(s & ~low)
            vals = [np.float32(costs[p] + costs[s ^ p]) for p in cands]
            best = min(vals)
            tied = [p for p, v in zip(cands, vals) if v == best]
            ties_seen += len(tied) > 1
            if size <= n - 2:
                want = tied[0]
            elif size == n - 1:
                by_lane = {}
                for b, p in enumerate(cands, start=1):
                    if p in tied:
                        by_lane.setdefault((b - 1) % 4, p)
                lane = min(by_lane, key=lambda r: _bitrev(r, 2))
                want = by_lane[lane]
            else:
                by_lane = {}
                for p in cands:  # all even p, ascending
                    if p in tied:
                        by_lane.setdefault((p // 2 - 1) % 32, p)
                lane = min(by_lane, key=lambda r: _bitrev(r, 5))
                want = by_lane[lane]
            assert (int(masks[s]) & 0xFF) == want, (trial, s, size, tied)
        assert ties_seen > 0
    assert full == 127


# ------------------------------------------------------------------ CPU SBVH baseline
def test_sbvh_trace_equals_brute_force_and_threads_agree():
    """The CPU split-BVH (reference CPU path) finds the brute-force closest
    hits; its multi-threaded build gives the single-threaded tree."""
    from helpers import camera_rays_host, compare_hits

    v = ab.gen_soup(31, 6000, sigma=0.03)
    one = oracle.Sbvh(v, threads=1)
    many = oracle.Sbvh(v, threads=8)
    assert one.stats == many.stats
    assert one.stats["tri_refs"] >= 6000 and one.stats["depth"] < 64
    assert one.stats["sah_bvhrt"] > 0
    rays = camera_rays_host(48)
    ref = oracle.trace_brute(v, rays)
    got1 = one.trace(rays, threads=1)
    got8 = many.trace(rays, threads=8)
    assert np.array_equal(got1, got8)
    compare_hits(got1, ref)
    assert (ref[:, 0] >= 0).mean() > 0.15


def test_rays_diffuse_restatement_invariants():
    """orc_rays_diffuse against what the mathematics fixes: sample 1 of the
    Halton (2, 3) sequence is (1/2, 1/3), so the cosine warp puts every ray at
    r = sqrt(1/2) from the pole: its component along the (ray-facing) normal
    is sqrt(1/2) and the in-plane part has length sqrt(1/2) at polar angle
    2*pi/3 relative to the hashed frame; directions are unit vectors; the
    origin is backed off 1e-4 along the primary ray; misses give tmax = -1.
    A dropped normal flip, a swapped Halton base or a missing back-off fails."""
    rng = np.random.default_rng(3)
    n = 64
    v = np.zeros((3 * n, 4), np.float32)
    v[:, :3] = rng.uniform(-1, 1, (3 * n, 3)).astype(np.float32)
    r = 4096
    prim = np.zeros((r, 8), np.float32)
    prim[:, 0:3] = rng.uniform(-2, 2, (r, 3))
    d = rng.normal(size=(r, 3))
    prim[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    prim[:, 7] = np.inf
    res = np.zeros((r, 4), np.int32)
    res[:, 0] = rng.integers(-1, n, r)
    tt = rng.uniform(0.0, 3.0, r).astype(np.float32)
    tt[:5] = [0.0, 5e-5, 1e-4, 2e-4, 1.0]
    res[:, 1] = tt.view(np.int32)
    out = oracle.rays_diffuse(v, prim, res, seed=77)
    hit = res[:, 0] >= 0
    assert np.all(out[~hit, 7] == -1.0) and np.all(out[hit, 7] == np.inf)
    back = np.maximum(tt.astype(np.float64) - 1e-4, 0.0)
    assert np.allclose(out[:, 0:3], prim[:, 0:3] + prim[:, 4:7] * back[:, None], atol=1e-6)
    dirs = out[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-6)
    tri = v[:, :3].reshape(n, 3, 3).astype(np.float64)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nn = np.where(hit[:, None], nrm[np.maximum(res[:, 0], 0)], [1.0, 0.0, 0.0])
    flip = np.sum(nn * prim[:, 4:7], axis=1) > 0
    nn[flip] *= -1
    cos_pole = np.sum(dirs * nn, axis=1)
    assert np.allclose(cos_pole, np.sqrt(0.5), atol=2e-6)
    # the hashed rotation angle makes the in-plane direction vary from ray to ray
    inplane = dirs - cos_pole[:, None] * nn
    assert np.std(inplane[:, 0]) > 0.2
