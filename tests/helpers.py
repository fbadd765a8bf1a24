"""Shared test helpers: workloads and exact comparisons."""
from __future__ import annotations

import numpy as np

# Config 1 of the workload set: 65,536-triangle soup, sigma 0.005, 30-bit keys,
# ATRBVH treelet 9 / 2 iterations, 256x256 pinhole camera at (0.5, 0.5, -1)
# looking +z with a 90 degree field of view.
C1_N = 65536
C1_SEED = 1
CAMERA_EYE = (0.5, 0.5, -1.0)
CAMERA_BASIS = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)

TREE_INT = ("parent", "left", "right", "data")


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def tree_arrays(t) -> dict:
    """Oracle Tree -> dict of arrays (same keys as atrbvh_amd.Tree.download())."""
    return {"parent": t.parent, "left": t.left, "right": t.right, "data": t.data,
            "bbmin": t.bbmin, "bbmax": t.bbmax, "area": t.area}


def assert_tree_equal(dev: dict, ref: dict, what: str = "tree"):
    for k in TREE_INT:
        d = np.asarray(dev[k])
        r = np.asarray(ref[k])
        if not np.array_equal(d, r):
            bad = np.nonzero(d != r)[0]
            raise AssertionError(f"{what}.{k}: {bad.size} mismatches, first at {bad[:8]}: "
                                 f"dev={d[bad[:8]]} ref={r[bad[:8]]}")
    for k in ("bbmin", "bbmax"):
        d = bits(np.asarray(dev[k])[:, :3])
        r = bits(np.asarray(ref[k])[:, :3])
        if not np.array_equal(d, r):
            bad = np.nonzero((d != r).any(axis=1))[0]
            raise AssertionError(f"{what}.{k}: {bad.size} boxes differ, first at {bad[:8]}")
    d = bits(dev["area"])
    r = bits(ref["area"])
    if not np.array_equal(d, r):
        bad = np.nonzero(d != r)[0]
        raise AssertionError(f"{what}.area: {bad.size} mismatches, first at {bad[:8]}")


def check_tree_invariants(t: dict, n: int):
    """Structural invariants of a BVHRT SoA tree (topology + conservative boxes)."""
    parent, left, right, data = (np.asarray(t[k]) for k in TREE_INT)
    m = 2 * n - 1
    assert parent[0] == -1
    internal = np.arange(n - 1)
    leaves = np.arange(n - 1, m)
    assert np.all(left[leaves] == -1) and np.all(right[leaves] == -1)
    assert np.all(data[internal] == -1)
    assert np.array_equal(np.sort(data[leaves]), np.arange(n))
    kids = np.concatenate([left[internal], right[internal]])
    assert np.array_equal(np.sort(kids), np.arange(1, m)), "every node except the root is a child once"
    assert np.all(parent[left[internal]] == internal)
    assert np.all(parent[right[internal]] == internal)
    bmin = np.asarray(t["bbmin"])[:, :3]
    bmax = np.asarray(t["bbmax"])[:, :3]
    for side in (left, right):
        c = side[internal]
        assert np.all(bmin[internal] <= bmin[c]) and np.all(bmax[internal] >= bmax[c])


def collapsed_leaf_sets(c: dict, n: int):
    """Map every reachable leaf entity of a collapsed tree to its triangle set."""
    left, right, data, count, tri = (np.asarray(c[k]) for k in
                                     ("left", "right", "data", "count", "tri_index"))
    out = {}
    stack = [0]
    while stack:
        node = stack.pop()
        for ch in (left[node], right[node]):
            if data[ch] >= 0:
                if ch < n - 1:  # a collapsed subtree's list
                    out[int(ch)] = tuple(int(x) for x in tri[data[ch]: data[ch] + count[ch]])
                else:
                    out[int(ch)] = (int(data[ch]),)
            else:
                stack.append(int(ch))
    return out


def camera_rays_host(g: int, eye=CAMERA_EYE) -> np.ndarray:
    """Host copy of the pinhole rays (row-major, 90 degree FOV, tmax = inf)."""
    i, j = np.meshgrid(np.arange(g), np.arange(g), indexing="xy")
    d = np.stack([(2.0 * (i + 0.5) / g - 1.0), (2.0 * (j + 0.5) / g - 1.0),
                  np.ones_like(i, dtype=np.float64)], -1).reshape(-1, 3)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays = np.zeros((g * g, 8), np.float32)
    rays[:, 0:3] = eye
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    return rays


def compare_hits(dev: np.ndarray, ref: np.ndarray, rtol: float = 1e-6, atol: float = 0.0):
    """Identical hit ids except exact-distance ties; t within rtol (plus atol,
    for comparisons of two different hit tests whose t carry absolute
    rounding of the scene's coordinates)."""
    dev = np.asarray(dev).reshape(-1, 4)
    ref = np.asarray(ref).reshape(-1, 4)
    td = dev[:, 1].copy().view(np.float32)
    tr = ref[:, 1].copy().view(np.float32)
    hit_r = ref[:, 0] >= 0
    assert np.array_equal(dev[:, 0] >= 0, hit_r), "hit/miss pattern differs"
    same = dev[:, 0] == ref[:, 0]
    tie = ~same & np.isclose(td, tr, rtol=rtol, atol=atol)
    assert np.all(same | tie), f"{np.count_nonzero(~(same | tie))} rays hit different triangles"
    h = hit_r
    assert np.allclose(td[h], tr[h], rtol=rtol, atol=atol)
    return int(np.count_nonzero(~same))


def _edge_distance(v: np.ndarray, tri: int, ray: np.ndarray) -> float:
    """World-space distance, in float64, from the point where the ray's line
    crosses triangle tri's plane to the nearest edge of the triangle,
    relative to 1 + the distance of that point from the origin."""
    p = v.reshape(-1, 3, 4)[tri, :, :3].astype(np.float64)
    o = ray[0:3].astype(np.float64)
    d = ray[4:7].astype(np.float64)
    nrm = np.cross(p[1] - p[0], p[2] - p[0])
    den = d @ nrm
    if den == 0.0:
        return 0.0
    x = o + ((p[0] - o) @ nrm) / den * d
    dist = np.inf
    for k in range(3):
        a, b = p[k], p[(k + 1) % 3]
        e = b - a
        s = np.clip((x - a) @ e / (e @ e), 0.0, 1.0)
        dist = min(dist, float(np.linalg.norm(x - (a + s * e))))
    return dist / (1.0 + float(np.linalg.norm(x)))


def compare_hits_two_tests(a: np.ndarray, b: np.ndarray, v: np.ndarray, rays: np.ndarray,
                           rtol: float = 1e-5, atol: float = 1e-6, edge_eps: float = 1e-5,
                           max_frac: float = 1e-3) -> int:
    """Hits of two different ray/triangle tests on the same tree (the Woop
    unit-triangle test and Moller-Trumbore) agree except where a ray grazes a
    triangle edge: identical ids or exact-distance ties elsewhere, t within
    rtol + atol, and on every ray where one test reports a nearer hit than the
    other, the ray crosses that nearer triangle's plane within edge_eps
    (relative, float64) of one of its edges.  Such rays are at most max_frac
    of all rays.
    Returns their number."""
    a = np.asarray(a).reshape(-1, 4)
    b = np.asarray(b).reshape(-1, 4)
    ta = np.where(a[:, 0] >= 0, a[:, 1].copy().view(np.float32), np.float32(np.inf))
    tb = np.where(b[:, 0] >= 0, b[:, 1].copy().view(np.float32), np.float32(np.inf))
    same = a[:, 0] == b[:, 0]
    both = (a[:, 0] >= 0) & (b[:, 0] >= 0)
    close = both & np.isclose(ta, tb, rtol=rtol, atol=atol)
    assert np.allclose(ta[same & both], tb[same & both], rtol=rtol, atol=atol)
    odd = np.nonzero(~same & ~close)[0]
    for r in odd:
        tri = int(a[r, 0]) if ta[r] < tb[r] else int(b[r, 0])
        m = _edge_distance(v, tri, rays[r])
        assert m <= edge_eps, f"ray {r}: tests disagree away from an edge (distance {m:.3g})"
    assert len(odd) <= max(3, int(max_frac * len(a))), f"{len(odd)} edge disagreements"
    return int(len(odd))


def _clip(poly: np.ndarray, axis: int, c: float, keep_below: bool) -> np.ndarray:
    """Sutherland-Hodgman clip of a convex polygon (k x 3, float64) to one side
    of the plane x[axis] = c."""
    out = []
    k = len(poly)
    for i in range(k):
        a, b = poly[i], poly[(i + 1) % k]
        ina = a[axis] <= c if keep_below else a[axis] >= c
        inb = b[axis] <= c if keep_below else b[axis] >= c
        if ina:
            out.append(a)
        if ina != inb:
            s = (c - a[axis]) / (b[axis] - a[axis])
            p = a + s * (b - a)
            p[axis] = c
            out.append(p)
    return np.array(out)


def split_fragments(v: np.ndarray, every: int = 3, levels: int = 1):
    """Triangle fragments for the triangle-splitting build: triangle i with
    i % every == 0 is cut `levels` times, each piece at the midpoint of its
    box's longest axis; a fragment's box is the box of its exactly clipped
    polygon (float64), rounded outwards to float32 so the fragments' boxes
    cover the triangle.  Other triangles keep one box.  Returns (bbmin,
    bbmax) as float4 rows (w = 0) and the triangle of each fragment."""
    tris = np.asarray(v, np.float32).reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    mins, maxs, ids = [], [], []
    for i, t in enumerate(tris):
        pieces = [t]
        if i % every == 0:
            for _ in range(levels):
                nxt = []
                for p in pieces:
                    lo, hi = p.min(0), p.max(0)
                    ax = int(np.argmax(hi - lo))
                    c = 0.5 * (lo[ax] + hi[ax])
                    for below in (True, False):
                        q = _clip(p, ax, c, below)
                        if len(q) >= 3:
                            nxt.append(q)
                pieces = nxt
        for p in pieces:
            lo = p.min(0).astype(np.float32)
            hi = p.max(0).astype(np.float32)
            lo = np.where(lo.astype(np.float64) > p.min(0), np.nextafter(lo, np.float32(-np.inf)), lo)
            hi = np.where(hi.astype(np.float64) < p.max(0), np.nextafter(hi, np.float32(np.inf)), hi)
            mins.append(np.r_[lo, 0.0])
            maxs.append(np.r_[hi, 0.0])
            ids.append(i)
    return (np.array(mins, np.float32), np.array(maxs, np.float32), np.array(ids, np.int32))


def triangle_boxes(v: np.ndarray):
    """Each triangle's own box as float4 rows (w = 0) and the identity ids."""
    t = np.asarray(v, np.float32).reshape(-1, 3, 4)[:, :, :3]
    n = t.shape[0]
    mn = np.zeros((n, 4), np.float32)
    mx = np.zeros((n, 4), np.float32)
    mn[:, :3] = t.min(1)
    mx[:, :3] = t.max(1)
    return mn, mx, np.arange(n, dtype=np.int32)


COMPAT_SRC = "tests/compat_main.cpp"
COMPAT_BIN = "tests/_bin/compat_main"


def build_compat_driver(root: str) -> str:
    """Compile tests/compat_main.cpp (a host program using the C++ facade
    include/bvhrt_compat.hpp) against libatrbvh.so with g++."""
    import os
    import subprocess

    out = os.path.join(root, COMPAT_BIN)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    libdir = os.path.join(root, "atrbvh_amd")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
           "-I" + os.path.join(root, "include"), os.path.join(root, COMPAT_SRC),
           "-L" + libdir, "-latrbvh", "-Wl,-rpath,$ORIGIN/../../atrbvh_amd", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def run_compat_driver(root: str, v: np.ndarray, morton64: bool, opt: str, tmpdir: str):
    """Runs the facade driver on vertices v; returns (returncode, stdout,
    parsed results or None)."""
    import os
    import subprocess

    exe = os.path.join(root, COMPAT_BIN)
    n = v.size // 12
    vin = os.path.join(tmpdir, "v.f32")
    out = os.path.join(tmpdir, "out.bin")
    np.ascontiguousarray(v, np.float32).tofile(vin)
    r = subprocess.run([exe, vin, str(n), "1" if morton64 else "0", opt, out],
                       capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        return r.returncode, r.stdout, None
    raw = np.fromfile(out, np.uint8)
    head = raw[:12]
    sah_tree, sah_c = head[:8].view(np.float32)
    used = int(head[8:12].view(np.uint32)[0])
    m = 2 * n - 1
    ints = raw[12:].view(np.int32)
    res = {"sah_tree": sah_tree, "sah_collapsed": sah_c, "used": used}
    for i, k in enumerate(("parent", "left", "right", "data", "c_data", "c_count")):
        res[k] = ints[i * m:(i + 1) * m]
    res["tri_index"] = ints[6 * m:6 * m + used]
    return 0, r.stdout, res
