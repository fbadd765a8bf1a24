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
                if count[ch] > 1:
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
