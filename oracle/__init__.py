"""ctypes binding of the serial CPU oracle (test infrastructure only).

The oracle restates the reference BVHRT-Core build path and the closest-hit
traversal on the CPU.  It is the checker: only tests/, the smoke check and the
CPU-baseline leg of bench.py load it.  The product path (atrbvh_amd) never
imports this package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ATRBVH_ORACLE_LIB selects another build of the same sources (the mutation
# tests load builds with a planted mistake through it).
_LIB_PATH = os.environ.get("ATRBVH_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_REF_SCHED = os.path.join(_HERE, "_ref", "trbvh_sched_dump")

_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)


class OrcTree(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32),
        ("root", ctypes.c_int32),
        ("parent", _i32p),
        ("left", _i32p),
        ("right", _i32p),
        ("data", _i32p),
        ("bbmin", _f32p),
        ("bbmax", _f32p),
        ("area", _f32p),
    ]


class OrcCollapsed(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32),
        ("root", ctypes.c_int32),
        ("parent", _i32p),
        ("left", _i32p),
        ("right", _i32p),
        ("data", _i32p),
        ("count", _i32p),
        ("bbmin", _f32p),
        ("bbmax", _f32p),
        ("tri_index", _i32p),
        ("used", ctypes.c_uint32),
    ]


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is mounted)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/BVHRT-Core/src"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run oracle.build())")
        _lib = ctypes.CDLL(_LIB_PATH)
        L = _lib
        L.orc_sah.restype = ctypes.c_float
        L.orc_trbvh_dp.restype = ctypes.c_float
        L.orc_expand_bits10.restype = ctypes.c_uint32
        L.orc_expand_bits10.argtypes = [ctypes.c_uint32]
        L.orc_expand_bits21.restype = ctypes.c_uint64
        L.orc_expand_bits21.argtypes = [ctypes.c_uint64]
        L.orc_morton30.restype = ctypes.c_uint32
        L.orc_morton30.argtypes = [ctypes.c_float] * 3
        L.orc_morton63.restype = ctypes.c_uint64
        L.orc_morton63.argtypes = [ctypes.c_float] * 3
        L.orc_sah.argtypes = [ctypes.POINTER(OrcTree), ctypes.c_float, ctypes.c_float]
        L.orc_trbvh_optimize.argtypes = [ctypes.POINTER(OrcTree), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_float, ctypes.c_float]
        L.orc_atrbvh_optimize.argtypes = L.orc_trbvh_optimize.argtypes
        L.orc_collapse.argtypes = [ctypes.POINTER(OrcTree), ctypes.c_float, ctypes.c_float,
                                   ctypes.POINTER(OrcCollapsed), _f32p]
        L.orc_lower_trm_row.argtypes = [ctypes.c_int]
        L.orc_lower_trm_col.argtypes = [ctypes.c_int]
        L.orc_trbvh_dp.argtypes = [ctypes.c_int, _f32p, _f32p, ctypes.c_int, ctypes.c_float,
                                   ctypes.c_float, _f32p, _i32p]
        L.orc_atrbvh_cluster.argtypes = [ctypes.c_int, _f32p, _f32p, ctypes.c_float, _i32p,
                                         _f32p]
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------- scene / Morton
def scene_bbox(v4: np.ndarray):
    v4 = _f32(v4)
    n = v4.shape[0] // 3
    mn = np.zeros(3, np.float32)
    mx = np.zeros(3, np.float32)
    lib().orc_scene_bbox(_p(v4, _f32p), ctypes.c_uint32(n), _p(mn, _f32p), _p(mx, _f32p))
    return mn, mx


def encode(v4, bbmin, bbmax, bits):
    v4 = _f32(v4)
    n = v4.shape[0] // 3
    bbmin = _f32(bbmin)
    bbmax = _f32(bbmax)
    idx = np.zeros(n, np.uint32)
    if bits == 30:
        k = np.zeros(n, np.uint32)
        lib().orc_encode30(_p(v4, _f32p), ctypes.c_uint32(n), _p(bbmin, _f32p), _p(bbmax, _f32p),
                           _p(k, _u32p), _p(idx, _u32p))
    else:
        k = np.zeros(n, np.uint64)
        lib().orc_encode63(_p(v4, _f32p), ctypes.c_uint32(n), _p(bbmin, _f32p), _p(bbmax, _f32p),
                           _p(k, _u64p), _p(idx, _u32p))
    return k, idx


def sort_pairs(keys, vals):
    keys = keys.copy()
    vals = np.ascontiguousarray(vals, dtype=np.uint32).copy()
    n = keys.shape[0]
    if keys.dtype == np.uint32:
        lib().orc_sort_pairs32(_p(keys, _u32p), _p(vals, _u32p), ctypes.c_uint32(n))
    else:
        lib().orc_sort_pairs64(_p(keys, _u64p), _p(vals, _u32p), ctypes.c_uint32(n))
    return keys, vals


# ---------------------------------------------------------------- trees
class Tree:
    """Host SoA tree in the BVHRT convention (owns its numpy arrays)."""

    def __init__(self, n: int):
        m = 2 * n - 1
        self.n = n
        self.parent = np.full(m, -7, np.int32)
        self.left = np.full(m, -7, np.int32)
        self.right = np.full(m, -7, np.int32)
        self.data = np.full(m, -7, np.int32)
        self.bbmin = np.zeros((m, 4), np.float32)
        self.bbmax = np.zeros((m, 4), np.float32)
        self.area = np.zeros(m, np.float32)
        self.root = 0

    def struct(self) -> OrcTree:
        s = OrcTree()
        s.n = self.n
        s.root = self.root
        s.parent = _p(self.parent, _i32p)
        s.left = _p(self.left, _i32p)
        s.right = _p(self.right, _i32p)
        s.data = _p(self.data, _i32p)
        s.bbmin = _p(self.bbmin, _f32p)
        s.bbmax = _p(self.bbmax, _f32p)
        s.area = _p(self.area, _f32p)
        return s

    def pull(self, s: OrcTree):
        self.root = s.root

    def copy(self) -> "Tree":
        t = Tree(self.n)
        for f in ("parent", "left", "right", "data", "bbmin", "bbmax", "area"):
            getattr(t, f)[...] = getattr(self, f)
        t.root = self.root
        return t


class Collapsed:
    def __init__(self, n: int):
        m = 2 * n - 1
        self.n = n
        self.parent = np.zeros(m, np.int32)
        self.left = np.zeros(m, np.int32)
        self.right = np.zeros(m, np.int32)
        self.data = np.zeros(m, np.int32)
        self.count = np.zeros(m, np.int32)
        self.bbmin = np.zeros((m, 4), np.float32)
        self.bbmax = np.zeros((m, 4), np.float32)
        self.tri_index = np.full(n, -1, np.int32)
        self.used = 0
        self.root = 0

    def struct(self) -> OrcCollapsed:
        s = OrcCollapsed()
        s.n = self.n
        s.root = self.root
        for f in ("parent", "left", "right", "data", "count", "tri_index"):
            setattr(s, f, _p(getattr(self, f), _i32p))
        s.bbmin = _p(self.bbmin, _f32p)
        s.bbmax = _p(self.bbmax, _f32p)
        s.used = self.used
        return s


def karras(keys, idx) -> Tree:
    n = keys.shape[0]
    t = Tree(n)
    s = t.struct()
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    if keys.dtype == np.uint32:
        lib().orc_karras32(_p(keys, _u32p), _p(idx, _u32p), ctypes.c_uint32(n), ctypes.byref(s))
    else:
        lib().orc_karras64(_p(keys, _u64p), _p(idx, _u32p), ctypes.c_uint32(n), ctypes.byref(s))
    t.pull(s)
    return t


def refit(v4, t: Tree) -> Tree:
    v4 = _f32(v4)
    s = t.struct()
    lib().orc_refit(_p(v4, _f32p), ctypes.byref(s))
    return t


def lbvh(v4, bbmin, bbmax, bits) -> Tree:
    v4 = _f32(v4)
    n = v4.shape[0] // 3
    t = Tree(n)
    s = t.struct()
    bbmin = _f32(bbmin)
    bbmax = _f32(bbmax)
    rc = lib().orc_lbvh(_p(v4, _f32p), ctypes.c_uint32(n), _p(bbmin, _f32p), _p(bbmax, _f32p),
                        ctypes.c_int(bits), ctypes.byref(s))
    if rc != 0:
        raise ValueError("orc_lbvh failed")
    t.pull(s)
    return t


def lbvh_boxes(bmin4, bmax4, tri, bbmin, bbmax, bits) -> Tree:
    """LBVH over m primitive boxes (float4 rows) with the triangle of each."""
    bmin4 = _f32(bmin4).reshape(-1, 4)
    bmax4 = _f32(bmax4).reshape(-1, 4)
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    m = bmin4.shape[0]
    t = Tree(m)
    s = t.struct()
    bbmin = _f32(bbmin)
    bbmax = _f32(bbmax)
    rc = lib().orc_lbvh_boxes(_p(bmin4, _f32p), _p(bmax4, _f32p), _p(tri, _i32p),
                              ctypes.c_uint32(m), _p(bbmin, _f32p), _p(bbmax, _f32p),
                              ctypes.c_int(bits), ctypes.byref(s))
    if rc != 0:
        raise ValueError("orc_lbvh_boxes failed")
    t.pull(s)
    return t


def ray_box(rays):
    """findAABBKernel (RayBufferKernels.cu:64-135): (lo, hi) of the rays."""
    rays = _f32(rays).reshape(-1, 8)
    box = np.zeros(6, np.float32)
    lib().orc_ray_box(_p(rays, _f32p), ctypes.c_uint32(rays.shape[0]), _p(box, _f32p))
    return box


def ray_keys(rays, box=None):
    """genMortonKeysKernel (RayBufferKernels.cu:139-171): (3, r) uint64 key words."""
    rays = _f32(rays).reshape(-1, 8)
    r = rays.shape[0]
    box = ray_box(rays) if box is None else _f32(box)
    keys = np.zeros((3, r), np.uint64)
    lib().orc_ray_keys(_p(rays, _f32p), ctypes.c_uint32(r), _p(box, _f32p), _p(keys, _u64p))
    return keys


def ray_sort_order(rays):
    """Old slot of every sorted slot: stable ascending order of the keys."""
    k = ray_keys(rays)
    return np.lexsort((k[0], k[1], k[2])).astype(np.uint32)


def sah(t: Tree, ci=1.2, ct=1.0) -> float:
    s = t.struct()
    return float(lib().orc_sah(ctypes.byref(s), ci, ct))


def trbvh(t: Tree, treelet=7, iterations=3, ci=1.2, ct=1.0) -> Tree:
    s = t.struct()
    rc = lib().orc_trbvh_optimize(ctypes.byref(s), treelet, iterations, ci, ct)
    if rc != 0:
        raise ValueError(f"orc_trbvh_optimize failed ({rc})")
    return t


def atrbvh(t: Tree, treelet=9, iterations=2, ci=1.2, ct=1.0) -> Tree:
    s = t.struct()
    rc = lib().orc_atrbvh_optimize(ctypes.byref(s), treelet, iterations, ci, ct)
    if rc != 0:
        raise ValueError(f"orc_atrbvh_optimize failed ({rc})")
    return t


def treelet_pass(t: Tree, kind: str, treelet: int, gamma: int, ci=1.2, ct=1.0) -> int:
    """One bottom-up treelet pass with threshold gamma (kind "trbvh" or
    "atrbvh"), in place; returns the number of treelets processed."""
    s = t.struct()
    k = lib().orc_treelet_pass(ctypes.byref(s), 0 if kind == "trbvh" else 1, treelet, gamma,
                               ctypes.c_float(ci), ctypes.c_float(ct))
    if k < 0:
        raise RuntimeError("oracle treelet pass failed")
    t.pull(s)
    return int(k)


def collapse(t: Tree, ci=1.2, ct=1.0):
    c = Collapsed(t.n)
    cs = c.struct()
    ts = t.struct()
    out = ctypes.c_float(0.0)
    rc = lib().orc_collapse(ctypes.byref(ts), ci, ct, ctypes.byref(cs), ctypes.byref(out))
    if rc != 0:
        raise ValueError("orc_collapse failed")
    c.used = cs.used
    c.root = cs.root
    return c, float(out.value)


def trace(c: Collapsed, v4, rays, any_hit=False):
    v4 = _f32(v4)
    rays = _f32(rays)
    r = rays.shape[0]
    res = np.zeros((r, 4), np.int32)
    cs = c.struct()
    lib().orc_trace(ctypes.byref(cs), _p(v4, _f32p), _p(rays, _f32p), ctypes.c_uint32(r),
                    ctypes.c_int(1 if any_hit else 0), _p(res, _i32p))
    return res


class SbvhStats(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_uint32), ("tri_refs", ctypes.c_uint32),
                ("duplicates", ctypes.c_uint32), ("depth", ctypes.c_int32),
                ("sah", ctypes.c_float), ("sah_bvhrt", ctypes.c_float)]


class Sbvh:
    """CPU split-BVH of the reference's CPU path (SplitBVHBuilder) and its
    multi-threaded closest-hit tracer."""

    def __init__(self, v4, threads=1):
        L = lib()
        L.orc_sbvh_build.restype = ctypes.c_void_p
        L.orc_sbvh_build.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.POINTER(SbvhStats)]
        L.orc_sbvh_trace.argtypes = [ctypes.c_void_p, _f32p, _f32p, ctypes.c_uint32,
                                     ctypes.c_int, _i32p]
        L.orc_sbvh_free.argtypes = [ctypes.c_void_p]
        self.v4 = _f32(v4)
        st = SbvhStats()
        self.h = L.orc_sbvh_build(_p(self.v4, _f32p), self.v4.shape[0] // 3, threads,
                                  ctypes.byref(st))
        self.stats = {k: getattr(st, k) for k, _ in SbvhStats._fields_}

    def trace(self, rays, threads=1):
        rays = _f32(rays)
        r = rays.shape[0]
        res = np.zeros((r, 4), np.int32)
        lib().orc_sbvh_trace(self.h, _p(self.v4, _f32p), _p(rays, _f32p), r, threads,
                             _p(res, _i32p))
        return res

    def close(self):
        if self.h:
            lib().orc_sbvh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def woopify(v0, v1, v2):
    """woopifyTri (CudaBVH.cpp:579-600): the three Woop rows, float32[3, 4]."""
    out = np.zeros(12, np.float32)
    a = [_f32(x).reshape(-1)[:3].copy() for x in (v0, v1, v2)]
    lib().orc_woopify(_p(a[0], _f32p), _p(a[1], _f32p), _p(a[2], _f32p), _p(out, _f32p))
    return out.reshape(3, 4)


def compact(c: Collapsed, v4, div=16):
    """createCompact (CudaBVH.cpp:469-574): (nodes int32[M, 16], woop
    float32[R, 4], tri_index int32[R])."""
    v4 = _f32(v4)
    cs = c.struct()
    m = ctypes.c_uint32()
    r = ctypes.c_uint32()
    if lib().orc_compact(ctypes.byref(cs), _p(v4, _f32p), div, None, ctypes.byref(m), None, None,
                         ctypes.byref(r)) != 0:
        raise ValueError("orc_compact failed")
    nodes = np.zeros((m.value, 16), np.int32)
    woop = np.zeros((r.value, 4), np.float32)
    tri = np.zeros(r.value, np.int32)
    lib().orc_compact(ctypes.byref(cs), _p(v4, _f32p), div, _p(nodes, _i32p), ctypes.byref(m),
                      _p(woop, _f32p), _p(tri, _i32p), ctypes.byref(r))
    return nodes, woop, tri


def trace_compact(nodes, woop, tri_index, rays, any_hit=False):
    """kepler_dynamic_fetch `trace` semantics on a Compact2 layout; int32[r, 4]
    {id or -1, t bits, 0, 0}."""
    nodes = np.ascontiguousarray(nodes, np.int32)
    woop = _f32(woop)
    tri_index = np.ascontiguousarray(tri_index, np.int32)
    rays = _f32(rays)
    r = rays.shape[0]
    res = np.zeros((r, 4), np.int32)
    lib().orc_trace_compact(_p(nodes, _i32p), _p(woop, _f32p), _p(tri_index, _i32p),
                            _p(rays, _f32p), ctypes.c_uint32(r), ctypes.c_int(1 if any_hit else 0),
                            _p(res, _i32p))
    return res


def rays_diffuse(v4, prim, res, seed, tmax=np.inf):
    """rayGenAOKernel with one sample per ray (RayGenKernels.cu:114-221)."""
    v4 = _f32(v4)
    prim = _f32(prim)
    res = np.ascontiguousarray(res, np.int32)
    r = prim.shape[0]
    out = np.zeros((r, 8), np.float32)
    lib().orc_rays_diffuse(_p(v4, _f32p), _p(prim, _f32p), _p(res, _i32p), ctypes.c_uint32(r),
                           ctypes.c_uint32(seed), ctypes.c_float(tmax), _p(out, _f32p))
    return out


def trace_brute(v4, rays):
    v4 = _f32(v4)
    rays = _f32(rays)
    r = rays.shape[0]
    res = np.zeros((r, 4), np.int32)
    lib().orc_trace_brute(_p(v4, _f32p), ctypes.c_uint32(v4.shape[0] // 3), _p(rays, _f32p),
                          ctypes.c_uint32(r), _p(res, _i32p))
    return res


def ray_triangle(v0, v1, v2, o, d, tmin=0.0, tmax=np.inf):
    t = ctypes.c_float()
    u = ctypes.c_float()
    v = ctypes.c_float()
    a = [_f32(x) for x in (v0, v1, v2, o, d)]
    hit = lib().orc_ray_triangle(*[_p(x, _f32p) for x in a], ctypes.c_float(tmin),
                                 ctypes.c_float(tmax), ctypes.byref(t), ctypes.byref(u),
                                 ctypes.byref(v))
    return bool(hit), t.value, u.value, v.value


# ---------------------------------------------------------------- schedules
def trbvh_schedule(treelet=7, warp=32) -> np.ndarray:
    buf = np.zeros(4096, np.int32)
    rounds = lib().orc_trbvh_schedule(treelet, warp, _p(buf, _i32p), buf.size)
    if rounds < 0:
        raise ValueError("orc_trbvh_schedule failed")
    return buf[: rounds * warp].reshape(rounds, warp)


def atrbvh_schedule(treelet=9, warp=32) -> np.ndarray:
    buf = np.zeros(4096, np.int32)
    size = lib().orc_atrbvh_schedule(treelet, warp, _p(buf, _i32p), buf.size)
    if size < 0:
        raise ValueError("orc_atrbvh_schedule failed")
    return buf[:size].copy()


def reference_trbvh_schedule(treelet=7, warp=32):
    """Schedule printed by the reference TRBVHScheduler compiled into oracle/_ref,
    or None when that build is absent."""
    if not os.path.exists(_REF_SCHED):
        return None
    out = subprocess.run([_REF_SCHED, str(treelet), str(warp)], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    rounds, w = (int(x) for x in out[0].split())
    rows = [list(map(int, line.split())) for line in out[1 : 1 + rounds]]
    return np.array(rows, np.int32).reshape(rounds, w)


def lower_trm_row(i):
    return lib().orc_lower_trm_row(i)


def lower_trm_col(i):
    return lib().orc_lower_trm_col(i)


def trbvh_dp(leaf_cost, subset_area, treelet_tris, ci=1.2, ct=1.0):
    n = len(leaf_cost)
    lc = _f32(leaf_cost)
    sa = _f32(subset_area)
    costs = np.zeros(1 << n, np.float32)
    masks = np.zeros(1 << n, np.int32)
    v = lib().orc_trbvh_dp(n, _p(lc, _f32p), _p(sa, _f32p), treelet_tris, ci, ct,
                           _p(costs, _f32p), _p(masks, _i32p))
    return float(v), costs, masks


def find_minimum(D, last_row: int) -> int:
    d = _f32(np.zeros(496, np.float32))
    src = _f32(D)
    d[: src.size] = src
    return lib().orc_find_minimum(_p(d, _f32p), last_row)


def find_leaf_to_expand(areas) -> int:
    a = _f32(areas)
    return lib().orc_find_leaf_to_expand(a.size, _p(a, _f32p))


def atrbvh_cluster(boxes, leaf_sah, ci=1.2):
    n = len(leaf_sah)
    b = _f32(boxes).reshape(n, 6)
    s = _f32(leaf_sah)
    merges = np.zeros((n - 1, 2), np.int32)
    areas = np.zeros(n, np.float32)
    rc = lib().orc_atrbvh_cluster(n, _p(b, _f32p), _p(s, _f32p), ci, _p(merges, _i32p),
                                  _p(areas, _f32p))
    if rc != 0:
        raise ValueError("orc_atrbvh_cluster failed")
    return merges, areas
