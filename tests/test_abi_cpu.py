"""CPU checks of the C-ABI library: it loads without a GPU, exports every
symbol the public headers declare, validates arguments before touching the
device, and its host utilities (scene generators, lane schedule) behave."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

import atrbvh_amd as ab
import oracle

HDRS = [os.path.join(ab.INCLUDE_DIR, h) for h in ("atrbvh.h", "atrbvh_scenes.h")]


def declared_symbols():
    names = set()
    for h in HDRS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(atr_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ab.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_symbols()) >= 40


def test_oracle_library_loads():
    assert oracle.lib().orc_expand_bits10(1) == 1


def test_version_string():
    assert b"gfx950" in ab.lib().atr_version()


def _gpu_present():
    import subprocess
    try:
        out = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=20).stdout
    except Exception:
        return False
    return "gfx9" in out


def test_argument_validation_without_device():
    L = ab.lib()
    out = ctypes.c_void_p()
    assert L.atr_context_create(0, None) == 1
    assert L.atr_lbvh_build(None, 30, ctypes.byref(out)) == 1
    assert L.atr_trbvh_optimize(None, 7, 3, 1.2, 1.0) == 1
    assert L.atr_atrbvh_optimize(None, 9, 2, 1.2, 1.0) == 1
    assert L.atr_collapse(None, 1.2, 1.0, ctypes.byref(out), None) == 1
    assert L.atr_trace(None, None, 0, 0, None) == 1
    assert L.atr_sync(None) == 1
    assert L.atr_scene_create(None, None, 10, None, None, ctypes.byref(out)) == 1
    assert L.atr_sort_pairs(None, 32, None, None, 0) == 1
    # the stage calls
    assert L.atr_build_tree(None, None, None, 10, 30, ctypes.byref(out)) == 1
    assert L.atr_build_tree_boxes(None, None, None, None, None, None, 10, 30,
                                  ctypes.byref(out)) == 1
    assert L.atr_trbvh_pass(None, 7, 7, 1.2, 1.0) == 1
    assert L.atr_atrbvh_pass(None, 9, 9, 1.2, 1.0) == 1
    assert L.atr_collapse_costs(None, 1.2, 1.0, ctypes.byref(out)) == 1
    assert L.atr_collapse_lists(None, None, None) == 1
    assert b"NULL" in L.atr_last_error() or len(L.atr_last_error()) > 0
    # size limits are checked before anything else
    assert L.atr_morton_encode(None, None, 1 << 30, None, None, 30, None, None) == 4
    assert L.atr_scene_create(None, None, 1 << 30, None, None, ctypes.byref(out)) == 4
    assert L.atr_scene_create_device(None, None, 1 << 30, None, None, ctypes.byref(out)) == 4
    assert L.atr_scene_create(None, None, (1 << 30) - 1, None, None, ctypes.byref(out)) == 1
    assert L.atr_trace(None, None, 2**31 + 1, 0, None) == 1
    assert b"2^31" in L.atr_last_error()
    eye = (ctypes.c_float * 3)()
    basis = (ctypes.c_float * 9)()
    assert L.atr_rays_pinhole(None, eye, basis, ctypes.c_float(1.0), 65536, 32769,
                              ctypes.c_float(1.0), None) == 1
    assert b"2^31" in L.atr_last_error()
    assert L.atr_rays_pinhole(None, eye, basis, ctypes.c_float(1.0), 65536, 32768,
                              ctypes.c_float(1.0), None) == 1
    assert b"NULL" in L.atr_last_error()
    # destroy of NULL handles is a no-op
    assert L.atr_tree_destroy(None) == 0 and L.atr_scene_destroy(None) == 0


def test_context_without_gpu_reports_device_error():
    if _gpu_present():
        pytest.skip("a GPU is present")
    with pytest.raises(ab.AtrError) as e:
        ab.Context(0)
    assert e.value.status == 3


def test_scene_generators_are_deterministic_and_shaped():
    a = ab.gen_soup(1, 1000)
    b = ab.gen_soup(1, 1000)
    c = ab.gen_soup(2, 1000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    t = a.reshape(-1, 3, 4)
    cen = t[:, :, :3].mean(axis=1)
    assert np.all(t[:, :, 3] == 0)
    assert 0.45 < cen.mean() < 0.55
    dev = (t[:, :, :3] - cen[:, None, :]).std()
    assert 0.003 < dev < 0.006

    g = ab.gen_gmm(2, 20000).reshape(-1, 3, 4)[:, :, :3]
    e = np.linalg.norm(g[:, 1] - g[:, 0], axis=1)
    assert 0.0015 < np.median(e) < 0.0025  # median edge 0.002
    e2 = np.linalg.norm(g[:, 2] - g[:, 1], axis=1)
    assert np.allclose(e, e2, rtol=1e-3, atol=1e-7)  # equilateral


def test_library_schedule_equals_oracle():
    for n in (3, 4, 5, 6, 7):
        assert np.array_equal(ab.trbvh_schedule(n, 32), oracle.trbvh_schedule(n, 32))


def test_cpp_facade_compiles_and_reports_errors(tmp_path):
    """include/bvhrt_compat.hpp builds warning-free under -Wall -Wextra -Werror
    and turns a library failure into BVHRT::Error (without a GPU: the device
    error of context creation)."""
    from helpers import build_compat_driver, run_compat_driver

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    build_compat_driver(root)
    if _gpu_present():
        pytest.skip("a GPU is present")
    rc, out, res = run_compat_driver(root, ab.gen_soup(1, 64), False, "atrbvh", str(tmp_path))
    assert rc == 3 and res is None
    assert out.startswith("status 3:")


def test_batch_argument_validation_without_device():
    L = ab.lib()
    out = ctypes.c_void_p()
    off = np.array([0, 5, 9], np.uint32)
    v = np.zeros((27, 4), np.float32)
    p32 = off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    assert L.atr_batch_create(None, v.ctypes.data, p32, 2, None, ctypes.byref(out)) == 1
    assert b"NULL" in L.atr_last_error()
    # offsets are checked before the context: a mesh of fewer than 2 triangles,
    # also where an offset near 2^32 would wrap a 32-bit check
    for bad, mesh in (([0, 5, 6], 1), ([0, 4294967295, 5], 1), ([0, 1, 9], 0)):
        o = np.array(bad, np.uint32)
        assert L.atr_batch_create(None, v.ctypes.data, o.ctypes.data_as(
            ctypes.POINTER(ctypes.c_uint32)), 2, None, ctypes.byref(out)) == 1
        assert f"mesh {mesh} has fewer".encode() in L.atr_last_error()
    big = np.array([0, 1 << 30], np.uint32)
    assert L.atr_batch_create(None, v.ctypes.data, big.ctypes.data_as(
        ctypes.POINTER(ctypes.c_uint32)), 1, None, ctypes.byref(out)) == 4
    assert L.atr_batch_build(None, 9, 2, 1.2, 1.0, 1) == 1
    assert L.atr_batch_destroy(None) == 0


def test_batch_generators():
    off = ab.gen_batch_sizes(3, 4096, 64, 65536)
    sizes = np.diff(off.astype(np.int64))
    assert sizes.min() >= 64 and sizes.max() <= 65536
    assert 30e6 < off[-1] < 50e6  # about 40M triangles
    # log-uniform: the median size is near the geometric mean of the range
    assert 1500 < np.median(sizes) < 3000
    small = off[:9]
    v = ab.gen_soup_batch(3000, small)
    for m in range(8):
        vm = v[3 * int(small[m]): 3 * int(small[m + 1])]
        assert np.array_equal(vm, ab.gen_soup(3000 + m, int(small[m + 1] - small[m])))


def _view(t: dict):
    keep = {k: np.ascontiguousarray(t[k], np.int32) for k in ("parent", "left", "right", "data")}
    v = ab.TreeView(int(t["n"]), 0, ab._ptr(keep["parent"], ab._i32p),
                    ab._ptr(keep["left"], ab._i32p), ab._ptr(keep["right"], ab._i32p),
                    ab._ptr(keep["data"], ab._i32p), None, None, None)
    return v, keep


def _oracle_tree(n=64, seed=5):
    v = ab.gen_soup(seed, n)
    mn, mx = oracle.scene_bbox(v)
    t = oracle.lbvh(v, mn, mx, 30)
    oracle.atrbvh(t, 9, 1)
    return {"n": n, "parent": t.parent.copy(), "left": t.left.copy(), "right": t.right.copy(),
            "data": t.data.copy()}


def _check(t: dict) -> int:
    v, keep = _view(t)
    return ab.lib().atr_tree_view_check(ctypes.byref(v))


def test_tree_view_check_accepts_valid_trees():
    assert _check(_oracle_tree()) == 0
    assert _check(_oracle_tree(2, 9)) == 0


def test_tree_view_check_rejects_bad_parent_array():
    t = _oracle_tree()
    t["parent"][t["left"][0]] = 3  # a child whose parent entry names another node
    assert _check(t) == 1
    assert b"parent" in ab.lib().atr_last_error()
    t = _oracle_tree()
    t["parent"][0] = 0
    assert _check(t) == 1


def test_tree_view_check_rejects_cycles_and_bad_leaves():
    n = 64
    t = _oracle_tree(n)
    # Detach the root's left subtree and make a two-node cycle of internal
    # nodes: every non-root node is still a child exactly once and every
    # parent entry matches, but the cycle is unreachable from the root.
    a = int(t["left"][0])
    assert a < n - 1
    b = int(t["left"][a])
    if b >= n - 1:
        b = int(t["right"][a])
    assert b < n - 1, "need two internal nodes"
    # root.left <- b.left and b.left <- a: a and b are each other's parent
    bl = int(t["left"][b])
    t["left"][0] = bl
    t["parent"][bl] = 0
    t["left"][b] = a
    t["parent"][a] = b
    assert _check(t) == 1
    assert b"unreachable" in ab.lib().atr_last_error()
    t = _oracle_tree(n)
    leaf = n - 1 + 3
    t["left"][leaf] = 1
    assert _check(t) == 1
    t = _oracle_tree(n)
    t["data"][n - 1] = t["data"][n]  # duplicate triangle
    assert _check(t) == 1
    t = _oracle_tree(n)
    t["data"][0] = 7  # internal node with data
    assert _check(t) == 1


def test_no_alloc_sort_argument_validation_without_device():
    L = ab.lib()
    assert L.atr_sort_pairs_noalloc(None, 32, None, None, 0) == 1


class _Buf:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = None


def test_traversal_wrappers_refuse_short_buffers():
    """The C-ABI sees plain pointers; the Python layer refuses a count that
    runs past a DeviceBuffer before anything reaches a kernel (with a null
    handle, a call that got past the check would fail with another message)."""
    c = ab.Collapsed.__new__(ab.Collapsed)
    c.h = None
    ctx = ab.Context.__new__(ab.Context)
    ctx.h = None
    n = 70001
    with pytest.raises(ab.AtrError, match="rays: buffer of"):
        c.trace(_Buf(21846 * 32), n, _Buf(n * 16))
    with pytest.raises(ab.AtrError, match="results: buffer of"):
        c.trace(_Buf(n * 32), n, _Buf(n * 16 - 1))
    with pytest.raises(ab.AtrError, match="buffer of"):
        ctx.trace_compact(n, _Buf(n * 32 - 32), _Buf(n * 16), None)
    with pytest.raises(ab.AtrError, match="buffer of"):
        ctx.rays_pinhole((0, 0, 0), np.eye(3), 1.0, 8, 8, 1.0, _Buf(63 * 32))
    with pytest.raises(ab.AtrError, match="buffer of"):
        ctx.rays_diffuse(None, _Buf(n * 32), _Buf(n * 16), n, 1, 1.0, _Buf(n * 31))
    with pytest.raises(ab.AtrError, match="buffer of"):
        ctx.rays_morton_sort(_Buf(n * 32), n, _Buf(n * 32), _Buf(n * 4 - 4))
    with pytest.raises(ab.AtrError, match="buffer of"):
        ctx.rays_unsort_results(_Buf(n * 16), _Buf(n * 4), n, _Buf(n * 8))
    with pytest.raises(ab.AtrError, match="keys: buffer of"):
        ctx.morton_encode(_Buf(n * 48), n, (0, 0, 0), (1, 1, 1), 63, _Buf(n * 4), None)
    with pytest.raises(ab.AtrError, match="values: buffer of"):
        ctx.sort_pairs(32, _Buf(n * 4), _Buf(n * 2), n)
    # sizes that fit pass the check and reach the library, which refuses the
    # null handle
    with pytest.raises(ab.AtrError) as e:
        c.trace(_Buf(n * 32), n, _Buf(n * 16))
    assert "buffer of" not in str(e.value)


class _OkLib:
    """Stands in for the library: every call succeeds (exercises the Python
    side of the wrappers only)."""

    def __getattr__(self, name):
        return lambda *a: 0


def test_traversal_wrappers_accept_exact_sizes(monkeypatch):
    monkeypatch.setattr(ab, "lib", lambda: _OkLib())
    c = ab.Collapsed.__new__(ab.Collapsed)
    c.h = None
    ctx = ab.Context.__new__(ab.Context)
    ctx.h = None
    n = 1000
    c.trace(_Buf(n * 32), n, _Buf(n * 16))
    lay = ab.CompactLayout()
    ctx.trace_compact(n, _Buf(n * 32), _Buf(n * 16), lay)
    ctx.rays_pinhole((0, 0, 0), np.eye(3), 1.0, 40, 25, 1.0, _Buf(n * 32))
    s = ab.Scene.__new__(ab.Scene)
    s.h = None
    ctx.rays_diffuse(s, _Buf(n * 32), _Buf(n * 16), n, 1, 1.0, _Buf(n * 32))
    ctx.rays_morton_sort(_Buf(n * 32), n, _Buf(n * 32), _Buf(n * 4), _Buf(n * 4), _Buf(n * 4))
    ctx.rays_morton_sort(_Buf(n * 32), n, _Buf(n * 32), _Buf(n * 4))
    ctx.rays_unsort_results(_Buf(n * 16), _Buf(n * 4), n, _Buf(n * 16))
    ctx.morton_encode(_Buf(n * 48), n, (0, 0, 0), (1, 1, 1), 63, _Buf(n * 8), _Buf(n * 4))
    ctx.morton_encode(_Buf(n * 48), n, (0, 0, 0), (1, 1, 1), 30, _Buf(n * 4), None)
    ctx.sort_pairs(64, _Buf(n * 8), _Buf(n * 4), n)
    ctx.sort_pairs(32, _Buf(n * 4), None, n)
    b = ab.Batch.__new__(ab.Batch)
    b.h = None
    b.meshes = 10
    b.trace(_Buf(n * 32), n // 10, _Buf(n * 16))
    with pytest.raises(ab.AtrError, match="hits: buffer of"):
        b.trace(_Buf(n * 32), n // 10, _Buf(n * 16 - 16))
