"""Python binding of the MI355X BVH builder / traverser (libatrbvh.so, C ABI in
include/atrbvh.h).

The host-side names mirror the BVHRT-Core builder API so tests read like the
reference's call sequence (gpu-ray-traversal/src/rt/cuda/Renderer.cpp:170-216):

    ctx = Context(0)
    scene = ctx.scene(vertices)                    # SceneWrapper
    tree = LBVHBuilder(use64=True).BuildTree(scene) # LBVHBuilder::BuildTree
    AgglomerativeTreeletOptimizer(9, 2).Optimize(tree)
    collapsed, sah = BVHTreeCollapser().Collapse(tree)

Every call goes to the gfx950 kernels; there is no CPU fallback.  Importing
works without a GPU (the library loads); creating a Context needs one.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libatrbvh.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p

ATR_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "OUT_OF_MEMORY", 3: "DEVICE", 4: "UNSUPPORTED",
                5: "INTERNAL"}


class AtrError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class TreeView(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_uint32), ("root", ctypes.c_int32),
                ("parent", _i32p), ("left", _i32p), ("right", _i32p), ("data", _i32p),
                ("bbmin", _f32p), ("bbmax", _f32p), ("area", _f32p)]


class CollapsedView(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_uint32), ("root", ctypes.c_int32),
                ("used", ctypes.c_uint32), ("parent", _i32p), ("left", _i32p),
                ("right", _i32p), ("data", _i32p), ("triangle_count", _i32p),
                ("bbmin", _f32p), ("bbmax", _f32p), ("triangle_indices", _i32p)]


_lib = None


def lib():
    """Load libatrbvh.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with atrbvh_amd/build.py "
                               "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.atr_last_error.restype = ctypes.c_char_p
        L.atr_version.restype = ctypes.c_char_p
        L.atr_context_stream.restype = _vp
        L.atr_context_stream.argtypes = [_vp]
        L.atr_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
        for name in ("atr_profile_count", "atr_context_destroy", "atr_sync",
                     "atr_profile_reset"):
            getattr(L, name).argtypes = [_vp]
        L.atr_profile_enable.argtypes = [_vp, ctypes.c_int]
        L.atr_malloc.argtypes = [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]
        L.atr_free.argtypes = [_vp, _vp]
        L.atr_memcpy_h2d.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        L.atr_memcpy_d2h.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        L.atr_memset_d32.argtypes = [_vp, _vp, ctypes.c_uint32, ctypes.c_size_t]
        L.atr_profile_get.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                      ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
        L.atr_scene_create.argtypes = [_vp, _vp, ctypes.c_uint32, _vp, _vp, ctypes.POINTER(_vp)]
        L.atr_scene_create_device.argtypes = L.atr_scene_create.argtypes
        L.atr_scene_bounds.argtypes = [_vp, _f32p, _f32p]
        L.atr_scene_vertices.argtypes = [_vp, ctypes.POINTER(_vp)]
        L.atr_scene_destroy.argtypes = [_vp]
        L.atr_lbvh_build.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp)]
        L.atr_rays_morton_sort.argtypes = [_vp, _vp, ctypes.c_uint32, _vp, _vp, _vp, _vp]
        L.atr_rays_morton_keys.argtypes = [_vp, _vp, ctypes.c_uint32, _vp, _vp]
        L.atr_rays_unsort_results.argtypes = [_vp, _vp, _vp, ctypes.c_uint32, _vp]
        L.atr_lbvh_build_boxes.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_uint32, ctypes.c_int,
                                           ctypes.POINTER(_vp)]
        L.atr_trbvh_optimize.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                         ctypes.c_float]
        L.atr_atrbvh_optimize.argtypes = L.atr_trbvh_optimize.argtypes
        L.atr_collapse.argtypes = [_vp, ctypes.c_float, ctypes.c_float, ctypes.POINTER(_vp),
                                   _f32p]
        L.atr_build_tree.argtypes = [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.POINTER(_vp)]
        L.atr_build_tree_boxes.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_uint32,
                                           ctypes.c_int, ctypes.POINTER(_vp)]
        L.atr_trbvh_pass.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_float]
        L.atr_atrbvh_pass.argtypes = L.atr_trbvh_pass.argtypes
        L.atr_collapse_costs.argtypes = [_vp, ctypes.c_float, ctypes.c_float, ctypes.POINTER(_vp)]
        L.atr_collapse_lists.argtypes = [_vp, _vp, _f32p]
        L.atr_collapsed_stats.argtypes = [_vp, _f32p, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_uint32)]
        L.atr_tree_sah.argtypes = [_vp, ctypes.c_float, ctypes.c_float, _f32p]
        L.atr_tree_num_triangles.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint32)]
        L.atr_collapsed_num_triangles.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint32)]
        L.atr_tree_download.argtypes = [_vp, ctypes.POINTER(TreeView)]
        L.atr_tree_device_view.argtypes = [_vp, ctypes.POINTER(TreeView)]
        L.atr_tree_upload.argtypes = [_vp, ctypes.POINTER(TreeView)]
        L.atr_tree_view_check.argtypes = [ctypes.POINTER(TreeView)]
        L.atr_tree_create_from_host.argtypes = [_vp, ctypes.POINTER(TreeView),
                                                ctypes.POINTER(_vp)]
        L.atr_tree_bind_scene.argtypes = [_vp, _vp]
        L.atr_tree_destroy.argtypes = [_vp]
        L.atr_collapsed_download.argtypes = [_vp, ctypes.POINTER(CollapsedView)]
        L.atr_collapsed_destroy.argtypes = [_vp]
        L.atr_trace.argtypes = [_vp, _vp, ctypes.c_uint32, ctypes.c_int, _vp]
        L.atr_collapsed_layout.argtypes = [_vp]
        L.atr_rays_pinhole.argtypes = [_vp, _f32p, _f32p, ctypes.c_float, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.c_float, _vp]
        L.atr_rays_diffuse.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_float, _vp]
        L.atr_morton_encode.argtypes = [_vp, _vp, ctypes.c_uint32, _f32p, _f32p, ctypes.c_int,
                                        _vp, _vp]
        L.atr_sort_pairs.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint32]
        L.atr_sort_pairs_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint32]
        L.atr_refit.argtypes = [_vp, _vp]
        L.atr_mesh_load_obj.argtypes = [ctypes.c_char_p, ctypes.POINTER(_vp)]
        L.atr_mesh_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.POINTER(ctypes.c_uint32)]
        L.atr_mesh_soup.argtypes = [_vp, _vp]
        L.atr_mesh_destroy.argtypes = [_vp]
        L.atr_scene_create_from_mesh.argtypes = [_vp, _vp, ctypes.POINTER(_vp)]
        L.atr_collapsed_compact.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(CompactLayout)]
        L.atr_trace_compact.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                        _vp, _vp, _vp, _vp, _vp]
        L.atr_gen_soup.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_float, _f32p]
        L.atr_gen_soup.restype = None
        L.atr_gen_gmm.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                  _f32p]
        L.atr_gen_gmm.restype = None
        L.atr_trbvh_schedule.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int]
        L.atr_gen_batch_sizes.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint32, _u32p]
        L.atr_gen_batch_sizes.restype = None
        L.atr_gen_soup_batch.argtypes = [ctypes.c_uint64, ctypes.c_uint32, _u32p, ctypes.c_float,
                                         _f32p]
        L.atr_gen_soup_batch.restype = None
        L.atr_batch_create.argtypes = [_vp, _vp, _u32p, ctypes.c_uint32, _vp,
                                       ctypes.POINTER(_vp)]
        L.atr_batch_create_device.argtypes = L.atr_batch_create.argtypes
        L.atr_batch_vertices.argtypes = [_vp, ctypes.POINTER(_vp)]
        L.atr_batch_build.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_int]
        L.atr_batch_sah.argtypes = [_vp, ctypes.c_float, ctypes.c_float, _f32p]
        L.atr_batch_extract.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(_vp)]
        L.atr_batch_destroy.argtypes = [_vp]
        L.atr_batch_sah_device.argtypes = [_vp, ctypes.c_float, ctypes.c_float, _vp]
        L.atr_tree_sah_device.argtypes = [_vp, ctypes.c_float, ctypes.c_float, _vp]
        L.atr_batch_set_config.argtypes = [_vp, ctypes.POINTER(BatchConfig)]
        L.atr_batch_get_config.argtypes = [_vp, ctypes.POINTER(BatchConfig)]
        L.atr_batch_trace.argtypes = [_vp, _vp, ctypes.c_uint32, _vp]
        L.atr_batch_hit_adjoint.argtypes = [_vp, _vp, _vp, ctypes.c_uint32, _vp, _vp]
        L.atr_batch_step.argtypes = [_vp, _vp, ctypes.c_float, ctypes.c_float]
        L.atr_batch_refit.argtypes = [_vp]
        L.atr_batch_rebuild.argtypes = [_vp, _u32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_float, ctypes.c_float]
        L.atr_hit_adjoint.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_uint32, _vp, _vp,
                                      ctypes.c_uint32, _vp]
        L.atr_depth_loss.argtypes = [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_int, _vp, _vp]
        L.atr_mesh_step.argtypes = [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_float,
                                    ctypes.c_float, _vp, _vp]
        _lib = L
    return _lib


def check(status: int):
    if status != ATR_OK:
        raise AtrError(status, lib().atr_last_error().decode())


def _ptr(a: np.ndarray, t=_f32p):
    return a.ctypes.data_as(t)


RAY_BYTES = 32      # float4 (origin, tmin) + float4 (direction, tmax)
RESULT_BYTES = 16   # int4 (triangle, t bits, u bits, v bits)


def _need(buf, nbytes: int, what: str):
    """The C-ABI takes plain pointers and cannot see a buffer's size: a call
    whose count runs past a DeviceBuffer is refused here, before any kernel
    reads or writes beyond the allocation."""
    if buf is not None and int(nbytes) > buf.nbytes:
        raise AtrError(1, f"{what}: buffer of {buf.nbytes} bytes, the call needs {int(nbytes)}")


# ---------------------------------------------------------------- scenes (host)
def gen_soup(seed: int, n: int, sigma: float = 0.005) -> np.ndarray:
    v = np.zeros((3 * n, 4), np.float32)
    lib().atr_gen_soup(seed, n, sigma, _ptr(v))
    return v


def gen_gmm(seed: int, n: int, clusters: int = 64, sigma_lo: float = 0.005, sigma_hi: float = 0.1,
            edge_median: float = 0.002, edge_sigma: float = 0.7) -> np.ndarray:
    v = np.zeros((3 * n, 4), np.float32)
    lib().atr_gen_gmm(seed, n, clusters, sigma_lo, sigma_hi, edge_median, edge_sigma, _ptr(v))
    return v


def gen_batch_sizes(seed: int, meshes: int, lo: int = 64, hi: int = 65536) -> np.ndarray:
    """Prefix offsets (meshes + 1) of log-uniform mesh sizes in [lo, hi]."""
    off = np.zeros(meshes + 1, np.uint32)
    lib().atr_gen_batch_sizes(seed, meshes, lo, hi, _ptr(off, _u32p))
    return off


def gen_soup_batch(seed_base: int, offsets: np.ndarray, sigma: float = 0.005) -> np.ndarray:
    """Mesh m = gen_soup(seed_base + m, N_m, sigma), concatenated mesh-major."""
    off = np.ascontiguousarray(offsets, np.uint32)
    v = np.zeros((3 * int(off[-1]), 4), np.float32)
    lib().atr_gen_soup_batch(seed_base, off.size - 1, _ptr(off, _u32p), sigma, _ptr(v))
    return v


class Mesh:
    """A Wavefront OBJ scene read as the reference's importer reads it."""

    def __init__(self, path: str):
        h = _vp()
        check(lib().atr_mesh_load_obj(os.fsencode(path), ctypes.byref(h)))
        self.h = h.value

    def info(self) -> dict:
        n = ctypes.c_uint32()
        p = ctypes.c_uint32()
        check(lib().atr_mesh_info(self.h, ctypes.byref(n), ctypes.byref(p)))
        return {"triangles": n.value, "positions": p.value}

    def soup(self) -> np.ndarray:
        v = np.zeros((3 * self.info()["triangles"], 4), np.float32)
        if v.size:
            check(lib().atr_mesh_soup(self.h, v.ctypes.data))
        return v

    def scene(self, ctx: "Context") -> "Scene":
        """SceneConverter::ConvertToCudaSceneDeviceMemory."""
        h = _vp()
        check(lib().atr_scene_create_from_mesh(ctx.h, self.h, ctypes.byref(h)))
        s = Scene.__new__(Scene)
        s.ctx = ctx
        s.n = self.info()["triangles"]
        s.h = h.value
        return s

    def close(self):
        if self.h:
            lib().atr_mesh_destroy(self.h)
            self.h = None


def trbvh_schedule(treelet: int = 7, warp: int = 32) -> np.ndarray:
    buf = np.zeros(4096, np.int32)
    rounds = lib().atr_trbvh_schedule(treelet, warp, _ptr(buf, _i32p), buf.size)
    if rounds < 0:
        raise AtrError(1, "schedule generation failed")
    return buf[: rounds * warp].reshape(rounds, warp)


# ---------------------------------------------------------------- device objects
class DeviceBuffer:
    """Device memory from the context's caching allocator."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        check(lib().atr_malloc(ctx.h, max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().atr_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes))

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(lib().atr_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().atr_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0):
        h = _vp()
        check(lib().atr_context_create(device, ctypes.byref(h)))
        self.h = h.value

    def close(self):
        if self.h:
            lib().atr_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(lib().atr_sync(self.h))

    def buffer(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def to_device(self, a: np.ndarray) -> DeviceBuffer:
        b = DeviceBuffer(self, a.nbytes)
        b.upload(a)
        return b

    def scene(self, vertices: np.ndarray, bbmin=None, bbmax=None) -> "Scene":
        return Scene(self, vertices, bbmin, bbmax)

    # profiling
    def profile(self, on: bool = True):
        check(lib().atr_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        check(lib().atr_profile_reset(self.h))

    def profile_records(self):
        out = {}
        for i in range(lib().atr_profile_count(self.h)):
            name = ctypes.c_char_p()
            launches = ctypes.c_int()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            check(lib().atr_profile_get(self.h, i, ctypes.byref(name), ctypes.byref(launches),
                                        ctypes.byref(ms), ctypes.byref(by)))
            out[name.value.decode()] = {"launches": launches.value, "ms": ms.value,
                                        "bytes": by.value}
        return out

    # low-level stages (DeviceGenerateMortonCodes / DeviceSort)
    def morton_encode(self, d_vertices: DeviceBuffer, n: int, bbmin, bbmax, bits: int,
                      d_keys: DeviceBuffer, d_idx: DeviceBuffer | None):
        mn = np.asarray(bbmin, np.float32)
        mx = np.asarray(bbmax, np.float32)
        _need(d_vertices, n * 48, "vertices")
        _need(d_keys, n * (8 if bits == 63 else 4), "keys")
        _need(d_idx, n * 4, "indices")
        check(lib().atr_morton_encode(self.h, d_vertices.ptr, n, _ptr(mn), _ptr(mx), bits,
                                      d_keys.ptr, d_idx.ptr if d_idx is not None else None))

    def sort_pairs(self, key_bits: int, d_keys: DeviceBuffer, d_vals: DeviceBuffer | None, n: int):
        _need(d_keys, n * (key_bits // 8), "keys")
        _need(d_vals, n * 4, "values")
        check(lib().atr_sort_pairs(self.h, key_bits, d_keys.ptr,
                                   d_vals.ptr if d_vals is not None else None, n))

    # differentiable hits
    def hit_adjoint(self, scene: "Scene", d_rays: DeviceBuffer, d_hits: DeviceBuffer, n: int,
                    d_grad_tuv: DeviceBuffer, d_index: DeviceBuffer | None, n_grad: int,
                    d_grad: DeviceBuffer):
        check(lib().atr_hit_adjoint(self.h, scene.h, d_rays.ptr, d_hits.ptr, n, d_grad_tuv.ptr,
                                    d_index.ptr if d_index is not None else None, n_grad,
                                    d_grad.ptr))

    def depth_loss(self, d_hits: DeviceBuffer, d_target: DeviceBuffer, n: int,
                   d_grad_tuv: DeviceBuffer, d_loss: DeviceBuffer, offset: int = 0,
                   same_triangle: bool = False):
        check(lib().atr_depth_loss(self.h, d_hits.ptr, d_target.ptr, n, int(same_triangle),
                                   d_grad_tuv.ptr, d_loss.ptr + 8 * offset))

    def mesh_step(self, d_mesh: DeviceBuffer, n_vertices: int, d_grad: DeviceBuffer | None,
                  lr: float, d_index: DeviceBuffer, scene: "Scene", max_step: float = 0.0):
        check(lib().atr_mesh_step(self.h, d_mesh.ptr, n_vertices,
                                  d_grad.ptr if d_grad is not None else None, lr, max_step,
                                  d_index.ptr, scene.h))

    def trace_compact(self, n: int, d_rays: DeviceBuffer, d_results: DeviceBuffer,
                      layout: "CompactLayout", any_hit: bool = False):
        """The reference's `trace` kernel ABI on a Compact2 layout."""
        _need(d_rays, n * RAY_BYTES, "rays")
        _need(d_results, n * RESULT_BYTES, "results")
        check(lib().atr_trace_compact(self.h, n, 1 if any_hit else 0, d_rays.ptr, d_results.ptr,
                                      layout.nodes, None, None, None, layout.tris, None, None,
                                      layout.tri_indices))

    def rays_pinhole(self, eye, basis, tan_half: float, w: int, h: int, tmax: float,
                     d_rays: DeviceBuffer):
        e = np.asarray(eye, np.float32)
        b = np.asarray(basis, np.float32).reshape(9)
        _need(d_rays, w * h * RAY_BYTES, "rays")
        check(lib().atr_rays_pinhole(self.h, _ptr(e), _ptr(b), tan_half, w, h, tmax, d_rays.ptr))

    def rays_morton_sort(self, d_rays: DeviceBuffer, n: int, d_out: DeviceBuffer,
                         d_slot_to_id: DeviceBuffer, d_id_to_slot: DeviceBuffer | None = None,
                         d_slot_to_id_in: DeviceBuffer | None = None):
        """RayBuffer::mortonSort (gpu-ray-traversal/src/rt/ray/RayBuffer.cpp:103-163)."""
        for b, size, what in ((d_rays, RAY_BYTES, "rays"), (d_out, RAY_BYTES, "sorted rays"),
                              (d_slot_to_id, 4, "slot_to_id"), (d_id_to_slot, 4, "id_to_slot"),
                              (d_slot_to_id_in, 4, "slot_to_id_in")):
            _need(b, n * size, what)
        check(lib().atr_rays_morton_sort(self.h, d_rays.ptr, n,
                                         d_slot_to_id_in.ptr if d_slot_to_id_in else None,
                                         d_out.ptr, d_slot_to_id.ptr,
                                         d_id_to_slot.ptr if d_id_to_slot else None))

    def rays_morton_keys(self, d_rays: DeviceBuffer, n: int, d_keys: DeviceBuffer,
                         d_box: DeviceBuffer):
        check(lib().atr_rays_morton_keys(self.h, d_rays.ptr, n, d_keys.ptr, d_box.ptr))

    def rays_unsort_results(self, d_sorted: DeviceBuffer, d_slot_to_id: DeviceBuffer, n: int,
                            d_out: DeviceBuffer):
        _need(d_sorted, n * RESULT_BYTES, "sorted results")
        _need(d_slot_to_id, n * 4, "slot_to_id")
        _need(d_out, n * RESULT_BYTES, "results")
        check(lib().atr_rays_unsort_results(self.h, d_sorted.ptr, d_slot_to_id.ptr, n,
                                            d_out.ptr))

    def rays_diffuse(self, scene: "Scene", d_primary: DeviceBuffer, d_results: DeviceBuffer,
                     n: int, seed: int, tmax: float, d_rays: DeviceBuffer):
        _need(d_primary, n * RAY_BYTES, "primary rays")
        _need(d_results, n * RESULT_BYTES, "primary results")
        _need(d_rays, n * RAY_BYTES, "rays")
        check(lib().atr_rays_diffuse(self.h, scene.h, d_primary.ptr, d_results.ptr, n, seed, tmax,
                                     d_rays.ptr))


class Scene:
    """SceneWrapper (BVHRT-Core/src/SceneWrapper.cpp:8-30); vertices float4[3n]."""

    def __init__(self, ctx: Context, vertices: np.ndarray, bbmin=None, bbmax=None):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4)
        assert v.shape[0] % 3 == 0
        self.ctx = ctx
        self.n = v.shape[0] // 3
        h = _vp()
        if bbmin is None:
            check(lib().atr_scene_create(ctx.h, v.ctypes.data, self.n, None, None,
                                         ctypes.byref(h)))
        else:
            mn = np.asarray(bbmin, np.float32)
            mx = np.asarray(bbmax, np.float32)
            check(lib().atr_scene_create(ctx.h, v.ctypes.data, self.n, mn.ctypes.data,
                                         mx.ctypes.data, ctypes.byref(h)))
        self.h = h.value

    @classmethod
    def from_device(cls, ctx: Context, d_vertices: int, n: int) -> "Scene":
        """Scene from float4[3n] already on the device (copied device-to-device)."""
        s = cls.__new__(cls)
        s.ctx = ctx
        s.n = int(n)
        h = _vp()
        check(lib().atr_scene_create_device(ctx.h, d_vertices, s.n, None, None, ctypes.byref(h)))
        s.h = h.value
        return s

    def bounds(self):
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        check(lib().atr_scene_bounds(self.h, _ptr(mn), _ptr(mx)))
        return mn, mx

    def device_vertices(self) -> int:
        p = _vp()
        check(lib().atr_scene_vertices(self.h, ctypes.byref(p)))
        return p.value

    def close(self):
        if self.h:
            lib().atr_scene_destroy(self.h)
            self.h = None


class Tree:
    """Device tree in the BVHRT SoA layout (SoABVHTree, BVHRT-Core/src/SoABVHTree.h)."""

    def __init__(self, ctx: Context, handle: int, n: int):
        self.ctx = ctx
        self.h = handle
        self.n = n

    @classmethod
    def from_host(cls, ctx: Context, host: dict, scene: Scene | None = None) -> "Tree":
        n = int(host["n"])
        arrs = {k: np.ascontiguousarray(host[k], np.int32) for k in ("parent", "left", "right",
                                                                       "data")}
        bmin = np.ascontiguousarray(host["bbmin"], np.float32).reshape(-1, 4)
        bmax = np.ascontiguousarray(host["bbmax"], np.float32).reshape(-1, 4)
        area = np.ascontiguousarray(host["area"], np.float32)
        v = TreeView(n, 0, _ptr(arrs["parent"], _i32p), _ptr(arrs["left"], _i32p),
                     _ptr(arrs["right"], _i32p), _ptr(arrs["data"], _i32p), _ptr(bmin),
                     _ptr(bmax), _ptr(area))
        h = _vp()
        check(lib().atr_tree_create_from_host(ctx.h, ctypes.byref(v), ctypes.byref(h)))
        t = cls(ctx, h.value, n)
        if scene is not None:
            check(lib().atr_tree_bind_scene(t.h, scene.h))
        return t

    def download(self) -> dict:
        m = 2 * self.n - 1
        out = {"n": self.n, "root": 0,
               "parent": np.empty(m, np.int32), "left": np.empty(m, np.int32),
               "right": np.empty(m, np.int32), "data": np.empty(m, np.int32),
               "bbmin": np.empty((m, 4), np.float32), "bbmax": np.empty((m, 4), np.float32),
               "area": np.empty(m, np.float32)}
        v = TreeView(self.n, 0, _ptr(out["parent"], _i32p), _ptr(out["left"], _i32p),
                     _ptr(out["right"], _i32p), _ptr(out["data"], _i32p), _ptr(out["bbmin"]),
                     _ptr(out["bbmax"]), _ptr(out["area"]))
        check(lib().atr_tree_download(self.h, ctypes.byref(v)))
        return out

    def upload(self, host: dict):
        """HostToDeviceTree (BVHTreeInstanceManager.cpp:56-66) into this tree."""
        arrs = {k: np.ascontiguousarray(host[k], np.int32) for k in ("parent", "left", "right",
                                                                       "data")}
        bmin = np.ascontiguousarray(host["bbmin"], np.float32).reshape(-1, 4)
        bmax = np.ascontiguousarray(host["bbmax"], np.float32).reshape(-1, 4)
        area = np.ascontiguousarray(host["area"], np.float32)
        v = TreeView(self.n, 0, _ptr(arrs["parent"], _i32p), _ptr(arrs["left"], _i32p),
                     _ptr(arrs["right"], _i32p), _ptr(arrs["data"], _i32p), _ptr(bmin),
                     _ptr(bmax), _ptr(area))
        check(lib().atr_tree_upload(self.h, ctypes.byref(v)))

    def SAH(self, ci: float = 1.2, ct: float = 1.0) -> float:
        out = ctypes.c_float()
        check(lib().atr_tree_sah(self.h, ci, ct, ctypes.byref(out)))
        return out.value

    def sah_device(self, d_out: DeviceBuffer, offset: int = 0, ci: float = 1.2,
                   ct: float = 1.0):
        check(lib().atr_tree_sah_device(self.h, ci, ct, d_out.ptr + 4 * offset))

    def refit(self, scene: Scene | None = None):
        check(lib().atr_refit(self.h, scene.h if scene is not None else None))

    def close(self):
        if self.h:
            lib().atr_tree_destroy(self.h)
            self.h = None


class Collapsed:
    """BVHCollapsedTree plus its traversal layout."""

    def __init__(self, ctx: Context, handle: int, n: int):
        self.ctx = ctx
        self.h = handle
        self.n = n

    def stats(self):
        sah = ctypes.c_float()
        height = ctypes.c_int()
        used = ctypes.c_uint32()
        check(lib().atr_collapsed_stats(self.h, ctypes.byref(sah), ctypes.byref(height),
                                        ctypes.byref(used)))
        return {"sah": sah.value, "height": height.value, "used": used.value}

    def download(self) -> dict:
        m = 2 * self.n - 1
        out = {"n": self.n, "root": 0,
               "parent": np.empty(m, np.int32), "left": np.empty(m, np.int32),
               "right": np.empty(m, np.int32), "data": np.empty(m, np.int32),
               "count": np.empty(m, np.int32),
               "bbmin": np.empty((m, 4), np.float32), "bbmax": np.empty((m, 4), np.float32),
               "tri_index": np.empty(self.n, np.int32)}
        v = CollapsedView(self.n, 0, 0, _ptr(out["parent"], _i32p), _ptr(out["left"], _i32p),
                          _ptr(out["right"], _i32p), _ptr(out["data"], _i32p),
                          _ptr(out["count"], _i32p), _ptr(out["bbmin"]), _ptr(out["bbmax"]),
                          _ptr(out["tri_index"], _i32p))
        check(lib().atr_collapsed_download(self.h, ctypes.byref(v)))
        out["used"] = v.used
        return out

    def layout(self):
        """Builds the traversal layout now (trace builds it on first use)."""
        check(lib().atr_collapsed_layout(self.h))

    def trace(self, d_rays: DeviceBuffer, n: int, d_results: DeviceBuffer, any_hit=False):
        _need(d_rays, n * RAY_BYTES, "rays")
        _need(d_results, n * RESULT_BYTES, "results")
        check(lib().atr_trace(self.h, d_rays.ptr, n, 1 if any_hit else 0, d_results.ptr))

    def compact(self, node_offset_size_div: int = 16) -> CompactLayout:
        """CudaBVH::createCompact on the GPU (Compact2 by default)."""
        out = CompactLayout()
        check(lib().atr_collapsed_compact(self.h, node_offset_size_div, ctypes.byref(out)))
        return out

    def close(self):
        if self.h:
            lib().atr_collapsed_destroy(self.h)
            self.h = None


class CompactLayout(ctypes.Structure):
    """atr_compact_layout: the reference tracer's layout on the device."""
    _fields_ = [("nodes", _vp), ("num_nodes", ctypes.c_uint32), ("tris", _vp),
                ("tri_indices", _vp), ("num_rows", ctypes.c_uint32),
                ("node_offset_size_div", ctypes.c_int)]

    def download(self, ctx: "Context"):
        """(nodes int32[M, 16], Woop rows float32[R, 4], tri_indices int32[R])."""
        nodes = np.empty((self.num_nodes, 16), np.int32)
        woop = np.empty((self.num_rows, 4), np.float32)
        tri = np.empty(self.num_rows, np.int32)
        check(lib().atr_memcpy_d2h(ctx.h, nodes.ctypes.data, self.nodes, nodes.nbytes))
        check(lib().atr_memcpy_d2h(ctx.h, woop.ctypes.data, self.tris, woop.nbytes))
        check(lib().atr_memcpy_d2h(ctx.h, tri.ctypes.data, self.tri_indices, tri.nbytes))
        return nodes, woop, tri


class BatchConfig(ctypes.Structure):
    """atr_batch_config: launch shape of batched builds."""
    _fields_ = [("encode_items", ctypes.c_int32), ("sort_items", ctypes.c_int32),
                ("treelet_block", ctypes.c_int32)]


class Batch:
    """Many independent meshes built as one forest in a single launch sequence
    (optionally a captured HIP graph); each mesh's tree equals the one
    LBVHBuilder(False) + AgglomerativeTreeletOptimizer gives for it alone."""

    def __init__(self, ctx: Context, vertices: np.ndarray, offsets: np.ndarray, bboxes=None):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4)
        self.offsets = np.ascontiguousarray(offsets, np.uint32)
        self.meshes = self.offsets.size - 1
        assert v.shape[0] == 3 * int(self.offsets[-1])
        self.ctx = ctx
        bb = None
        if bboxes is not None:
            bb = np.ascontiguousarray(bboxes, np.float32).reshape(self.meshes, 6)
        h = _vp()
        check(lib().atr_batch_create(ctx.h, v.ctypes.data, _ptr(self.offsets, _u32p),
                                     self.meshes, bb.ctypes.data if bb is not None else None,
                                     ctypes.byref(h)))
        self.h = h.value

    def build(self, treelet: int = 9, iterations: int = 2, ci: float = 1.2, ct: float = 1.0,
              graph: bool = True):
        check(lib().atr_batch_build(self.h, treelet, iterations, ci, ct, 1 if graph else 0))

    def sah(self, ci: float = 1.2, ct: float = 1.0) -> np.ndarray:
        out = np.zeros(self.meshes, np.float32)
        check(lib().atr_batch_sah(self.h, ci, ct, _ptr(out)))
        return out

    def set_config(self, encode_items: int = 16, sort_items: int = 16, treelet_block: int = 128):
        cfg = BatchConfig(encode_items, sort_items, treelet_block)
        check(lib().atr_batch_set_config(self.h, ctypes.byref(cfg)))

    def config(self) -> dict:
        cfg = BatchConfig()
        check(lib().atr_batch_get_config(self.h, ctypes.byref(cfg)))
        return {"encode_items": cfg.encode_items, "sort_items": cfg.sort_items,
                "treelet_block": cfg.treelet_block}

    def trace(self, d_rays: DeviceBuffer, rays_per_mesh: int, d_hits: DeviceBuffer):
        _need(d_rays, self.meshes * rays_per_mesh * RAY_BYTES, "rays")
        _need(d_hits, self.meshes * rays_per_mesh * RESULT_BYTES, "hits")
        check(lib().atr_batch_trace(self.h, d_rays.ptr, rays_per_mesh, d_hits.ptr))

    def hit_adjoint(self, d_rays: DeviceBuffer, d_hits: DeviceBuffer, n: int,
                    d_grad_tuv: DeviceBuffer, d_grad: DeviceBuffer):
        check(lib().atr_batch_hit_adjoint(self.h, d_rays.ptr, d_hits.ptr, n, d_grad_tuv.ptr,
                                          d_grad.ptr))

    def step(self, d_grad: DeviceBuffer, lr: float, max_step: float = 0.0):
        check(lib().atr_batch_step(self.h, d_grad.ptr, lr, max_step))

    def refit(self):
        check(lib().atr_batch_refit(self.h))

    def sah_device(self, d_out: DeviceBuffer, ci: float = 1.2, ct: float = 1.0):
        check(lib().atr_batch_sah_device(self.h, ci, ct, d_out.ptr))

    def rebuild(self, meshes, treelet: int = 9, iterations: int = 2, ci: float = 1.2,
                ct: float = 1.0):
        m = np.ascontiguousarray(meshes, np.uint32)
        check(lib().atr_batch_rebuild(self.h, _ptr(m, _u32p), m.size, treelet, iterations, ci,
                                      ct))

    def extract(self, mesh: int) -> Tree:
        h = _vp()
        check(lib().atr_batch_extract(self.h, mesh, ctypes.byref(h)))
        return Tree(self.ctx, h.value, int(self.offsets[mesh + 1] - self.offsets[mesh]))

    def device_vertices(self) -> int:
        p = _vp()
        check(lib().atr_batch_vertices(self.h, ctypes.byref(p)))
        return p.value

    def close(self):
        if self.h:
            lib().atr_batch_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------- BVHRT mirror
class LBVHBuilder:
    """LBVHBuilder(bool use64BitMortonCodes) (BVHRT-Core/src/LBVHBuilder.h:31)."""

    def __init__(self, use64BitMortonCodes: bool = False):
        self.bits = 63 if use64BitMortonCodes else 30

    def BuildTree(self, scene: Scene) -> Tree:
        h = _vp()
        check(lib().atr_lbvh_build(scene.h, self.bits, ctypes.byref(h)))
        return Tree(scene.ctx, h.value, scene.n)

    def BuildTreeFromBoxes(self, scene: Scene, d_bbmin: DeviceBuffer, d_bbmax: DeviceBuffer,
                           d_tri: DeviceBuffer, m: int) -> Tree:
        """The triangle-splitting build (LBVH.cu:99-161, 320-428, 572-669) over m
        device boxes (float4 rows), each naming a triangle of `scene`."""
        h = _vp()
        check(lib().atr_lbvh_build_boxes(scene.h, d_bbmin.ptr, d_bbmax.ptr, d_tri.ptr, m,
                                         self.bits, ctypes.byref(h)))
        return Tree(scene.ctx, h.value, m)


class TRBVHOptimizer:
    """TRBVHOptimizer(int treeletSize = 7, int iterations = 3) (TRBVHOptimizer.h:27)."""

    def __init__(self, treeletSize: int = 7, iterations: int = 3, ci: float = 1.2,
                 ct: float = 1.0):
        self.treeletSize, self.iterations, self.ci, self.ct = treeletSize, iterations, ci, ct

    def Optimize(self, tree: Tree):
        check(lib().atr_trbvh_optimize(tree.h, self.treeletSize, self.iterations, self.ci,
                                       self.ct))


class AgglomerativeTreeletOptimizer:
    """AgglomerativeTreeletOptimizer(int treeletSize = 9, int iterations = 2)
    (AgglomerativeTreeletOptimizer.h:24)."""

    def __init__(self, treeletSize: int = 9, iterations: int = 2, ci: float = 1.2,
                 ct: float = 1.0):
        self.treeletSize, self.iterations, self.ci, self.ct = treeletSize, iterations, ci, ct

    def Optimize(self, tree: Tree):
        check(lib().atr_atrbvh_optimize(tree.h, self.treeletSize, self.iterations, self.ci,
                                        self.ct))


class BVHTreeCollapser:
    """BVHTreeCollapser::Collapse(tree, &sah, ci, ct) (BVHTreeCollapser.h:39)."""

    def Collapse(self, tree: Tree, ci: float = 1.2, ct: float = 1.0, want_sah: bool = True):
        h = _vp()
        sah = ctypes.c_float()
        check(lib().atr_collapse(tree.h, ci, ct, ctypes.byref(h),
                                 ctypes.byref(sah) if want_sah else None))
        return Collapsed(tree.ctx, h.value, tree.n), (sah.value if want_sah else None)


# ------------------------------------------------------- stages one at a time
def DeviceBuildTree(scene: Scene, d_sorted_keys: DeviceBuffer, d_sort_indices: DeviceBuffer,
                    n: int, morton_bits: int) -> Tree:
    """DeviceBuildTree + DeviceCalculateNodeBoundingBoxes (LBVH.h:80-95,137-139)
    over keys sorted by Context.sort_pairs (uint32 keys for 30 bits, uint64 for 63)."""
    h = _vp()
    check(lib().atr_build_tree(scene.h, d_sorted_keys.ptr, d_sort_indices.ptr, n, morton_bits,
                               ctypes.byref(h)))
    return Tree(scene.ctx, h.value, n)


def DeviceBuildTreeTS(scene: Scene, d_sorted_keys: DeviceBuffer, d_routing: DeviceBuffer,
                      d_bbmin: DeviceBuffer, d_bbmax: DeviceBuffer, d_tri: DeviceBuffer, m: int,
                      morton_bits: int) -> Tree:
    """DeviceBuildTree with routing indices + DeviceCalculateInternalNodeBoundingBoxes
    (LBVH.h:108-125,153-155) over m primitive boxes."""
    h = _vp()
    check(lib().atr_build_tree_boxes(scene.h, d_sorted_keys.ptr, d_routing.ptr, d_bbmin.ptr,
                                     d_bbmax.ptr, d_tri.ptr, m, morton_bits, ctypes.byref(h)))
    return Tree(scene.ctx, h.value, m)


def DeviceTreeletReestructureOptimizer(tree: Tree, treeletSize: int, gamma: int,
                                       ci: float = 1.2, ct: float = 1.0):
    """One TRBVH pass at threshold gamma (TRBVH.h:42-46)."""
    check(lib().atr_trbvh_pass(tree.h, treeletSize, gamma, ci, ct))


def DeviceAgglomerativeTreeletOptimizer(tree: Tree, treeletSize: int, gamma: int,
                                        ci: float = 1.2, ct: float = 1.0):
    """One ATRBVH pass at threshold gamma (AgglomerativeTreelet.h:39-75)."""
    check(lib().atr_atrbvh_pass(tree.h, treeletSize, gamma, ci, ct))


def DeviceCalculateCosts(tree: Tree, ci: float = 1.2, ct: float = 1.0) -> Collapsed:
    """DeviceCopyTree + DeviceCalculateCosts (TreeCollapser.h:19-42): an
    unfinished collapsed tree that DeviceCollapseTree completes."""
    h = _vp()
    check(lib().atr_collapse_costs(tree.h, ci, ct, ctypes.byref(h)))
    return Collapsed(tree.ctx, h.value, tree.n)


def DeviceCollapseTree(tree: Tree, collapsed: Collapsed, want_sah: bool = True):
    """DeviceCollapseTree (TreeCollapser.h:60-62); returns the SAH (or None)."""
    sah = ctypes.c_float()
    check(lib().atr_collapse_lists(tree.h, collapsed.h, ctypes.byref(sah) if want_sah else None))
    return sah.value if want_sah else None
