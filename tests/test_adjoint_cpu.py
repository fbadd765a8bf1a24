"""CPU checks of the differentiable workload's inputs and of the FP64
checkers used for the hit adjoint: the sphere mesh is a closed, outward
oriented 2-manifold of 2V - 4 triangles; the analytic FP64 gradients agree
with FP64 central differences (h = 1e-4) on hits from the oracle tracer."""
from __future__ import annotations

import numpy as np

import atrbvh_amd.workloads as wl
import oracle
from adjoint_ref import (fd_fixed, fd_retrace, fd_tuv, loss_grad_fp64, rays_od, rel_err,
                         target_depth, tuv_adjoint_fp64, tuv_fp64, well_conditioned_vertices)


def _trace(v4, rays):
    mn, mx = oracle.scene_bbox(v4)
    t = oracle.lbvh(v4, mn, mx, 30)
    c, _ = oracle.collapse(t)
    return oracle.trace(c, v4, rays)


def test_sphere_mesh_is_closed_outward_manifold():
    m = wl.sphere_mesh(seed=4, n_vertices=2562)
    idx = m["index"]
    assert idx.shape == (2 * 2562 - 4, 3)
    # every directed edge once, its reverse once: closed and consistently wound
    e = np.concatenate([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]])
    key = e[:, 0].astype(np.int64) * 10**6 + e[:, 1]
    rkey = e[:, 1].astype(np.int64) * 10**6 + e[:, 0]
    assert np.unique(key).size == key.size
    assert np.array_equal(np.sort(key), np.sort(rkey))
    u = m["unit"]
    n = np.cross(u[idx[:, 1]] - u[idx[:, 0]], u[idx[:, 2]] - u[idx[:, 0]])
    assert np.all(np.einsum("ij,ij->i", n, u[idx].sum(axis=1)) > 0)
    r = np.linalg.norm(m["vertices"], axis=1)
    assert 0.7 < r.min() < 1.0 < r.max() < 1.3
    assert len(m["modes"]) == 16 and all(1 <= l <= 8 and abs(mm) <= l for l, mm, _ in m["modes"])
    # the Fibonacci lattice is the documented one
    f = wl.fibonacci_sphere(5)
    assert np.allclose(np.linalg.norm(f, axis=1), 1.0) and np.isclose(f[0, 2], 0.8)


def _small_case():
    m = wl.sphere_mesh(seed=4, n_vertices=2562)
    rays = wl.ortho_rays(96)
    hits = _trace(wl.soup(m["vertices"], m["index"]), rays)
    target = _trace(wl.soup(m["unit"], m["index"]), rays)
    tt = target_depth(m["unit"], m["index"], rays, target[:, 0])
    ok = well_conditioned_vertices(m["vertices"], m["index"])
    front = ok[m["vertices"][ok, 2] < -0.3]
    sel = np.random.default_rng(0).choice(front, 64, replace=False)
    return m, rays, hits, tt, sel


def test_fp64_loss_gradient_matches_central_differences():
    m, rays, hits, tt, sel = _small_case()
    g, loss = loss_grad_fp64(m["vertices"], m["index"], rays, hits[:, 0], tt)
    assert loss > 0 and np.count_nonzero(g[sel]) > 100
    # h = 1e-4: the 3-point stencil carries its O(h^2) truncation error, the
    # 5-point stencil none that matters; re-tracing adds the kinks of rays that
    # change triangle within h
    fd2 = fd_fixed(m["vertices"], m["index"], rays, hits[:, 0], tt, sel)
    assert rel_err(g[sel], fd2) < 2e-5
    fd4 = fd_fixed(m["vertices"], m["index"], rays, hits[:, 0], tt, sel, order=4)
    assert rel_err(g[sel], fd4) < 1e-7
    fr = fd_retrace(m["vertices"], m["index"], rays, hits[:, 0], tt, sel)
    assert rel_err(g[sel], fr) < 1e-2


def test_fp64_tuv_adjoint_matches_central_differences():
    m, rays, hits, tt, sel = _small_case()
    g_tuv = np.random.default_rng(1).normal(size=(len(rays), 3))
    g = tuv_adjoint_fp64(m["vertices"], m["index"], rays, hits[:, 0], g_tuv)
    fd = fd_tuv(m["vertices"], m["index"], rays, hits[:, 0], g_tuv, sel)
    assert rel_err(g[sel], fd) < 1e-7
    # the t-only case equals the closed form n b_i / (n . d) used by loss_grad_fp64
    g_t = np.zeros_like(g_tuv)
    valid = (hits[:, 0] >= 0) & ~np.isnan(tt)
    o, d = rays_od(rays)
    tri = m["index"][hits[valid, 0]]
    v = m["vertices"]
    t = tuv_fp64(o[valid], d[valid], v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]])[0]
    g_t[valid, 0] = 2.0 * (t - tt[valid])
    h_ = np.where(valid, hits[:, 0], -1)
    g2 = tuv_adjoint_fp64(v, m["index"], rays, h_, g_t)
    g3, _ = loss_grad_fp64(v, m["index"], rays, hits[:, 0], tt)
    assert rel_err(g2, g3) < 1e-12
