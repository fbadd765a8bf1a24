"""FP64 host checkers for the hit adjoint: the analytic gradient of the depth
loss sum (t - t_target)^2, and central finite differences of that loss (with
the per-ray triangle held fixed, and with rays re-traced against the moved
surface).  Test infrastructure only."""
from __future__ import annotations

import numpy as np


def tuv_fp64(o, d, v0, v1, v2):
    """(t, u, v) of the ray/plane hit in the triangle's barycentric frame
    (Moller-Trumbore without the inside tests), all float64."""
    e1 = v1 - v0
    e2 = v2 - v0
    p = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, p)
    s = o - v0
    q = np.cross(s, e1)
    inv = 1.0 / det
    return (np.einsum("ij,ij->i", e2, q) * inv, np.einsum("ij,ij->i", s, p) * inv,
            np.einsum("ij,ij->i", d, q) * inv)


def hit_fp64(o, d, v0, v1, v2):
    """Two-sided Moller-Trumbore hit distance (inf on a miss), float64."""
    e1 = v1 - v0
    e2 = v2 - v0
    p = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, p)
    s = o - v0
    u = np.einsum("ij,ij->i", s, p)
    q = np.cross(s, e1)
    v = np.einsum("ij,ij->i", d, q)
    t = np.einsum("ij,ij->i", e2, q)
    with np.errstate(divide="ignore", invalid="ignore"):
        pos = (det > 0) & (u >= 0) & (u <= det) & (v >= 0) & (u + v <= det)
        neg = (det < 0) & (u <= 0) & (u >= det) & (v <= 0) & (u + v >= det)
        tt = t / det
    ok = (pos | neg) & (tt > 0)
    return np.where(ok, tt, np.inf)


def rays_od(rays: np.ndarray):
    r = np.asarray(rays, np.float32).astype(np.float64)
    return r[:, 0:3], r[:, 4:7]


def target_depth(unit: np.ndarray, index: np.ndarray, rays: np.ndarray, target_tri: np.ndarray):
    """FP64 depth of every ray on its target triangle (NaN where no target)."""
    o, d = rays_od(rays)
    tt = np.full(len(rays), np.nan)
    m = target_tri >= 0
    tri = index[target_tri[m]]
    tt[m] = tuv_fp64(o[m], d[m], unit[tri[:, 0]], unit[tri[:, 1]], unit[tri[:, 2]])[0]
    return tt


def loss_grad_fp64(verts, index, rays, hit_tri, target_t):
    """Analytic gradient of sum (t - t_target)^2 over rays hit in both
    traces, with t on the ray's hit triangle: dL/dv_i = 2 r b_i n / (n . d)
    (n = e1 x e2, b = barycentrics).  Returns (grad float64[V, 3], loss)."""
    o, d = rays_od(rays)
    m = (hit_tri >= 0) & ~np.isnan(target_t)
    tri = index[hit_tri[m]]
    v0, v1, v2 = verts[tri[:, 0]], verts[tri[:, 1]], verts[tri[:, 2]]
    t, u, v = tuv_fp64(o[m], d[m], v0, v1, v2)
    r = t - target_t[m]
    n = np.cross(v1 - v0, v2 - v0)
    w = (2.0 * r / np.einsum("ij,ij->i", n, d[m]))[:, None] * n
    g = np.zeros_like(verts, dtype=np.float64)
    for k, b in enumerate((1.0 - u - v, u, v)):
        np.add.at(g, tri[:, k], b[:, None] * w)
    return g, float(np.sum(r * r))


def tuv_adjoint_fp64(verts, index, rays, hit_tri, g_tuv):
    """Generic adjoint: sum over hit rays of g . d(t, u, v)/d(vertices), from
    the implicit-function form M x = o - v0 (M = [-d | e1 | e2]):
    dL/dv_i = -b_i M^{-T} g.  Returns float64[V, 3]."""
    o, d = rays_od(rays)
    m = hit_tri >= 0
    tri = index[hit_tri[m]]
    v0, v1, v2 = verts[tri[:, 0]], verts[tri[:, 1]], verts[tri[:, 2]]
    t, u, v = tuv_fp64(o[m], d[m], v0, v1, v2)
    M = np.stack([-d[m], v1 - v0, v2 - v0], axis=2)  # columns
    w = np.linalg.solve(np.transpose(M, (0, 2, 1)), np.asarray(g_tuv, np.float64)[m][:, :, None])
    w = w[:, :, 0]
    g = np.zeros_like(verts, dtype=np.float64)
    for k, b in enumerate((1.0 - u - v, u, v)):
        np.add.at(g, tri[:, k], -b[:, None] * w)
    return g


def fd_tuv(verts, index, rays, hit_tri, g_tuv, sel, h=1e-4, order=4):
    """Central differences of sum_r g_r . (t, u, v)_r with triangles fixed."""
    offs, wts = _stencil(order)
    o, d = rays_od(rays)
    g_tuv = np.asarray(g_tuv, np.float64)
    start, inc = _incidence(index, len(verts))
    out = np.zeros((len(sel), 3))
    for j, k in enumerate(sel):
        tris = inc[start[k]:start[k + 1]]
        rs = np.nonzero(np.isin(hit_tri, tris))[0]
        if rs.size == 0:
            continue
        tri = index[hit_tri[rs]]
        for a in range(3):
            acc = 0.0
            for c_, w_ in zip(offs, wts):
                p = [verts[tri[:, c]].copy() for c in range(3)]
                for c in range(3):
                    p[c][tri[:, c] == k, a] += c_ * h
                x = np.stack(tuv_fp64(o[rs], d[rs], p[0], p[1], p[2]), axis=1)
                acc += w_ * np.sum(x * g_tuv[rs])
            out[j, a] = acc / h
    return out


def _incidence(index: np.ndarray, n_vertices: int):
    """CSR vertex -> incident triangles."""
    flat = index.reshape(-1)
    order = np.argsort(flat, kind="stable")
    counts = np.bincount(flat, minlength=n_vertices)
    start = np.concatenate([[0], np.cumsum(counts)])
    return start, order // 3


def _stencil(order: int):
    """Central-difference offsets and weights (derivative = sum w f(x + c h) / h)."""
    if order == 2:
        return (1.0, -1.0), (0.5, -0.5)
    return (2.0, 1.0, -1.0, -2.0), (-1.0 / 12.0, 8.0 / 12.0, -8.0 / 12.0, 1.0 / 12.0)


def fd_fixed(verts, index, rays, hit_tri, target_t, sel, h=1e-4, order=2):
    """Central differences (2nd or 4th order) of the loss with each ray's
    triangle held fixed."""
    offs, wts = _stencil(order)
    o, d = rays_od(rays)
    start, inc = _incidence(index, len(verts))
    valid = (hit_tri >= 0) & ~np.isnan(target_t)
    idx = np.nonzero(valid)[0]
    order = np.argsort(hit_tri[idx], kind="stable")
    tri_sorted = hit_tri[idx][order]
    bounds = np.searchsorted(tri_sorted, np.arange(len(index) + 1))
    out = np.zeros((len(sel), 3))
    for j, k in enumerate(sel):
        tris = inc[start[k]:start[k + 1]]
        rs = np.concatenate([idx[order[bounds[t]:bounds[t + 1]]] for t in tris])
        if rs.size == 0:
            continue
        tri = index[hit_tri[rs]]
        for a in range(3):
            acc = 0.0
            for c_, w_ in zip(offs, wts):
                p = [verts[tri[:, c]].copy() for c in range(3)]
                for c in range(3):
                    p[c][tri[:, c] == k, a] += c_ * h
                t = tuv_fp64(o[rs], d[rs], p[0], p[1], p[2])[0]
                acc += w_ * np.sum((t - target_t[rs]) ** 2)
            out[j, a] = acc / h
    return out


def fd_retrace(verts, index, rays, hit_tri, target_t, sel, h=1e-4):
    """Central differences of the loss with the rays near each vertex re-traced
    (FP64, closest hit) against its two-ring of triangles after the move."""
    o, d = rays_od(rays)
    start, inc = _incidence(index, len(verts))
    valid = (hit_tri >= 0) & ~np.isnan(target_t)
    idx = np.nonzero(valid)[0]
    order = np.argsort(hit_tri[idx], kind="stable")
    tri_sorted = hit_tri[idx][order]
    bounds = np.searchsorted(tri_sorted, np.arange(len(index) + 1))
    out = np.zeros((len(sel), 3))
    for j, k in enumerate(sel):
        ring1 = inc[start[k]:start[k + 1]]
        vs = np.unique(index[ring1].reshape(-1))
        ring2 = np.unique(np.concatenate([inc[start[x]:start[x + 1]] for x in vs]))
        rs = np.concatenate([idx[order[bounds[t]:bounds[t + 1]]] for t in ring2])
        if rs.size == 0:
            continue
        tri = index[ring2]
        for a in range(3):
            vals = []
            for s in (+h, -h):
                p = [verts[tri[:, c]].copy() for c in range(3)]
                for c in range(3):
                    p[c][tri[:, c] == k, a] += s
                best = np.full(rs.size, np.inf)
                for q in range(len(ring2)):
                    tq = hit_fp64(o[rs], d[rs], np.broadcast_to(p[0][q], (rs.size, 3)),
                                  np.broadcast_to(p[1][q], (rs.size, 3)),
                                  np.broadcast_to(p[2][q], (rs.size, 3)))
                    best = np.minimum(best, tq)
                hitm = np.isfinite(best)
                vals.append(np.sum((best[hitm] - target_t[rs][hitm]) ** 2))
            out[j, a] = (vals[0] - vals[1]) / (2.0 * h)
    return out


def well_conditioned_vertices(verts, index, direction=(0.0, 0.0, 1.0), min_cos=0.3):
    """Vertices whose incident triangles all face the ray direction at an
    angle with |cos| >= min_cos (t is a smooth function of them well beyond
    the finite-difference step; near-grazing triangles make dt/dv blow up)."""
    d = np.asarray(direction, np.float64)
    n = np.cross(verts[index[:, 1]] - verts[index[:, 0]], verts[index[:, 2]] - verts[index[:, 0]])
    c = np.abs(n @ d) / np.linalg.norm(n, axis=1)
    bad = np.zeros(len(verts), bool)
    for k in range(3):
        np.logical_or.at(bad, index[:, k], c < min_cos)
    return np.nonzero(~bad)[0]


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))
