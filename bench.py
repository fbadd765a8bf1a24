"""Benchmark: ATRBVH build throughput on one MI355X.

Workload (the largest single-GPU configuration of the workload set):
2,097,152 fp32 triangles from a 64-cluster Gaussian mixture (centres uniform
in [0,1]^3, per-cluster sigma log-uniform in [0.005, 0.1], equilateral
triangles with log-normal edge length, median 0.002, sigma 0.7, random
rotation).  One step = LBVH with 63-bit Morton keys -> ATRBVH (treelet 9,
2 iterations) -> SAH leaf collapse (ci 1.2, ct 1) to the reference's
collapsed tree, with the scene resident in HBM.  value = triangles / step
time (Mtris/s).  The traversal layout built from the collapsed tree (the
reference builds its CudaBVH after the build, on the host) is timed apart.

Also reported: SAH after each stage, the TRBVH (treelet 7, 3 iterations)
build, traversal Mrays/s for 4,194,304 primary rays (2048x2048 pinhole at
(0.5, 0.5, -1) looking +z, 90 degree field of view) and 4,194,304 diffuse
rays from the primary hits, a per-kernel HIP-event breakdown, the roofline
of the dominant streaming kernel, and the CPU oracle timed on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import atrbvh_amd as ab  # noqa: E402

N_TRIS = 2_097_152
SEED = 2
GRID = 2048
EYE = (0.5, 0.5, -1.0)
BASIS = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
HBM_PEAK_GBS = 8000.0
METRIC = "ATRBVH build Mtris/s & SAH at 1M tris; batched builds/s vs plain; Mrays/s"
# streaming kernels priced against HBM bandwidth
STREAMING = ("encode63", "onesweep64", "build_refit63")


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def build_step(scene, treelet="atrbvh", want_sah=False, morton63=True):
    tree = ab.LBVHBuilder(morton63).BuildTree(scene)
    if treelet == "atrbvh":
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(tree)
    elif treelet == "trbvh":
        ab.TRBVHOptimizer(7, 3).Optimize(tree)
    collapsed, sah = ab.BVHTreeCollapser().Collapse(tree, want_sah=want_sah)
    return tree, collapsed, sah


def timed_builds(ctx, scene, n: int, morton63: bool = True, reps: int = 10,
                 warmup: int = 2) -> dict:
    """LBVH + ATRBVH (9, 2) + collapse of `scene` (n triangles): `warmup`
    untimed builds, then `reps` builds timed back to back between stream
    synchronisations; the rate and the collapsed tree's SAH."""
    for _ in range(warmup):
        t, c, _ = build_step(scene, morton63=morton63)
        c.close()
        t.close()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        t, c, _ = build_step(scene, morton63=morton63)
        c.close()
        t.close()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    t, c, sah = build_step(scene, want_sah=True, morton63=morton63)
    c.close()
    t.close()
    return {"triangles": n, "morton_bits": 63 if morton63 else 30, "reps": reps,
            "ms_per_build": round(ms, 4), "Mtris_s": round(n / (ms * 1e-3) / 1e6, 2),
            "sah": round(float(sah), 4)}


PMC_NAMES = {"build_refit63": "k_build<SingleKeys<unsigned long>, TriLeaves>",
             "onesweep64": "k_onesweep<unsigned long, 8, 1024>",
             "encode63": "k_encode<unsigned long, TriLeaves>"}


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE scaled by a plain-copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "gmm_2m_build_pmc_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"][PMC_NAMES[kernel]]
        return round(float(k["hbm_bytes_per_launch"]))
    except Exception:
        return None


def host_info() -> dict:
    """What the CPU numbers were measured on: the host's logical CPU count, its
    CPU model name and the compiler and flags of the CPU code (oracle/Makefile)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cc = "gcc"
    try:
        import subprocess

        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True,
                            timeout=10).stdout.splitlines()[0]
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model,
            "compiler": f"{cc}; C: -O2 -ffp-contract=off -fno-fast-math -std=c11; "
                        "C++ (SBVH): -O2 -ffp-contract=off -fno-fast-math -std=c++17 -pthread"}


def cpu_baseline(sample_tris: int, min_seconds: float = 10.0):
    """The oracle (serial C restatement of the reference build path) on the same
    workload, on one host core, repeated until at least min_seconds of work."""
    import oracle

    v = ab.gen_gmm(SEED, sample_tris)
    mn, mx = oracle.scene_bbox(v)
    runs = []
    while sum(runs) < min_seconds:
        t0 = time.perf_counter()
        t = oracle.lbvh(v, mn, mx, 63)
        oracle.atrbvh(t, 9, 2)
        oracle.collapse(t)
        runs.append(time.perf_counter() - t0)
    dt = float(np.median(runs))
    return {"value": round(sample_tris / dt / 1e6, 4), "unit": "Mtris/s", "cores": 1,
            "kind": "port", **host_info(),
            "sample": f"oracle LBVH63 + ATRBVH(9,2) + collapse of the {sample_tris:,}-triangle "
                      f"GMM scene, single thread, median of {len(runs)} runs "
                      f"({sum(runs):.1f} s total)"}


def cpu_paths(v: np.ndarray, rays: np.ndarray | None, gpu_hits: np.ndarray | None):
    """The reference's CPU path restated (CPU split-BVH builder and closest-hit
    tracer, oracle/sbvh.cpp) on the host cores: the SBVH build of a
    262,144-triangle prefix of the scene on one thread (the reference builder
    is single-threaded), the build of the whole scene and the trace of the
    primary rays on T threads (T = host CPUs, at most 16)."""
    import oracle

    threads = max(1, min(16, os.cpu_count() or 1))
    out = {"threads": threads, "threads_note": "host CPUs, capped at 16 (the box's CPU share)",
           **host_info()}
    sample = 262144
    t0 = time.perf_counter()
    one = oracle.Sbvh(v[: 3 * sample], threads=1)
    dt = time.perf_counter() - t0
    out["sbvh_build_1t"] = {"triangles": sample, "s": round(dt, 3),
                            "Mtris_s": round(sample / dt / 1e6, 4),
                            "sah_bvhrt": round(one.stats["sah_bvhrt"], 4)}
    one.close()
    t0 = time.perf_counter()
    full = oracle.Sbvh(v, threads=threads)
    dt = time.perf_counter() - t0
    out["sbvh_build_mt"] = {"triangles": v.shape[0] // 3, "s": round(dt, 3),
                            "Mtris_s": round(v.shape[0] / 3 / dt / 1e6, 4),
                            "sah": round(full.stats["sah"], 4),
                            "sah_bvhrt": round(full.stats["sah_bvhrt"], 4),
                            "duplicates": full.stats["duplicates"]}
    if rays is not None:
        t0 = time.perf_counter()
        hits = full.trace(rays, threads=threads)
        dt = time.perf_counter() - t0
        out["cpu_trace_mt"] = {"rays": rays.shape[0], "s": round(dt, 3),
                               "Mrays_s": round(rays.shape[0] / dt / 1e6, 3)}
        if gpu_hits is not None:
            out["cpu_trace_mt"]["same_hit_as_gpu"] = round(
                float(np.mean(hits[:, 0] == gpu_hits[:, 0])), 6)
    full.close()
    return out


def batched_leg(ctx, meshes: int, reps: int):
    """Batched builds (4,096 soups of 64..65,536 triangles, ~39M triangles,
    per-mesh seeds 3000 + m, ATRBVH 9/2, 30-bit keys) as one captured launch
    sequence, against the plain per-mesh loop of the same library calls."""
    off = ab.gen_batch_sizes(3, meshes, 64, 65536)
    v = ab.gen_soup_batch(3000, off)
    total = int(off[-1])
    b = ab.Batch(ctx, v, off)
    del v
    from atrbvh_amd.autotune import tune_launch

    tuned = tune_launch(ctx, b, 9, 2, reps=2)
    b.build(9, 2, graph=True)  # capture + first replay with the chosen shape
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.build(9, 2, graph=True)
    ctx.sync()
    graph_ms = (time.perf_counter() - t0) * 1000.0 / reps
    t0 = time.perf_counter()
    b.build(9, 2, graph=False)
    ctx.sync()
    eager_ms = (time.perf_counter() - t0) * 1000.0
    sah_batch = b.sah()
    base = b.device_vertices()
    scenes = [ab.Scene.from_device(ctx, base + 48 * int(off[m]), int(off[m + 1] - off[m]))
              for m in range(meshes)]
    ctx.sync()
    # plain loop: one untimed pass fills the allocator cache with every mesh
    # size, so the timed pass allocates nothing either
    for s in scenes:
        t = ab.LBVHBuilder(False).BuildTree(s)
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
        t.close()
    ctx.sync()
    t0 = time.perf_counter()
    for s in scenes:
        t = ab.LBVHBuilder(False).BuildTree(s)
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
        t.close()
    ctx.sync()
    loop_ms = (time.perf_counter() - t0) * 1000.0
    same = True
    for m, s in enumerate(scenes):
        t = ab.LBVHBuilder(False).BuildTree(s)
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
        same &= np.float32(t.SAH()).tobytes() == sah_batch[m].tobytes()
        t.close()
        s.close()
    b.close()
    cpp = cpp_plain_loop(meshes, tuned["best"], reps=3)
    return {"meshes": meshes, "triangles": total,
            "graph_ms": round(graph_ms, 3), "eager_ms": round(eager_ms, 3),
            "builds_per_s": round(meshes / (graph_ms * 1e-3), 1),
            "Mtris_s": round(total / (graph_ms * 1e-3) / 1e6, 1),
            "plain_loop_ms": round(loop_ms, 3),
            "plain_builds_per_s": round(meshes / (loop_ms * 1e-3), 1),
            "speedup_vs_plain": round(loop_ms / graph_ms, 2),
            "sah_identical": bool(same),
            "mean_sah": round(float(np.mean(sah_batch)), 4),
            "cpp_loop": cpp,
            "autotune": tuned}


def cpp_plain_loop(meshes: int, shape: dict, reps: int = 3):
    """The same batch against the per-mesh loop, both timed from one C++ host
    thread through the C ABI (atrbvh_amd/plain_loop.cpp), so that the loop
    pays no interpreter overhead per call."""
    import subprocess

    exe = os.path.join(ROOT, "atrbvh_amd", "plain_loop")
    if not os.path.exists(exe):
        return {"error": f"{exe} is not built"}
    cmd = [exe, str(meshes), "3", "64", "65536", "3000", str(reps),
           str(shape["encode_items"]), str(shape["sort_items"]), str(shape["treelet_block"])]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    line = (r.stdout.strip().splitlines() or ["{}"])[-1]
    try:
        out = json.loads(line)
    except json.JSONDecodeError:
        out = {"error": f"rc {r.returncode}: {r.stdout[-300:]} {r.stderr[-300:]}"}
    return out


def differentiable_leg(ctx, steps: int = 50, lr: float = 0.005, max_step: float = 0.002):
    """Refit-driven optimisation of the 262,144-triangle perturbed sphere
    toward the unit sphere's depths under 1024 x 1024 orthographic rays: per
    step trace -> depth loss -> hit adjoint (atomic scatter-add over the fixed
    topology) -> vertex step (moves capped at max_step: rays near the
    silhouette graze their triangles and their depth gradients are unbounded)
    + soup gather -> refit -> device SAH check (ATRBVH rebuild when it
    exceeds 1.10x the post-build SAH) -> collapse."""
    import atrbvh_amd.workloads as wl

    m = wl.sphere_mesh(seed=4)
    rays = wl.ortho_rays(1024)
    R = rays.shape[0]
    V = m["vertices"].shape[0]
    s = ctx.scene(wl.soup(m["vertices"], m["index"]))
    st = ctx.scene(wl.soup(m["unit"], m["index"]))
    d_rays = ctx.to_device(rays)
    d_hits = ctx.buffer(R * 16)
    d_target = ctx.buffer(R * 16)
    d_g = ctx.buffer(R * 12)
    d_index = ctx.to_device(m["index"].astype(np.uint32).reshape(-1))
    d_mesh = ctx.to_device(wl.mesh_float4(m["vertices"]))
    d_grad = ctx.to_device(np.zeros((V, 3), np.float32))
    d_loss = ctx.to_device(np.zeros(steps, np.float64))
    d_sah = ctx.buffer(4 * steps)

    def build(scene):
        t = ab.LBVHBuilder(False).BuildTree(scene)
        ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(t)
        return t

    tt = build(st)
    ct, _ = ab.BVHTreeCollapser().Collapse(tt)
    ct.trace(d_rays, R, d_target)
    ct.close()
    tt.close()
    t = build(s)
    sah_built = t.SAH()
    c, _ = ab.BVHTreeCollapser().Collapse(t)
    rebuilds = 0
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        c.trace(d_rays, R, d_hits)
        ctx.depth_loss(d_hits, d_target, R, d_g, d_loss, offset=k)
        ctx.hit_adjoint(s, d_rays, d_hits, R, d_g, d_index, V, d_grad)
        ctx.mesh_step(d_mesh, V, d_grad, lr, d_index, s, max_step)
        t.refit(s)
        t.sah_device(d_sah, offset=k)
        if float(d_sah.download(np.float32, steps)[k]) > 1.10 * sah_built:
            t.close()
            t = build(s)
            sah_built = t.SAH()
            rebuilds += 1
        c.close()
        c, _ = ab.BVHTreeCollapser().Collapse(t)
    ctx.sync()
    dt = time.perf_counter() - t0
    losses = d_loss.download(np.float64, steps)
    sah_refit = t.SAH()
    fresh = build(s)
    sah_fresh = fresh.SAH()
    fresh.close()
    c.close()
    t.close()
    s.close()
    st.close()
    return {"triangles": int(m["index"].shape[0]), "rays": R, "steps": steps,
            "ms_per_step": round(dt * 1000.0 / steps, 3), "steps_per_s": round(steps / dt, 2),
            "loss_first": round(float(losses[0]), 4), "loss_last": round(float(losses[-1]), 4),
            "rebuilds": rebuilds, "sah_refit_final": round(sah_refit, 4),
            "sah_rebuilt_final": round(sah_fresh, 4),
            "sah_drift": round(sah_refit / sah_fresh - 1.0, 5)}


def batch_centroid_rays(v: np.ndarray, off: np.ndarray, per_mesh: int, seed: int = 6) -> np.ndarray:
    """per_mesh rays along +z per mesh, mesh-major, each through the centroid
    of a random triangle of its mesh (a soup of small blobs covers little of
    its own box, so a regular grid would mostly miss)."""
    rng = np.random.default_rng(seed)
    meshes = off.size - 1
    out = np.zeros((meshes * per_mesh, 8), np.float32)
    for m in range(meshes):
        vm = v[3 * int(off[m]): 3 * int(off[m + 1]), :3].reshape(-1, 3, 3)
        cen = vm[rng.integers(0, vm.shape[0], per_mesh)].mean(axis=1)
        r = out[m * per_mesh:(m + 1) * per_mesh]
        r[:, 0] = cen[:, 0]
        r[:, 1] = cen[:, 1]
        r[:, 2] = vm[:, :, 2].min() - 0.1
        r[:, 6] = 1.0
        r[:, 7] = np.inf
    return out


def batched_differentiable_leg(ctx, meshes: int = 64, steps: int = 100, lr: float = 0.002,
                               max_step: float = 0.0005, per_side: int = 512):
    """64 soups of 1,024..524,288 triangles (per-mesh seeds 5000 + m), vertices
    perturbed by N(0, 0.002); every step: batched trace (262,144 rays per mesh
    along +z, each through a random triangle's centroid) -> depth loss against
    the unperturbed meshes on rays that hit the same triangle in both ->
    batched hit adjoint -> soup step (corner moves capped at max_step) ->
    batched refit -> per-mesh SAH, with an ATRBVH rebuild of every mesh whose
    refit SAH exceeds 1.10x its post-build SAH.  Compared with rebuilding
    every mesh every step."""
    off = ab.gen_batch_sizes(5, meshes, 1024, 524288)
    v0 = ab.gen_soup_batch(5000, off)
    vp = v0.copy()
    vp[:, :3] += np.random.default_rng(5).normal(0.0, 0.002, (vp.shape[0], 3)).astype(np.float32)
    rays = batch_centroid_rays(v0, off, per_side * per_side)
    per = per_side * per_side
    R = rays.shape[0]
    n = int(off[-1])
    d_rays = ctx.to_device(rays)
    del rays
    d_target = ctx.buffer(R * 16)
    d_hits = ctx.buffer(R * 16)
    d_g = ctx.buffer(R * 12)
    d_grad = ctx.to_device(np.zeros((3 * n, 3), np.float32))
    d_loss = ctx.to_device(np.zeros(steps, np.float64))
    d_sah = ctx.buffer(4 * meshes)
    bt = ab.Batch(ctx, v0, off)
    bt.build(9, 2, graph=False)
    bt.trace(d_rays, per, d_target)
    ctx.sync()
    bt.close()
    results = {}
    for mode in ("refit", "rebuild_every_step"):
        b = ab.Batch(ctx, vp, off)
        b.build(9, 2, graph=True)
        b.sah_device(d_sah)
        sah_built = d_sah.download(np.float32, meshes).copy()
        rebuilds = 0
        d_loss.upload(np.zeros(steps, np.float64))
        ctx.sync()
        t0 = time.perf_counter()
        for k in range(steps):
            b.trace(d_rays, per, d_hits)
            ctx.depth_loss(d_hits, d_target, R, d_g, d_loss, offset=k, same_triangle=True)
            b.hit_adjoint(d_rays, d_hits, R, d_g, d_grad)
            b.step(d_grad, lr, max_step)
            if mode == "refit":
                b.refit()
                b.sah_device(d_sah)
                sah = d_sah.download(np.float32, meshes)
                flagged = np.nonzero(sah > np.float32(1.10) * sah_built)[0]
                if flagged.size:
                    b.rebuild(flagged, 9, 2)
                    b.sah_device(d_sah)
                    sah_built[flagged] = d_sah.download(np.float32, meshes)[flagged]
                    rebuilds += int(flagged.size)
            else:
                b.build(9, 2, graph=True)
        ctx.sync()
        dt = time.perf_counter() - t0
        b.sah_device(d_sah)
        final_sah = d_sah.download(np.float32, meshes).copy()
        losses = d_loss.download(np.float64, steps)
        b.close()
        results[mode] = {"ms_per_step": round(dt * 1000.0 / steps, 3),
                         "steps_per_s": round(steps / dt, 2), "rebuilds": rebuilds,
                         "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                         "final_sah": final_sah}
    drift = results["refit"]["final_sah"] / results["rebuild_every_step"]["final_sah"] - 1.0
    for r in results.values():
        r["mean_final_sah"] = round(float(np.mean(r.pop("final_sah"))), 4)
    return {"meshes": meshes, "triangles": n, "rays": R, "steps": steps,
            "refit": results["refit"], "rebuild_every_step": results["rebuild_every_step"],
            "sah_drift_mean": round(float(np.mean(drift)), 5),
            "sah_drift_max": round(float(np.max(drift)), 5)}


def copy_calibration(nbytes: int = 256 << 20):
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: 3 copies of {nbytes} bytes")


def scaling_leg(reps: int = 10):
    """LBVH + ATRBVH (9, 2) + collapse at several sizes: the 65,536-triangle
    soup with 30-bit keys and its 256 x 256 pinhole rays (the reference's own
    CPU-runnable case), and the GMM scene family of the headline build with
    63-bit keys at 262,144 to 4,194,304 triangles.  Prints one JSON line of
    rows."""
    ctx = ab.Context(0)
    rows = []
    cases = [("soup", 65_536, False), ("gmm", 262_144, True), ("gmm", 1_048_576, True),
             ("gmm", 2_097_152, True), ("gmm", 4_194_304, True)]
    for kind, n, m63 in cases:
        v = ab.gen_soup(1, n) if kind == "soup" else ab.gen_gmm(SEED, n)
        scene = ctx.scene(v)
        for _ in range(2):
            t, c, _ = build_step(scene, morton63=m63)
            c.close()
            t.close()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            t, c, _ = build_step(scene, morton63=m63)
            c.close()
            t.close()
        ctx.sync()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        t, c, sah = build_step(scene, want_sah=True, morton63=m63)
        row = {"scene": kind, "triangles": n, "morton_bits": 63 if m63 else 30,
               "ms_per_build": round(ms, 4), "Mtris_s": round(n / (ms * 1e-3) / 1e6, 2),
               "sah": round(float(sah), 4)}
        if kind == "soup":
            g = 256
            d_rays = ctx.buffer(g * g * 32)
            d_res = ctx.buffer(g * g * 16)
            ctx.rays_pinhole((0.5, 0.5, -1.0), (1, 0, 0, 0, 1, 0, 0, 0, 1), 1.0, g, g, np.inf,
                             d_rays)
            c.trace(d_rays, g * g, d_res)
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                c.trace(d_rays, g * g, d_res)
            ctx.sync()
            tms = (time.perf_counter() - t0) * 1e3 / reps
            row["trace_ms"] = round(tms, 4)
            row["trace_Mrays_s"] = round(g * g / (tms * 1e-3) / 1e6, 1)
        c.close()
        t.close()
        scene.close()
        log(f"[scaling] {kind} {n:,} ({row['morton_bits']}-bit): {ms:.3f} ms, "
            f"{row['Mtris_s']} Mtris/s, SAH {row['sah']}")
        rows.append(row)
    print(json.dumps({"scaling": rows, "workload": "LBVH + ATRBVH(9,2) + collapse per size",
                      "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-tris", type=int, default=N_TRIS)
    ap.add_argument("--no-trace", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-build leg")
    ap.add_argument("--batch-meshes", type=int, default=4096)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable leg")
    ap.add_argument("--no-batch-adjoint", action="store_true",
                    help="skip the batched differentiable leg")
    ap.add_argument("--scaling", action="store_true",
                    help="only time the build at several sizes (one JSON line of rows)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.scaling:
        scaling_leg()
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the multi-GPU "
              "sharding of this path is design-only.", file=sys.stderr)
        sys.exit(2)

    log(f"[bench] generating {N_TRIS:,}-triangle GMM scene")
    v = ab.gen_gmm(SEED, N_TRIS)
    ctx = ab.Context(0)
    scene = ctx.scene(v)

    # ---------------------------------------------------------------- timed build
    for _ in range(args.warmup):
        t, c, _ = build_step(scene)
        c.close()
        t.close()
    ctx.sync()
    log(f"[bench] warmup done ({args.warmup} steps)")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t, c, _ = build_step(scene)
        c.close()
        t.close()
    ctx.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    mtris = N_TRIS / (ms_per_step * 1e-3) / 1e6
    log(f"[bench] build {ms_per_step:.3f} ms/step, {mtris:.1f} Mtris/s")

    # the same build of the same scene family at 1,048,576 triangles (the size
    # the metric names); a failure here is reported in the line, not fatal
    try:
        v1m = ab.gen_gmm(SEED, 1 << 20)
        s1m = ctx.scene(v1m)
        build_1m = timed_builds(ctx, s1m, 1 << 20, True, reps=max(args.steps, 1))
        s1m.close()
        del v1m
        log(f"[bench] 1M build {build_1m['ms_per_build']} ms, {build_1m['Mtris_s']} Mtris/s")
    except Exception as e:  # noqa: BLE001 -- keep the headline line
        build_1m = {"error": f"{type(e).__name__}: {e}"}

    # ---------------------------------------------------------------- profiled steps
    ctx.profile(True)
    ctx.profile_reset()
    prof_steps = max(3, min(args.steps, 10))
    for _ in range(prof_steps):
        t, c, _ = build_step(scene)
        c.close()
        t.close()
    ctx.sync()
    recs = ctx.profile_records()
    ctx.profile(False)
    kernels = {}
    for name, r in recs.items():
        per = r["launches"] / prof_steps
        kernels[name] = {"ms_per_step": round(r["ms"] / prof_steps, 4),
                         "launches_per_step": per,
                         "avg_launch_us": round(r["ms"] * 1000.0 / max(r["launches"], 1), 3)}
        if r["bytes"] > 0 and r["ms"] > 0:
            kernels[name]["GB_s"] = round(r["bytes"] / (r["ms"] * 1e-3) / 1e9, 1)
    gpu_ms = sum(r["ms"] for r in recs.values()) / prof_steps
    streaming = {k: recs[k] for k in STREAMING if k in recs}
    s_bytes = sum(r["bytes"] for r in streaming.values()) / prof_steps
    s_ms = sum(r["ms"] for r in streaming.values()) / prof_steps
    dom = max(streaming, key=lambda k: streaming[k]["ms"])
    dr = recs[dom]
    per_launch_bytes = dr["bytes"] / dr["launches"]
    per_launch_s = dr["ms"] * 1e-3 / dr["launches"]
    achieved = per_launch_bytes / per_launch_s / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(dom),
                "algorithmic_bytes_per_launch": per_launch_bytes,
                "avg_launch_us": round(per_launch_s * 1e6, 3),
                "streaming_group": {"kernels": list(streaming), "bytes_per_step": s_bytes,
                                    "ms_per_step": round(s_ms, 4),
                                    "GB_s": round(s_bytes / (s_ms * 1e-3) / 1e9, 1),
                                    "frac": round(s_bytes / (s_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                                  4)}}
    log(f"[bench] profiled: gpu {gpu_ms:.3f} ms/step; dominant streaming kernel {dom} "
        f"{achieved:.0f} GB/s")

    # ---------------------------------------------------------------- SAH per stage
    tree = ab.LBVHBuilder(True).BuildTree(scene)
    sah_lbvh = tree.SAH()
    ab.AgglomerativeTreeletOptimizer(9, 2).Optimize(tree)
    sah_atrbvh = tree.SAH()
    collapsed, sah_collapsed = ab.BVHTreeCollapser().Collapse(tree)
    stats = collapsed.stats()
    # the traversal layout (node records and triangle slots), built from the
    # collapsed tree after the build, as CudaBVH is after the reference's
    ctx.sync()
    lt0 = time.perf_counter()
    collapsed.layout()
    ctx.sync()
    tlayout_ms = (time.perf_counter() - lt0) * 1000.0

    # ---------------------------------------------------------------- TRBVH build
    ctx.sync()
    tt0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        t2, c2, _ = build_step(scene, "trbvh")
        c2.close()
        t2.close()
    ctx.sync()
    trbvh_ms = (time.perf_counter() - tt0) * 1000.0 / reps
    t2 = ab.LBVHBuilder(True).BuildTree(scene)
    ab.TRBVHOptimizer(7, 3).Optimize(t2)
    sah_trbvh = t2.SAH()
    c2, sah_trbvh_collapsed = ab.BVHTreeCollapser().Collapse(t2)
    c2.close()
    t2.close()
    log(f"[bench] SAH lbvh {sah_lbvh:.3f} atrbvh {sah_atrbvh:.3f} collapsed {sah_collapsed:.3f} "
        f"trbvh {sah_trbvh:.3f}; trbvh build {trbvh_ms:.2f} ms")

    # ---------------------------------------------------------------- traversal
    trace = None
    prim_rays_host = prim_hits_host = None
    if not args.no_trace:
        r = GRID * GRID
        d_prim = ctx.buffer(r * 32)
        d_res = ctx.buffer(r * 16)
        d_sec = ctx.buffer(r * 32)
        d_res2 = ctx.buffer(r * 16)
        ctx.rays_pinhole(EYE, BASIS, 1.0, GRID, GRID, float("inf"), d_prim)
        collapsed.trace(d_prim, r, d_res)
        ctx.rays_diffuse(scene, d_prim, d_res, r, 7, float("inf"), d_sec)
        ctx.sync()
        res = d_res.download(np.int32, (r, 4))
        prim_rays_host = d_prim.download(np.float32, (r, 8))
        prim_hits_host = res
        n_diffuse = int((res[:, 0] >= 0).sum())
        treps = 5
        ctx.sync()
        w0 = time.perf_counter()
        for _ in range(treps):
            collapsed.trace(d_prim, r, d_res)
        ctx.sync()
        w1 = time.perf_counter()
        for _ in range(treps):
            collapsed.trace(d_sec, r, d_res2)
        ctx.sync()
        w2 = time.perf_counter()
        prim_ms = (w1 - w0) * 1000.0 / treps
        sec_ms = (w2 - w1) * 1000.0 / treps
        trace = {"primary_rays": r, "primary_ms": round(prim_ms, 3),
                 "primary_Mrays_s": round(r / (prim_ms * 1e-3) / 1e6, 1),
                 "primary_hit_rate": round(n_diffuse / r, 4),
                 "diffuse_rays": n_diffuse, "diffuse_ms": round(sec_ms, 3),
                 "diffuse_Mrays_s": round(n_diffuse / (sec_ms * 1e-3) / 1e6, 1)}
        # secondary rays sorted by their 6D Morton key first (the reference's
        # sortSecondary option, Renderer.cpp:353-354, RayBuffer::mortonSort),
        # sort timed apart as the reference times trace kernels alone
        d_sorted = ctx.buffer(r * 32)
        d_s2i = ctx.buffer(r * 4)
        d_res4 = ctx.buffer(r * 16)
        ctx.rays_morton_sort(d_sec, r, d_sorted, d_s2i)
        ctx.sync()
        w0 = time.perf_counter()
        for _ in range(treps):
            ctx.rays_morton_sort(d_sec, r, d_sorted, d_s2i)
        ctx.sync()
        w1 = time.perf_counter()
        for _ in range(treps):
            collapsed.trace(d_sorted, r, d_res4)
        ctx.sync()
        w2 = time.perf_counter()
        sort_ms = (w1 - w0) * 1000.0 / treps
        ssec_ms = (w2 - w1) * 1000.0 / treps
        d_back = ctx.buffer(r * 16)
        ctx.rays_unsort_results(d_res4, d_s2i, r, d_back)
        same = bool(np.array_equal(d_back.download(np.int32, (r, 4))[:, :2],
                                   d_res2.download(np.int32, (r, 4))[:, :2]))
        trace["diffuse_sorted"] = {
            "sort_ms": round(sort_ms, 3), "trace_ms": round(ssec_ms, 3),
            "Mrays_s": round(n_diffuse / (ssec_ms * 1e-3) / 1e6, 1),
            "Mrays_s_with_sort": round(n_diffuse / ((ssec_ms + sort_ms) * 1e-3) / 1e6, 1),
            "hits_equal_unsorted": same}
        for b in (d_sorted, d_s2i, d_res4, d_back):
            b.free()
        # the reference tracer's layout (Compact2, Woop triangles) and its
        # `trace` ABI on the same rays
        ctx.sync()
        c0 = time.perf_counter()
        lay = collapsed.compact(16)
        layout_ms = (time.perf_counter() - c0) * 1000.0
        d_res3 = ctx.buffer(r * 16)
        ctx.trace_compact(r, d_prim, d_res3, lay)
        ctx.sync()
        w0 = time.perf_counter()
        for _ in range(treps):
            ctx.trace_compact(r, d_prim, d_res3, lay)
        ctx.sync()
        w1 = time.perf_counter()
        for _ in range(treps):
            ctx.trace_compact(r, d_sec, d_res3, lay)
        ctx.sync()
        w2 = time.perf_counter()
        cp_ms = (w1 - w0) * 1000.0 / treps
        cs_ms = (w2 - w1) * 1000.0 / treps
        trace["compact_woop"] = {
            "layout_ms": round(layout_ms, 3), "nodes": lay.num_nodes, "rows": lay.num_rows,
            "primary_ms": round(cp_ms, 3), "primary_Mrays_s": round(r / (cp_ms * 1e-3) / 1e6, 1),
            "diffuse_ms": round(cs_ms, 3),
            "diffuse_Mrays_s": round(n_diffuse / (cs_ms * 1e-3) / 1e6, 1)}
        log(f"[bench] trace primary {trace['primary_Mrays_s']} Mrays/s, diffuse "
            f"{trace['diffuse_Mrays_s']} Mrays/s; sorted {trace['diffuse_sorted']}; "
            f"Woop/Compact2 {trace['compact_woop']}")
    collapsed.close()
    tree.close()

    batched = None
    if not args.no_batch:
        log(f"[bench] batched leg: {args.batch_meshes} meshes")
        batched = batched_leg(ctx, args.batch_meshes, reps=3)
        log(f"[bench] batched {batched['graph_ms']} ms ({batched['builds_per_s']} builds/s), "
            f"plain loop {batched['plain_loop_ms']} ms; x{batched['speedup_vs_plain']}, "
            f"SAH identical: {batched['sah_identical']}; C++ loop {batched['cpp_loop']}")

    adjoint = None
    if not args.no_adjoint:
        log("[bench] differentiable refit leg: 50 steps")
        adjoint = differentiable_leg(ctx)
        log(f"[bench] adjoint loop {adjoint['ms_per_step']} ms/step, loss "
            f"{adjoint['loss_first']} -> {adjoint['loss_last']}, rebuilds {adjoint['rebuilds']}")

    batch_adjoint = None
    if not args.no_batch_adjoint:
        log("[bench] batched differentiable leg: 64 meshes, 100 steps")
        batch_adjoint = batched_differentiable_leg(ctx)
        log(f"[bench] batched adjoint: refit path {batch_adjoint['refit']['ms_per_step']} ms/step "
            f"({batch_adjoint['refit']['rebuilds']} rebuilds), rebuild-every-step "
            f"{batch_adjoint['rebuild_every_step']['ms_per_step']} ms/step, SAH drift "
            f"{batch_adjoint['sah_drift_mean']}")

    cpu = None
    cpu_ref = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline on {args.cpu_tris:,} triangles")
        cpu = cpu_baseline(args.cpu_tris)
        log("[bench] CPU SBVH builder / tracer")
        cpu_ref = cpu_paths(v, prim_rays_host, prim_hits_host)

    out = {
        "metric": METRIC,
        "value": round(mtris, 2),
        "unit": "Mtris/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "2,097,152-triangle 64-cluster GMM soup; LBVH 63-bit Morton + "
                               "ATRBVH (treelet 9, 2 iterations) + SAH collapse (ci 1.2, ct 1)",
                   "triangles": N_TRIS, "morton_bits": 63, "treelet": 9, "iterations": 2,
                   "seed": SEED},
        "sah": {"lbvh": round(sah_lbvh, 4), "atrbvh": round(sah_atrbvh, 4),
                "atrbvh_collapsed": round(sah_collapsed, 4), "trbvh": round(sah_trbvh, 4),
                "trbvh_collapsed": round(sah_trbvh_collapsed, 4)},
        "tree_height": stats["height"],
        "traversal_layout": {"ms": round(tlayout_ms, 3),
                             "build_plus_layout_Mtris_s": round(
                                 N_TRIS / ((ms_per_step + tlayout_ms) * 1e-3) / 1e6, 2)},
        "build_1m": build_1m,
        "trbvh_build": {"ms": round(trbvh_ms, 3),
                        "Mtris_s": round(N_TRIS / (trbvh_ms * 1e-3) / 1e6, 2)},
        "trace": trace,
        "batched": batched,
        "differentiable": adjoint,
        "batched_differentiable": batch_adjoint,
        "gpu_ms_per_step": round(gpu_ms, 4),
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_reference_path": cpu_ref,
    }
    print(json.dumps(out), flush=True)
    scene.close()
    ctx.close()


if __name__ == "__main__":
    main()
