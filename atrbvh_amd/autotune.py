"""Autotuner for batched builds (atr_batch_*).

Two searches, both measured on the caller's own batch with graph replays:

- tune_launch: the launch shape (triangles per encode thread and keys per
  radix-sort thread: 8 or 16 each; threads per workgroup of the treelet
  passes: 64, 128 or 256).  Every shape computes the same trees; the tuner
  checks that the per-mesh SAH of every candidate equals the first
  candidate's bit for bit and keeps the fastest shape on the batch.
- tune_treelet: treelet size and restructuring iterations, which do change
  the trees: the fastest candidate whose mean per-mesh SAH stays within
  `sah_tol` of the reference setting (treelet 9, 2 iterations) is chosen.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

# (encode_items, sort_items, treelet_block); the first is the library's
# default shape (atr_batch_config), the one gain_vs_default is relative to
LAUNCH_SHAPES = tuple(itertools.product((16, 8), (8, 16), (128, 64, 256)))
TREELET_CANDIDATES = ((9, 2), (7, 2), (9, 1), (11, 1), (11, 2), (13, 1))


def _time_builds(ctx, batch, treelet: int, iterations: int, reps: int) -> float:
    batch.build(treelet, iterations, graph=True)  # capture + first replay
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.build(treelet, iterations, graph=True)
    ctx.sync()
    return (time.perf_counter() - t0) * 1000.0 / reps


def tune_launch(ctx, batch, treelet: int = 9, iterations: int = 2, reps: int = 3,
                shapes=LAUNCH_SHAPES) -> dict:
    """Times every launch shape, checks identical per-mesh SAH, keeps the fastest
    shape set on the batch.  Returns {"best": {...}, "table": [...],
    "gain_vs_default": default time / best time} (the default shape is the
    first one)."""
    table = []
    ref_sah = None
    for enc, srt, blk in shapes:
        batch.set_config(encode_items=enc, sort_items=srt, treelet_block=blk)
        ms = _time_builds(ctx, batch, treelet, iterations, reps)
        sah = batch.sah()
        if ref_sah is None:
            ref_sah = sah
        same = bool(np.array_equal(sah.view(np.uint32), ref_sah.view(np.uint32)))
        if not same:
            raise RuntimeError(f"launch shape {enc}/{srt}/{blk} changed the trees")
        table.append({"encode_items": enc, "sort_items": srt, "treelet_block": blk,
                      "ms": round(ms, 3)})
    best = min(table, key=lambda r: r["ms"])
    batch.set_config(encode_items=best["encode_items"], sort_items=best["sort_items"],
                     treelet_block=best["treelet_block"])
    default = table[0]
    return {"best": best, "table": table,
            "gain_vs_default": round(default["ms"] / best["ms"], 4)}


def tune_treelet(ctx, batch, sah_tol: float = 0.01, reps: int = 3,
                 candidates=TREELET_CANDIDATES) -> dict:
    """Fastest (treelet, iterations) whose mean per-mesh SAH is within sah_tol of
    the reference setting's (the first candidate)."""
    table = []
    for n, it in candidates:
        ms = _time_builds(ctx, batch, n, it, reps)
        table.append({"treelet": n, "iterations": it, "ms": round(ms, 3),
                      "mean_sah": float(np.mean(batch.sah()))})
    limit = table[0]["mean_sah"] * (1.0 + sah_tol)
    ok = [r for r in table if r["mean_sah"] <= limit]
    best = min(ok, key=lambda r: r["ms"])
    return {"best": best, "table": table, "sah_limit": limit}
