"""CPU checks of bench.py's contract that need no GPU: one GPU only (another
count exits with a message and no JSON line), and the CPU-baseline object
carries the fields a reported baseline needs (rate, unit, threads used, kind,
sample, host CPU count and model, compiler and flags)."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2", "--steps", "1", "--warmup", "0"],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_cpu_baseline_fields():
    sys.path.insert(0, ROOT)
    try:
        import bench
    finally:
        sys.path.remove(ROOT)
    cb = bench.cpu_baseline(4096, min_seconds=0.05)
    assert cb["value"] > 0 and cb["unit"] == "Mtris/s"
    assert cb["cores"] == 1 and cb["kind"] == "port"
    assert "4,096-triangle" in cb["sample"]
    assert cb["nproc"] == os.cpu_count()
    assert cb["cpu_model"] and "-ffp-contract=off" in cb["compiler"]


class _Closable:
    def close(self):
        pass


class _Ctx:
    def __init__(self):
        self.syncs = 0

    def sync(self):
        self.syncs += 1


def test_timed_builds_counts_and_fields(monkeypatch):
    sys.path.insert(0, ROOT)
    try:
        import bench
    finally:
        sys.path.remove(ROOT)
    calls = []

    def fake_step(scene, treelet="atrbvh", want_sah=False, morton63=True):
        calls.append((want_sah, morton63))
        return _Closable(), _Closable(), 131.5 if want_sah else None

    monkeypatch.setattr(bench, "build_step", fake_step)
    ctx = _Ctx()
    r = bench.timed_builds(ctx, object(), 1 << 20, True, reps=4, warmup=2)
    assert len(calls) == 2 + 4 + 1 and calls[-1] == (True, True)
    assert ctx.syncs == 2
    assert set(r) == {"triangles", "morton_bits", "reps", "ms_per_build", "Mtris_s", "sah"}
    assert r["triangles"] == 1 << 20 and r["morton_bits"] == 63 and r["sah"] == 131.5
    assert r["Mtris_s"] > 0
