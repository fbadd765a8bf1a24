"""CPU model check of the persistent traversal's ray hand-out
(atrbvh_amd/csrc/device.hpp, RayQueue; used by k_trace_fetch in trace.hip and
k_trace_compact in compact.hip).  The arithmetic of span / ray_of / take is
evaluated in 32-bit unsigned integers as on the device, and waves are
interleaved in random orders: every ray below R is handed out exactly once,
positions past R (the last, partial 64-ray block) are dropped, and every
wave stops only after all eight counters ran dry.  At the largest count the
library accepts (2^31 rays per call) no intermediate value wraps."""
from __future__ import annotations

import random

import pytest

M32 = (1 << 32) - 1
COUNTERS = 8


def _span(R, k):
    nb = ((R + 63) & M32) // 64
    return (((nb - 1 - k) & M32) // COUNTERS + 1) * 64 & M32 if nb > k else 0


def _ray_of(s, ck):
    return (((s >> 6) * COUNTERS + ck) * 64 + (s & 63)) & M32


class _Wave:
    def __init__(self, block, R, waves):
        self.R = R
        self.ck = block % COUNTERS
        self.dry = 0
        self.next = self.end = 0
        self.exhausted = False
        self.chunk = min(128, max(16, R // (4 * waves)))

    def take(self, want, counters):
        if self.next == self.end:
            sk = _span(self.R, self.ck)
            base = counters[self.ck]
            counters[self.ck] = (base + self.chunk) & M32
            if base >= sk:
                self.dry += 1
                if self.dry == COUNTERS:
                    self.exhausted = True
                else:
                    self.ck = (self.ck + 1) % COUNTERS
                return 0, 0
            self.next = base
            self.end = min(base + self.chunk, sk)
        t = min(want, self.end - self.next)
        first = self.next
        self.next += t
        return t, first


def _run(R, blocks, waves_per_block, seed):
    rng = random.Random(seed)
    waves = [_Wave(b, R, blocks * waves_per_block)
             for b in range(blocks) for _ in range(waves_per_block)]
    counters = [0] * COUNTERS
    seen = [0] * R
    busy = [[0] * 64 for _ in waves]     # remaining steps of each lane's ray
    live = list(range(len(waves)))
    while live:
        w = rng.choice(live)
        wave, lanes = waves[w], busy[w]
        idle = [l for l in range(64) if lanes[l] == 0]
        if not wave.exhausted and len(idle) >= 16:
            t, first = wave.take(len(idle), counters)
            for rank, lane in enumerate(idle[:t]):
                k = _ray_of(first + rank, wave.ck)
                if k < R:
                    seen[k] += 1
                    lanes[lane] = rng.randint(1, 4)
        if all(x == 0 for x in lanes):
            if wave.exhausted:
                live.remove(w)
            continue
        for l in range(64):
            if lanes[l]:
                lanes[l] -= 1
    return seen


@pytest.mark.parametrize("R,blocks,wpb", [(1, 3, 4), (63, 1, 4), (64, 9, 4), (65, 2, 4),
                                          (1000, 8, 4), (4097, 16, 4), (70001, 24, 4),
                                          (20000, 1, 1)])
def test_every_ray_handed_out_once(R, blocks, wpb):
    for seed in range(3):
        seen = _run(R, blocks, wpb, seed)
        assert all(x == 1 for x in seen), (R, seed, [i for i, x in enumerate(seen) if x != 1][:10])


def test_no_wrap_at_the_largest_count():
    R = 1 << 31
    nb = (R + 63) // 64
    assert R + 63 <= M32
    total = 0
    for k in range(COUNTERS):
        sk = _span(R, k)
        total += sk
        # the last position of counter k maps below 2^32 and, for a full last
        # block, below R
        last = _ray_of(sk - 1, k)
        assert last == ((sk // 64 - 1) * COUNTERS + k) * 64 + 63 <= M32
        assert last < R + 64
        # the counter grows at most one chunk per wave past its span
        assert sk + 128 * 256 * 8 * 4 <= M32
    assert total == nb * 64
