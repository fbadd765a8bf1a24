"""Pins of the oracle's ray-sort keys (RayBuffer::mortonSort, gpu-ray-traversal/
src/rt/ray/RayBuffer.cpp:103-163; findAABBKernel / genMortonKeysKernel,
rt/ray/RayBufferKernels.cu:64-171) by what the definition fixes:

- the rays' box is the min / max of the origins and the finite end points;
- closed forms: an origin at the box's low corner quantises to 0, at its
  high corner to 2^24; a direction along +axis to 2^21 on that axis and 2^20
  on the others, along -axis to 0;
- decoding the key (bit k + 6i of the 192-bit key is bit i of coordinate k)
  gives coordinates within the fp32 rounding of (o - lo)/(hi - lo) * 2^24
  (four steps) and (d/|d| + 1)/2 * 2^21 (two steps), computed here in
  float64;
- ordering by the key is ordering by the interleaved coordinates: an
  interleave with transposed bit positions fails the decode."""
from __future__ import annotations

import numpy as np

import oracle


def _decode(keys: np.ndarray) -> np.ndarray:
    r = keys.shape[1]
    q = np.zeros((6, r), np.uint64)
    for pos in range(192):
        w, b = divmod(pos, 64)
        k, i = pos % 6, pos // 6
        bit = (keys[w] >> np.uint64(b)) & np.uint64(1)
        q[k] |= bit << np.uint64(i)
    return q


def _random_rays(r, seed, tmax=None):
    rng = np.random.default_rng(seed)
    rays = np.zeros((r, 8), np.float32)
    rays[:, 0:3] = rng.uniform(-2.0, 3.0, (r, 3))
    d = rng.normal(size=(r, 3))
    rays[:, 4:7] = d * rng.uniform(0.2, 3.0, (r, 1))  # directions need not be unit
    rays[:, 7] = rng.uniform(0.0, 2.0, r) if tmax is None else tmax
    return rays


def test_ray_box_is_min_max_of_origins_and_finite_end_points():
    rays = _random_rays(5000, 1)
    rays[::7, 7] = np.inf  # end points left out
    box = oracle.ray_box(rays)
    o = rays[:, 0:3]
    e = (o + rays[:, 4:7] * rays[:, 7:8]).astype(np.float32)
    fin = np.all(np.isfinite(e), axis=1)
    pts = np.concatenate([o, e[fin]])
    assert np.array_equal(box[:3], pts.min(0)) and np.array_equal(box[3:], pts.max(0))


def test_ray_key_closed_forms():
    rays = np.zeros((8, 8), np.float32)
    rays[:, 7] = 0.0
    rays[1, 0:3] = [4.0, 2.0, 1.0]  # high corner (box = [0,4]x[0,2]x[0,1])
    axes = np.eye(3, dtype=np.float32)
    rays[0, 4:7] = axes[0]
    rays[1, 4:7] = axes[1]
    for k in range(3):
        rays[2 + k, 4:7] = 5.0 * axes[k]
        rays[5 + k, 4:7] = -axes[k]
    q = _decode(oracle.ray_keys(rays))
    assert np.all(q[0:3, 0] == 0) and np.all(q[0:3, 1] == 1 << 24)
    for k in range(3):
        want = np.full(3, 1 << 20)
        want[k] = 1 << 21
        assert np.array_equal(q[3:6, 2 + k], want)
        want[k] = 0
        assert np.array_equal(q[3:6, 5 + k], want)


def test_ray_key_decodes_to_the_quantised_coordinates():
    rays = _random_rays(20000, 2)
    box = oracle.ray_box(rays).astype(np.float64)
    q = _decode(oracle.ray_keys(rays)).astype(np.float64)
    a = (rays[:, 0:3] - box[:3]) / (box[3:] - box[:3]) * 2.0**24
    d = rays[:, 4:7].astype(np.float64)
    b = (d / np.linalg.norm(d, axis=1, keepdims=True) + 1.0) * 0.5 * 2.0**21
    assert np.all(np.abs(q[0:3].T - np.floor(a)) <= 4.0)
    assert np.all(np.abs(q[3:6].T - np.floor(b)) <= 2.0)
    # the top 45 key bits are zero (origins reach bit 24, directions bit 21)
    keys = oracle.ray_keys(rays)
    assert np.all(keys[2] >> np.uint64(19) == 0)
    # sorting by key sorts by the interleaved coordinates: the first sorted
    # ray has the smallest top coordinate bit pattern
    order = oracle.ray_sort_order(rays)
    top = (q[2, order] >= 2.0**23)
    assert np.all(np.diff(top.astype(int)) >= 0)
