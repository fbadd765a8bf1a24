"""CPU checks of closed-form index arithmetic that the HIP kernels use instead
of tables: the treelet kernels decode a lower-triangular matrix index p into
(row, col) with a float square-root estimate plus integer corrections
(treelet.hip, trm_row).  The estimate must land within one row for every
p < 496 (n <= 32) even when the hardware square root is off by a few ulp;
the corrections then make it exact.  The reference decodes the same index
with a loop (TriangularMatrix.cuh:34-52)."""
import numpy as np


def _row_exact(p: int) -> int:
    r = 1
    while r * (r + 1) // 2 <= p:
        r += 1
    return r


def _row_device(p: int, sqrt_value: np.float32) -> int:
    r = int((np.float32(1.0) + sqrt_value) * np.float32(0.5))
    if (r * (r - 1)) // 2 > p:
        r -= 1
    if (r * (r + 1)) // 2 <= p:
        r += 1
    return r


def test_trm_row_exact_for_every_index_and_sqrt_error():
    for p in range(496):
        s = np.sqrt(np.float32(8 * p + 1))
        for ulps in range(-3, 4):
            v = s
            step = np.float32(np.inf) if ulps > 0 else np.float32(0.0)
            for _ in range(abs(ulps)):
                v = np.nextafter(v, step)
            r = _row_device(p, np.float32(v))
            assert r == _row_exact(p), (p, ulps)
            c = p - r * (r - 1) // 2
            assert 0 <= c < r


def test_trm_row_matches_row_major_enumeration():
    p = 0
    for r in range(1, 32):
        for c in range(r):
            s = np.sqrt(np.float32(8 * p + 1))
            assert _row_device(p, s) == r
            assert p - r * (r - 1) // 2 == c
            p += 1
    assert p == 496
