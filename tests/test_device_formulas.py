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


def _br5(lane: int) -> int:
    return int(f"{lane:05b}"[::-1], 2)


def _min_by_ballot(D, last_row: int) -> int:
    """The selection rule the ATRBVH kernel uses for findMinimum: every lane
    l of a 32-lane half keeps its first strict minimum over entries
    l, l + 32, ...; the winner is the smallest lane value, ties going to the
    lane with the smallest 5-bit-reversed index."""
    size = last_row * (last_row + 1) // 2
    best = []
    for lane in range(32):
        md, mi = np.float32(np.finfo(np.float32).max), 0
        for j in range(lane, size, 32):
            if D[j] < md:
                md, mi = D[j], j
        best.append((md, mi))
    m = min(b[0] for b in best)
    lane = min((l for l in range(32) if best[l][0] == m), key=_br5)
    return best[lane][1]


def test_min_by_ballot_equals_halving_tree_on_ties():
    """The oracle emulates the reference's width-32 __shfl_down halving tree
    (AgglomerativeTreelet.cu:65-99); the kernel's ballot formulation must pick
    the same index, ties included (few distinct values force many ties)."""
    import oracle

    rng = np.random.default_rng(11)
    for trial in range(3000):
        last_row = int(rng.integers(1, 32))
        size = last_row * (last_row + 1) // 2
        levels = int(rng.integers(1, 5))
        D = rng.integers(0, levels, size).astype(np.float32) * np.float32(0.25)
        assert _min_by_ballot(D, last_row) == oracle.find_minimum(D, last_row), trial
