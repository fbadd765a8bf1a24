"""Writes the golden fixtures in this directory.

- trbvh_schedule_ref.json: the TRBVH lane schedule printed by the reference's
  own TRBVHScheduler.cpp (BVHRT-Core/src/TRBVHScheduler.cpp:29-126) compiled
  into oracle/_ref by oracle/Makefile, for treelet sizes 3..7 and warp widths
  4, 8, 32, 64.  Only the printed integers are stored.
- c1_summary.json: the oracle's results on the 65,536-triangle soup workload
  (seed 1): scene box, SAH after LBVH / ATRBVH(9,2) / TRBVH(7,3) / collapse,
  and SHA-256 digests of every tree array.  A regression guard against
  unintended changes of the oracle, not a pin: it holds the oracle's own
  output.

Run from the repository root: python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import atrbvh_amd as ab  # noqa: E402
import oracle  # noqa: E402


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def tree_digests(t) -> dict:
    return {k: digest(getattr(t, k)) for k in ("parent", "left", "right", "data", "bbmin", "bbmax",
                                               "area")}


def f32(x) -> str:
    return np.float32(x).tobytes().hex()


def main():
    sched = {}
    for n in (3, 4, 5, 6, 7):
        for w in (4, 8, 32, 64):
            s = oracle.reference_trbvh_schedule(n, w)
            if s is None:
                raise SystemExit("oracle/_ref is not built (needs the reference sources)")
            sched[f"{n}_{w}"] = s.tolist()
    with open(os.path.join(HERE, "trbvh_schedule_ref.json"), "w") as f:
        json.dump(sched, f)

    v = ab.gen_soup(1, 65536)
    mn, mx = oracle.scene_bbox(v)
    lb = oracle.lbvh(v, mn, mx, 30)
    at = lb.copy()
    oracle.atrbvh(at, 9, 2)
    tr = lb.copy()
    oracle.trbvh(tr, 7, 3)
    c, csah = oracle.collapse(at)
    summary = {
        "bbmin": [f32(x) for x in mn], "bbmax": [f32(x) for x in mx],
        "sah_lbvh": f32(oracle.sah(lb)), "sah_atrbvh": f32(oracle.sah(at)),
        "sah_trbvh": f32(oracle.sah(tr)), "sah_collapsed": f32(csah),
        "lbvh": tree_digests(lb), "atrbvh": tree_digests(at), "trbvh": tree_digests(tr),
        "collapsed_used": int(c.used),
        "collapsed": {k: digest(getattr(c, k)) for k in ("data", "count")},
        "collapsed_tri_index": digest(c.tri_index[: c.used]),
    }
    with open(os.path.join(HERE, "c1_summary.json"), "w") as f:
        json.dump(summary, f, indent=1)

    print("fixtures written")


if __name__ == "__main__":
    main()
