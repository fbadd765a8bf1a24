"""The OBJ importer and scene converter (atr_mesh_*, host only) against
closed forms on hand-written files: exactly representable numbers come out
exact and others within a few ulp of the correctly rounded value (the
reference's parser accumulates in fp32, String.cpp:440-491); fans, relative
and out-of-range indices, "p/t/n" corner forms, continuation lines, ignored
malformed lines and material grouping follow MeshWavefrontIO.cpp:330-520 and
Scene.cpp:55-77."""
from __future__ import annotations

import os

import numpy as np
import pytest

import atrbvh_amd as ab


def _write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as f:
        f.write(text)
    return p


def _soup(path):
    m = ab.Mesh(path)
    try:
        return m.info(), m.soup()
    finally:
        m.close()


def test_positions_faces_and_corner_forms(tmp_path):
    obj = """# comment
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0.5 0.25 -2e1
v 1 2 3 4
vt 0.5 0.5
vn 0 0 1
f 1 2 3
f 1/1 3/1 4/1
f 1/1/1 2//1 5/1/1 4
f -5 -4 -3
f 1 2 99
f 1 2
f 1 2 x
f 1 2 \\
  3
"""
    info, v = _soup(_write(tmp_path, "a.obj", obj))
    assert info["positions"] == 5  # "v 1 2 3 4" is not a position
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.25, -20.0]], np.float32)
    want = [(0, 1, 2), (0, 2, 3), (0, 1, 4), (0, 4, 3),  # quad fan 1 2 5 4
            (0, 1, 2),  # -5 -4 -3 with five positions
            (0, 1, -1),  # 99 is out of range
            (0, 1, 2)]  # continuation line
    assert info["triangles"] == len(want)
    for t, (a, b, c) in enumerate(want):
        for k, idx in enumerate((a, b, c)):
            q = v[3 * t + k]
            assert q[3] == 0.0
            assert np.array_equal(q[:3], P[idx] if idx >= 0 else np.zeros(3, np.float32))


def test_number_parsing(tmp_path):
    vals = ["0.5", "-1.25", "3", "+2", "1e2", "2.5E-1", "0.1", "3.14159", "-0.007", "123.456",
            "1e-3", "6.02e5", "0.333333333"]
    # positions are read back through faces: one degenerate face per position
    body = "\n".join(["v %s 0 0" % s for s in vals]) + "\n" + "\n".join(
        "f %d %d %d" % (i + 1, i + 1, i + 1) for i in range(len(vals))) + "\n"
    info, v = _soup(_write(tmp_path, "n2.obj", body))
    got = v[0::3, 0]
    for s, g in zip(vals, got):
        want = np.float32(float(s))
        if s in ("0.5", "-1.25", "3", "+2", "1e2", "2.5E-1"):
            assert g == want, s
        ulp = np.spacing(np.abs(want))
        assert abs(float(g) - float(want)) <= 4 * ulp, (s, g, want)


def test_material_groups_order_triangles(tmp_path):
    sub = os.path.join(str(tmp_path), "sc")
    os.makedirs(sub)
    _write(sub, "m.mtl", "newmtl red\nKd 1 0 0\nnewmtl blue\nKd 0 0 1\n")
    obj = """mtllib m.mtl
v 0 0 0
v 1 0 0
v 0 1 0
v 0 0 1
f 1 2 3
usemtl blue
f 1 2 4
usemtl red
f 1 3 4
usemtl blue
f 2 3 4
usemtl unknown
f 4 3 2
"""
    path = _write(sub, "s.obj", obj)
    info, v = _soup(path)
    # groups in order of creation: default (first face), blue, red; the
    # unknown material's face joins the default group
    order = [(0, 1, 2), (3, 2, 1), (0, 1, 3), (1, 2, 3), (0, 2, 3)]
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    assert info["triangles"] == 5
    for t, tri in enumerate(order):
        for k, idx in enumerate(tri):
            assert np.array_equal(v[3 * t + k, :3], P[idx]), (t, k)
    # a file named without a directory reads no material library
    cwd = os.getcwd()
    try:
        os.chdir(sub)
        info2, v2 = _soup("s.obj")
    finally:
        os.chdir(cwd)
    flat = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3), (3, 2, 1)]
    for t, tri in enumerate(flat):
        for k, idx in enumerate(tri):
            assert np.array_equal(v2[3 * t + k, :3], P[idx])


def test_missing_file_is_an_error():
    with pytest.raises(ab.AtrError) as e:
        ab.Mesh("/nonexistent/x.obj")
    assert e.value.status == 1


def test_directory_path_is_an_error(tmp_path):
    """A directory opens as a FILE* on Linux but cannot be read: the loader
    reports it instead of sizing a buffer from a seek that gives no size."""
    with pytest.raises(ab.AtrError) as e:
        ab.Mesh(str(tmp_path))
    assert e.value.status == 1 and "cannot read" in str(e.value)


def test_mtllib_naming_a_directory_is_ignored(tmp_path):
    os.mkdir(os.path.join(str(tmp_path), "lib"))
    obj = "mtllib lib\nusemtl m\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"
    info, v = _soup(_write(tmp_path, "d.obj", obj))
    assert info["triangles"] == 1
    assert np.array_equal(v[:, :3], np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32))


def test_long_indices_wrap_like_int32_accumulation(tmp_path):
    """Indices accumulate digit by digit in 32-bit two's complement
    (String.cpp:384-399), so 4294967297 = 2^32 + 1 reads as 1 and
    99999999999 as 1215752191 (out of range: the origin)."""
    obj = ("v 0 0 0\nv 1 0 0\nv 0 1 0\n"
           "f 4294967297 2 3\n"
           "f 99999999999 2 3\n"
           "f 1 2 3\n")
    info, v = _soup(_write(tmp_path, "w.obj", obj))
    assert info["triangles"] == 3
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    assert np.array_equal(v[0:3, :3], v[6:9, :3])
    assert np.array_equal(v[0:3, :3], P)
    assert np.array_equal(v[3, :3], np.zeros(3, np.float32))
    assert np.array_equal(v[4:6, :3], P[1:])
