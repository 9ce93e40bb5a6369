"""The product's host half-edge builder equals the oracle's (CPU only)."""
import numpy as np
import pytest

import oracle as O
import svg_amd as S
from svg_amd import meshgen as mg


@pytest.mark.parametrize("gen", [
    lambda: mg.tetrahedron(),
    lambda: mg.planar_disk(),
    lambda: mg.saddle_cone()[:2],
    lambda: mg.heightfield64(),
    lambda: mg.noisy_sphere(),
    lambda: mg.bumpy_torus(),
])
def test_mesh_builder_matches_oracle(gen):
    xyz, tri = gen()
    a = S.Mesh.from_arrays(xyz, tri).export()
    b = O.Mesh(xyz, tri).export()
    for k in ("twin", "len", "angle", "firstEdge", "boundary"):
        assert np.array_equal(a[k], b[k]), k


def test_mesh_rejects_bad_indices():
    xyz, tri = mg.tetrahedron()
    tri = tri.copy()
    tri[0, 0] = 99
    with pytest.raises(S.SvgError):
        S.Mesh.from_arrays(xyz, tri)


def _fan_of_three_faces_on_one_edge():
    # edge (0, 1) shared by three faces
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, -1, 0], [0.5, 0, 1]], float)
    tri = np.array([[0, 1, 2], [1, 0, 3], [1, 0, 4]], np.int32)
    return xyz, tri


@pytest.mark.parametrize("case,msg", [
    ("repeat", "repeats a vertex"),
    ("three", "shared by more than two faces"),
    ("flip", "oriented oppositely"),
    ("nan", "non-finite coordinate"),
    ("coincident", "edge of length 0"),
])
def test_mesh_refuses_broken_topology(case, msg):
    # meshes on which walks around a vertex would not close (the reference
    # would mis-pair the twins, Mesh.cu:79-89), and meshes whose geometry
    # would put NaN into the window keys, are refused with SVG_E_ARG
    xyz, tri = mg.tetrahedron()
    tri = tri.copy()
    if case == "repeat":
        tri[1] = [tri[1, 0], tri[1, 0], tri[1, 2]]
    elif case == "three":
        xyz, tri = _fan_of_three_faces_on_one_edge()
    elif case == "flip":
        tri[2] = tri[2, ::-1]
    elif case == "nan":
        xyz = xyz.copy()
        xyz[2, 1] = np.nan
    else:                                  # two vertices at one position
        xyz = xyz.copy()
        xyz[1] = xyz[0]
    with pytest.raises(S.SvgError) as e:
        S.Mesh.from_arrays(xyz, tri)
    assert msg in str(e.value) and "error -1" in str(e.value)


def test_obj_roundtrip(tmp_path):
    xyz, tri = mg.noisy_sphere(seed=4, splits4=2, centroid_split=False)
    p = tmp_path / "m.obj"
    with open(p, "w") as f:
        for v in xyz:
            f.write("v %.17g %.17g %.17g\n" % tuple(v))
        f.write("vn 0 0 1\n")
        for t in tri:
            f.write("f %d %d %d\n" % tuple(t + 1))
    m = S.Mesh()
    assert m.LoadFromFile(str(p))
    a = m.export()
    b = S.Mesh.from_arrays(xyz, tri).export()
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert not S.Mesh().LoadFromFile(str(tmp_path / "missing.obj"))


def _write(p, text):
    with open(p, "w") as f:
        f.write(text)


def test_obj_corner_forms_polygons_and_comments(tmp_path):
    # one unit square split two ways: plain triangles (the reference's form)
    # against a quad with i/t/n, i//n and negative corners, comments, vt/vn
    plain = tmp_path / "plain.obj"
    _write(plain, "v 0 0 0\nv 1 0 0\nv 1 1 0.1\nv 0 1 0\nf 1 2 3\nf 1 3 4\n")
    fancy = tmp_path / "fancy.obj"
    _write(fancy, "# a quad\nv 0 0 0\nv 1 0 0\nvt 0 0\nvn 0 0 1\nv 1 1 0.1  # third\nv 0 1 0\n"
                  "f 1/1/1 2//1 -2/3 -1\n")
    a = S.Mesh()
    assert a.LoadFromFile(str(plain))
    b = S.Mesh()
    assert b.LoadFromFile(str(fancy))
    ea, eb = a.export(), b.export()
    for k in ea:
        assert np.array_equal(ea[k], eb[k]), k
    assert (a.vertNum, a.faceNum) == (4, 2)


def test_obj_bad_corner_is_an_error(tmp_path):
    bad = tmp_path / "bad.obj"
    _write(bad, "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 x 3\n")
    assert not S.Mesh().LoadFromFile(str(bad))
    short = tmp_path / "short.obj"
    _write(short, "v 0 0 0\nv 1 0 0\nf 1 2\n")
    assert not S.Mesh().LoadFromFile(str(short))


@pytest.mark.parametrize("field,value,msg", [
    ("adj", 7, "arc target out of range"),
    ("w", -1.0, "arc length is not a finite non-negative number"),
    ("w", float("nan"), "arc length is not a finite non-negative number"),
    ("se", 99, "key-point edge out of range"),
    ("de", -2, "key-point edge out of range"),
    ("sx", float("inf"), "key-point position is not finite"),
])
def test_svg_file_with_bad_records_is_refused(tmp_path, field, value, msg):
    # a .svg file (SVG.cu:208-239 layout: int32 row length, then 40-B records
    # int32 adj, 4 B padding, f64 dist, u32 src edge, u32 dst edge, f64 src x,
    # f64 dst x) whose record is out of range is refused before any upload
    import ctypes as C
    import struct
    xyz, tri = mg.tetrahedron()
    mesh = S.Mesh.from_arrays(xyz, tri)
    rec = dict(adj=1, w=0.5, se=0, de=3, sx=0.1, dx=0.2)
    rec[field] = value
    p = tmp_path / "bad.svg"
    with open(p, "wb") as f:
        f.write(struct.pack("<i", 1))
        f.write(struct.pack("<iidIIdd", rec["adj"], 0, rec["w"], rec["se"] & 0xFFFFFFFF, rec["de"] & 0xFFFFFFFF,
                            rec["sx"], rec["dx"]))
        for _ in range(3):
            f.write(struct.pack("<i", 0))
    g = C.c_void_p()
    rc = S.lib().svg_graph_load(mesh.h, 3, str(p).encode(), C.byref(g))
    assert rc == -4 and msg in S.lib().svg_last_error().decode()
