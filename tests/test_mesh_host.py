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
