"""The oracle reproduces its frozen outputs (tests/golden/oracle_small.npz,
written by tests/golden/make_fixtures.py after the closed-form pins passed):
rows of two small meshes and twelve surface-point queries on each."""
import os

import numpy as np
import pytest

from tests.golden.make_fixtures import compute

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "oracle_small.npz")


@pytest.fixture(scope="module")
def fixture():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", ["sphere", "disk"])
def test_oracle_matches_frozen_outputs(fixture, name):
    f = {k[len(name) + 1:]: v for k, v in fixture.items() if k.startswith(name + "_")}
    res = compute(f["xyz"], f["tri"], int(f["K"]), f["faces"], f["pts"])
    assert np.array_equal(res["tails"], f["tails"])
    for u, t in enumerate(f["tails"]):
        for k in ("adj", "dist", "src_edge", "dst_edge", "src_x", "dst_x"):
            assert np.array_equal(res[k][u, :t], f[k][u, :t]), (name, u, k)
    assert np.array_equal(res["point_i"], f["point_i"])
    assert np.array_equal(res["point_d"], f["point_d"], equal_nan=True)
