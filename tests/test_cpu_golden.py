"""Regression against the committed golden fixture (tests/golden/corner_8x4.npz,
written by tests/golden/make_golden.py from the oracle)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_golden_corner_8x4(afar):
    import tests.golden.make_golden as mg
    want = np.load(os.path.join(HERE, "golden", "corner_8x4.npz"))
    got = mg.build_fixture()
    for k in ("perm", "S", "C", "stencil", "cellFaces", "shock"):
        assert np.array_equal(got[k], want[k]), k
    for k in ("rbf", "refDistance"):
        assert np.allclose(got[k], want[k], rtol=0, atol=1e-12), k
    for k in ("dts", "U", "W"):
        assert np.allclose(got[k], want[k], rtol=1e-13, atol=1e-15), k
