"""tipLut (src/assembly.js:38-54) in the library's host code: bit-identical to the
reference's tables on the golden tools, and its argument checks."""
import numpy as np
import pytest

import golden_util as G


def test_tip_lut_matches_reference():
    from mcsim import lib
    g = G.load("tiplut")
    for t, want in zip(g["cases"], g["results"]):
        shape, D, rc, er, dz = G.tip_shape(t)
        got = lib.tip_lut(shape, D, rc, er, dz)
        assert np.array_equal(got.view(np.int32), G.f32(want).view(np.int32)), t


def test_tip_lut_shapes():
    from mcsim import lib
    ball = lib.tip_lut(G.MC_TIP_BALL, 6.0)
    assert len(ball) == 151 and ball[0] == 0 and ball[-1] == np.float32(3.0)
    assert np.all(np.diff(ball) >= 0)
    flat = lib.tip_lut(G.MC_TIP_FLAT, 6.0)
    assert not flat.any()
    # a corner radius as large as the cutter is a ball
    assert np.array_equal(lib.tip_lut(G.MC_TIP_BULL, 6.0, 3.0), ball)
    with pytest.raises(lib.McError):
        lib.tip_lut(7, 6.0)
    with pytest.raises(lib.McError):
        lib.tip_lut(G.MC_TIP_EDGE, 6.0)   # no envelope given
