"""Path B on the GPU: heightmap stamps with the reference's sequential semantics.
Final H / err / wallD / lineCut / lineWall bit-exact against the reference's golden
vectors and the oracle; per-stamp volume within 1e-12 relative (summation order)."""
import numpy as np
import pytest

import golden_util as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    from mcsim import lib
    lib.load()
    print("device:", lib.device_info(), flush=True)
    return lib


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.int32), np.asarray(b).view(np.int32))


def test_stamps_golden(mc):
    g = G.load("stock")
    for case, ref in zip(g["cases"], g["results"]):
        s = mc.Stock(case["w"], case["l"], case["h"], case["cell"])
        assert (s.nx, s.ny, s.hx, s.hy) == (ref["nx"], ref["ny"], ref["hx"], ref["hy"])
        if case["wallBand"]:
            s.set_wall_band(case["wallBand"])
        vol = s.stamps(G.stamps_from_json(case["stamps"]))
        e, w, lc, lw = s.side()
        assert same_bits(s.H(), G.f32(ref["H"]))
        assert same_bits(e, G.f32(ref["err"])) and same_bits(w, G.f32(ref["wallD"]))
        assert np.array_equal(lc, G.i32(ref["lineCut"])) and np.array_equal(lw, G.i32(ref["lineWall"]))
        np.testing.assert_allclose(vol, np.array(ref["vol"]), rtol=1e-12, atol=1e-12)
        xy = np.array(case["samples"])
        got = s.sample(xy)
        want = np.array([G.num(v) for v in ref["sample"]])
        assert np.array_equal(got, want)
        s.close()


def test_zigzag_prefix_matches_oracle(mc):
    import oracle
    from mcsim import workloads as W, path
    segs = W.zigzag_segments(400)
    rec, _ = path.schedule(segs, 0.1)
    arr = path.as_stamp_array(rec)
    s = mc.Stock(51.1, 51.1, 20, 0.1)
    vol = s.stamps(arr)
    o = oracle.OracleStock(51.1, 51.1, 20, 0.1)
    ovol = o.stamps(arr, len(arr))
    assert same_bits(s.H(), o.H())
    e, w, lc, lw = s.side()
    assert same_bits(e, o.err()) and same_bits(w, o.wallD())
    assert np.array_equal(lc, o.lineCut()) and np.array_equal(lw, o.lineWall())
    np.testing.assert_allclose(vol, ovol, rtol=1e-11, atol=1e-9)
    # the same stamps split over several calls give the same map
    s2 = mc.Stock(51.1, 51.1, 20, 0.1)
    for a in range(0, len(arr), 7919):
        sub = path.as_stamp_array(rec[a:a + 7919])
        s2.stamps(sub, want_vol=False)
    assert same_bits(s2.H(), o.H())
