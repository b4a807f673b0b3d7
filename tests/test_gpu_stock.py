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
        luts = G.luts_from_json(ref.get("luts", []))
        vol, dirty, disk = [], [], []
        qs = np.array([[x, y, r, G.num(a)] for x, y, r, a in case.get("queries", [])]).reshape(-1, 4)
        for part, half in enumerate(G.split_at(case)):
            if part == 1:
                dirty.append(s.dirty())
                disk += [tuple(r) for r in s.max_in_disk(qs)]
                s.clear_dirty()
            if half:
                vol += list(s.stamps_ex(G.stamps_from_json(half), G.opts_from_json(half), None, *luts))
        e, w, lc, lw = s.side()
        assert same_bits(s.H(), G.f32(ref["H"]))
        assert same_bits(e, G.f32(ref["err"])) and same_bits(w, G.f32(ref["wallD"]))
        assert np.array_equal(lc, G.i32(ref["lineCut"])) and np.array_equal(lw, G.i32(ref["lineWall"]))
        ms, mf = s.marks()
        assert same_bits(ms, G.f32(ref["markS"])) and same_bits(mf, G.f32(ref["markF"]))
        np.testing.assert_allclose(vol, np.array(ref["vol"]), rtol=1e-12, atol=1e-12)
        xy = np.array(case["samples"])
        got = s.sample(xy)
        want = np.array([G.num(v) for v in ref["sample"]])
        assert np.array_equal(got, want)
        if "queries" in case:
            dirty.append(s.dirty())
            assert dirty == [G.rect(d) for d in ref["dirty"]]
            disk += [tuple(r) for r in s.max_in_disk(qs)]
            assert disk == [(G.num(h), x, y) for h, x, y in ref["disk"]]
            assert same_bits(s.block_maxima(), G.f32(ref["bmax"]))
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


def marked_profiled_run(mc, n_batches=6):
    """A zigzag of flat stamps with feed marks, interleaved with profiled (bull and ball
    tables) and plain ball stamps, applied in batches on the GPU and in the oracle"""
    import oracle
    from mcsim import workloads as W, path
    from mcsim._abi import StampOpt, MC_TIP_BULL, MC_TIP_BALL
    segs = W.zigzag_segments(120)
    rec, _ = path.schedule(segs, 0.1)
    arr = path.as_stamp_array(rec)
    n = len(arr)
    rng = np.random.default_rng(11)
    tabs = [oracle.tip_lut(MC_TIP_BULL, 2 * arr[0].R, 1.5), oracle.tip_lut(MC_TIP_BALL, 2 * arr[0].R)]
    luts = np.concatenate(tabs)
    off = np.array([0, len(tabs[0])], np.int32)
    ln = np.array([len(t) for t in tabs], np.int32)
    opts = (StampOpt * n)()
    kind = rng.integers(0, 10, n)
    for i in range(n):
        a = arr[i]
        if kind[i] < 6:       # feed marks along the path direction
            j = max(0, i - 1)
            dx, dy = a.cx - arr[j].cx, a.cy - arr[j].cy
            d = np.hypot(dx, dy) or 1.0
            opts[i] = StampOpt(dx / d, dy / d, 0.12, -1, 0)
        elif kind[i] < 8:     # tip profile
            opts[i] = StampOpt(0.0, 0.0, 0.0, int(kind[i] - 6), 0)
        else:
            opts[i] = StampOpt(0.0, 0.0, 0.0, -1, 0)
            if kind[i] == 9:
                arr[i].type = 1
    s = mc.Stock(51.1, 51.1, 20, 0.1)
    o = oracle.OracleStock(51.1, 51.1, 20, 0.1)
    cuts = np.linspace(0, n, n_batches + 1).astype(int)
    rects = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        sub = (type(arr[0]) * (b - a)).from_buffer_copy(arr, a * C_sizeof(arr[0]))
        osub = (StampOpt * (b - a)).from_buffer_copy(opts, a * C_sizeof(opts[0]))
        v = s.stamps_ex(sub, osub, None, luts, off, ln)
        ov = o.stamps_ex(sub, osub, b - a, luts, off, ln)
        np.testing.assert_allclose(v, ov, rtol=1e-11, atol=1e-9)
        rects.append((s.dirty(), o.dirty()))
        s.clear_dirty()
        o.clear_dirty()
    return s, o, rects


def C_sizeof(x):
    import ctypes
    return ctypes.sizeof(x)


def test_marks_and_profiles_match_oracle(mc):
    s, o, rects = marked_profiled_run(mc)
    assert same_bits(s.H(), o.H())
    e, w, lc, lw = s.side()
    assert same_bits(e, o.err()) and same_bits(w, o.wallD())
    assert np.array_equal(lc, o.lineCut()) and np.array_equal(lw, o.lineWall())
    ms, mf = s.marks()
    assert same_bits(ms, o.markS()) and same_bits(mf, o.markF())
    assert np.count_nonzero(mf) > 10000
    for got, want in rects:
        assert got == want
    assert same_bits(s.block_maxima(), o.block_maxima())
    # interference queries over the cut stock: value and reported position
    rng = np.random.default_rng(3)
    q = np.column_stack([rng.uniform(-5, 56, 400), rng.uniform(-5, 56, 400), rng.uniform(-0.5, 15, 400),
                         np.where(rng.uniform(size=400) < 0.6, -np.inf, rng.uniform(-20, 0.5, 400))])
    got = s.max_in_disk(q)
    want = np.array([o.max_in_disk(*row) for row in q])
    assert np.array_equal(got, want)
    assert np.count_nonzero(np.isfinite(got[:, 0])) > 200
    s.close()


def test_dirty_rect_only_for_stamps_that_change(mc):
    """Stock.dirty grows by a stamp's box only when the stamp changed a node: a repeated
    stamp can still rewrite wall distances (its FP64 distance below their f32 copy), a
    shallower one inside the pocket changes nothing"""
    import oracle
    s = mc.Stock(10.0, 10.0, 5.0, 0.1)
    o = oracle.OracleStock(10.0, 10.0, 5.0, 0.1)
    seq = [dict(cx=5.0, cy=5.0, tipZ=-2.0, R=1.0, type="flat", errVal=0.0, line=1),
           dict(cx=5.0, cy=5.0, tipZ=-2.0, R=1.0, type="flat", errVal=0.0, line=2),
           dict(cx=5.0, cy=5.0, tipZ=-1.0, R=0.5, type="flat", errVal=0.0, line=3)]
    seen = []
    for st in seq:
        arr = G.stamps_from_json([st])
        s.stamps(arr)
        o.stamps(arr, 1)
        assert s.dirty() == o.dirty()
        seen.append(s.dirty())
        s.clear_dirty()
        o.clear_dirty()
    assert seen[0][2] > seen[0][0] and seen[2] == [2 ** 31 - 1, 2 ** 31 - 1, -1, -1]
    s.close()


def test_engagement_golden(mc):
    """Simulator.fillEngagement / immersion (src/sim.js:301-348) against the reference's own
    answers on the cut golden stocks: Hb bit for bit, hmax, the immersion arc"""
    g = G.load("engage")
    st = G.load("stock")
    for case, ref in zip(g["cases"], g["results"]):
        sc, sr = st["cases"][case["stock"]], st["results"][case["stock"]]
        s, arr, opt, luts = G.stock_from_case(mc.Stock, sc, sr)
        s.stamps_ex(arr, opt, None, *luts)
        res, Hb = s.engagement(case["queries"])
        assert same_bits(Hb.ravel(), G.f32(ref["Hb"]))
        for r, hm, im, q in zip(res, ref["hmax"], ref["imm"], case["queries"]):
            assert r.hmax == G.num(hm)
            assert r.ap == min(float("inf"), max(0.0, G.num(hm) - q["tipZ"]))
            assert (im is None and r.imm == 0) or (r.imm == 1 and [r.phiSt, r.phiEx] == im)
        s.close()


def test_engagement_matches_oracle_on_zigzag(mc):
    import oracle
    from mcsim import workloads as W, path
    segs = W.zigzag_segments(300)
    rec, _ = path.schedule(segs, 0.1)
    arr = path.as_stamp_array(rec)
    s = mc.Stock(51.1, 51.1, 20, 0.1)
    s.stamps(arr, want_vol=False)
    o = oracle.OracleStock(51.1, 51.1, 20, 0.1)
    o.stamps(arr, len(arr))
    rng = np.random.default_rng(9)
    qs = []
    for _ in range(3000):
        a = rng.uniform(-np.pi, np.pi)
        qs.append(dict(cx=float(rng.uniform(-3, 54)), cy=float(rng.uniform(-3, 54)), ux=float(np.cos(a)),
                       uy=float(np.sin(a)), R=float(rng.uniform(0.5, 6)), tipZ=float(rng.uniform(-4, 0.5)),
                       fluteLen=float(rng.uniform(1, 30))))
    res, Hb = s.engagement(qs)
    n_imm = 0
    for i, q in enumerate(qs):
        h, hb = oracle.fill_engagement(o, q["cx"], q["cy"], q["ux"], q["uy"], q["R"])
        assert same_bits(Hb[i], hb) and res[i].hmax == h
        assert res[i].ap == min(q["fluteLen"], max(0.0, h - q["tipZ"]))
        im = oracle.immersion(hb, q["tipZ"])
        assert (im is None and res[i].imm == 0) or (res[i].imm == 1 and (res[i].phiSt, res[i].phiEx) == im)
        n_imm += im is not None
    assert n_imm > 300
    s.close()


def test_simulator_run_replayed(mc):
    """Whole runs of the reference's simulator: the stamps it issued, replayed on the device
    stock, give its final actual and commanded maps bit for bit and its per-stamp volumes;
    the engagement kernel answers every sample it took, on the stock as it stood then"""
    g = G.load("sim")

    def stamp_fn(s, a, o, luts):
        return s.stamps_ex(a, o, None, *luts)

    def engage_fn(s, cx, cy, ux, uy, R):
        res, _ = s.engagement([dict(cx=cx, cy=cy, ux=ux, uy=uy, R=R, tipZ=0.0)], want_hb=False)
        return res[0].hmax

    for c, r in zip(g["cases"], g["results"]):
        s, ideal, vol, vol_ref, ivol, ivol_ref = G.replay_sim(mc.Stock, stamp_fn, engage_fn, c, r)
        for st, ref in ((s, r["stock"]), (ideal, r["commanded"])):
            e, w, lc, lw = st.side()
            ms, mf = st.marks()
            maps = dict(H=st.H(), err=e, wallD=w, markS=ms, markF=mf, lineCut=lc, lineWall=lw)
            G.check_maps(maps, ref, lambda a: np.asarray(a, np.float32).view(np.int32))
        np.testing.assert_allclose(vol, vol_ref, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(ivol, ivol_ref, rtol=1e-12, atol=1e-12)
        s.close()
        ideal.close()


def test_full_zigzag_batch_matches_reference(mc):
    """the whole heightmap workload (the 5,000-segment zigzag, about 2.4 million flat stamps
    on the 512 x 512 stock) in one batch: every final map is the reference's Stock.stamp
    result byte for byte (SHA-256 of each map against tests/golden/golden_zigzag.json, which
    the reference wrote), and the removed volume equals its within 1e-12 relative"""
    import hashlib
    import json
    import os
    from mcsim import path, workloads as W
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden_zigzag.json")))
    rec, _ = path.schedule(W.zigzag_segments(5000), 0.1)
    assert len(rec) == g["n"]
    s = mc.Stock(51.1, 51.1, 20, 0.1)
    assert (s.nx, s.ny) == (g["nx"], g["ny"])
    vol = s.stamps(path.as_stamp_array(rec))
    err, wallD, lineCut, lineWall = s.side()
    markS, markF = s.marks()
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    got = dict(H=sha(s.H()), err=sha(err), wallD=sha(wallD), lineCut=sha(lineCut), lineWall=sha(lineWall),
               markS=sha(markS), markF=sha(markF))
    assert got == {k: g[k] for k in got}
    assert abs(float(np.sum(vol)) - g["removed"]) <= 1e-12 * g["removed"]
    s.close()


def test_segment_volumes_and_maps_of_line_batch(mc):
    """mc_stamp_batch_lines: the same maps and dirty rectangle as the per-stamp batch, and each
    line's volume the sum of its stamps' volumes (within 1e-12 relative: the device sums in
    another order); a line outside [0, nlines) is refused before any device work"""
    import ctypes as C
    import oracle
    from mcsim import path, workloads as W
    rec, _ = path.schedule(W.zigzag_segments(300), 0.1)
    arr = path.as_stamp_array(rec)
    nl = int(rec["line"].max()) + 1
    a, b = mc.Stock(51.1, 51.1, 20, 0.1), mc.Stock(51.1, 51.1, 20, 0.1)
    vs = a.stamps(arr)
    lv = b.stamps_lines(arr, nl)
    assert same_bits(a.H(), b.H())
    for x, y in zip(a.side(), b.side()):
        assert same_bits(x, y)
    assert a.dirty() == b.dirty()
    want = np.bincount(rec["line"], weights=vs, minlength=nl)
    assert np.all(np.abs(lv - want) <= 1e-12 * np.maximum(1.0, np.abs(want)))
    o = oracle.OracleStock(51.1, 51.1, 20, 0.1)
    ov = o.stamps(arr, len(arr))
    assert same_bits(b.H(), o.H())
    ow = np.bincount(rec["line"], weights=ov, minlength=nl)
    assert np.all(np.abs(lv - ow) <= 1e-11 * np.maximum(1.0, np.abs(ow)))
    assert lv[0] == 0 and np.count_nonzero(lv) > 100
    out = np.zeros(nl)
    assert mc.load().mc_stamp_batch_lines(b.p, arr, len(arr), out.ctypes.data, nl - 1) != 0
