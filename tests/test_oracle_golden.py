"""The oracle against the reference's own outputs (tests/golden/*.json, produced by
running the reference's JavaScript under Node): chatterAt verdicts and chatter indices,
timeDomainLimit brackets, per-step forces, Stock.stamp maps, lobes, toolModel."""
import numpy as np

import golden_util as G
import oracle


def test_chatter_cases_bit_exact():
    g = G.load("chatter")
    assert len(g["cases"]) >= 50
    for case, ref in zip(g["cases"], g["results"]):
        t = oracle.chatter_at(G.cell_from_json(case["cell"]), case["ap"])
        assert t.chatter == ref["chatter"]
        assert t.index == ref["index"] and t.amp == ref["amp"] and t.nz == ref["nz"]


def test_time_domain_limits_bit_exact():
    g = G.load("tdl")
    for case, ref in zip(g["cases"], g["results"]):
        o = case["opts"]
        r = oracle.time_domain_limit(G.cell_from_json(case["cell"]), o["aMin"], o["aMax"], o["iters"])
        assert (r.alim, r.lo, r.hi) == (ref["alim"], ref["lo"], G.num(ref["hi"]))


def test_force_traces():
    g = G.load("trace")
    for case, ref in zip(g["cases"], g["results"]):
        fx, fy, x, y = oracle.engine_trace(G.cell_from_json(case["cell"]), case["ap"], case["steps"])
        linear = not case["cell"]["nonlinear"] or case["cell"]["mat"]["mc"] == 0
        for a, key in ((fx, "Fx"), (fy, "Fy"), (x, "x"), (y, "y")):
            b = G.f64(ref[key])
            if linear:
                assert np.array_equal(a.view(np.int64), b.view(np.int64)), key
            else:   # size effect: pow's last bits (see test_oracle_math)
                assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * np.max(np.abs(b))), key


def bits(a):
    return np.asarray(a, np.float32).view(np.int32)


def test_stock_maps_bit_exact():
    g = G.load("stock")
    for case, ref in zip(g["cases"], g["results"]):
        s = oracle.OracleStock(case["w"], case["l"], case["h"], case["cell"])
        assert (s.nx, s.ny, s.hx, s.hy) == (ref["nx"], ref["ny"], ref["hx"], ref["hy"])
        if case["wallBand"]:
            s.set_wall_band(case["wallBand"])
        luts = G.luts_from_json(ref.get("luts", []))
        vol, dirty, disk = [], [], []

        def ask():
            for x, y, r, a in case.get("queries", []):
                disk.append(s.max_in_disk(x, y, r, G.num(a)))

        for part, half in enumerate(G.split_at(case)):
            if part == 1:
                dirty.append(s.dirty())
                ask()
                s.clear_dirty()
            if half:
                vol += list(s.stamps_ex(G.stamps_from_json(half), G.opts_from_json(half), len(half),
                                        *luts))
        assert np.array_equal(bits(s.H()), bits(G.f32(ref["H"])))
        assert np.array_equal(s.err(), G.f32(ref["err"]))
        assert np.array_equal(s.wallD(), G.f32(ref["wallD"]))
        assert np.array_equal(s.lineCut(), G.i32(ref["lineCut"]))
        assert np.array_equal(s.lineWall(), G.i32(ref["lineWall"]))
        assert np.array_equal(vol, np.array(ref["vol"]))
        for (x, y), v in zip(case["samples"], ref["sample"]):
            assert s.sample(x, y) == G.num(v)
        assert np.array_equal(bits(s.markS()), bits(G.f32(ref["markS"])))
        assert np.array_equal(bits(s.markF()), bits(G.f32(ref["markF"])))
        if "queries" not in case:
            continue
        assert np.count_nonzero(s.markF()) > 100     # the marks were exercised
        dirty.append(s.dirty())
        assert dirty == [G.rect(d) for d in ref["dirty"]]
        ask()
        want = [(G.num(h), x, y) for h, x, y in ref["disk"]]
        assert disk == want
        assert sum(d[0] != float("-inf") for d in disk) > len(disk) // 2
        assert np.array_equal(bits(s.block_maxima()), bits(G.f32(ref["bmax"])))
        s.clear_dirty()
        for w in case["walls"]:
            s.mark_wall(*w)
        assert np.array_equal(bits(s.wallErr()), bits(G.f32(ref["wallErr"])))
        assert s.dirty() == G.rect(ref["wallDirty"])
        assert np.count_nonzero(s.wallErr()) > 100


def test_tip_lut_bit_exact():
    g = G.load("tiplut")
    for t, want in zip(g["cases"], g["results"]):
        shape, D, rc, er, dz = G.tip_shape(t)
        got = oracle.tip_lut(shape, D, rc, er, dz)
        assert np.array_equal(bits(got), bits(G.f32(want))), t


def test_stock_profile_reduces_to_ball():
    """A ball's own tipLut table cuts to within the table's radial step of the ball
    formula: the table path and the closed form bound each other."""
    R = 2.0
    lut = oracle.tip_lut(G.MC_TIP_BALL, 2 * R)
    a = oracle.OracleStock(10.0, 10.0, 5.0, 0.05)
    b = oracle.OracleStock(10.0, 10.0, 5.0, 0.05)
    st = G.stamps_from_json([dict(cx=5.0, cy=5.0, tipZ=-2.5, R=R, type="ball", errVal=1.0, line=1)])
    opt = G.opts_from_json([dict(lut=0)])
    a.stamps_ex(st, opt, 1, lut, np.zeros(1, np.int32), np.array([len(lut)], np.int32))
    b.stamps(st, 1)
    Ha, Hb = a.H(), b.H()
    assert np.count_nonzero(Ha < 0) == np.count_nonzero(Hb < 0)
    # slope of a ball of radius 2 below r = 1.9 is < 3.1: one half-step of 0.01 in r
    inner = np.zeros_like(Ha, bool).reshape(a.ny, a.nx)
    yy, xx = np.mgrid[0:a.ny, 0:a.nx]
    inner[(xx * a.hx - 5) ** 2 + (yy * a.hy - 5) ** 2 < 1.9 ** 2] = True
    assert np.max(np.abs(Ha - Hb)[inner.ravel()]) < 0.031


def test_max_in_disk_brute_force():
    """maxInDisk's value is the highest node inside the disk whenever that node is above
    `above` (the block walk only chooses which equal node reports its position)."""
    s = oracle.OracleStock(9.0, 7.0, 5.0, 0.1)
    rng = np.random.default_rng(5)
    st = [dict(cx=float(rng.uniform(0, 9)), cy=float(rng.uniform(0, 7)), tipZ=float(rng.uniform(-4, -0.1)),
               R=float(rng.uniform(0.3, 2)), type="flat", errVal=0.0, line=1) for _ in range(60)]
    s.stamps(G.stamps_from_json(st), len(st))
    H = s.H().reshape(s.ny, s.nx).astype(np.float64)
    yy, xx = np.mgrid[0:s.ny, 0:s.nx]
    for _ in range(300):
        cx, cy, r = rng.uniform(-1, 10), rng.uniform(-1, 8), rng.uniform(0.05, 4)
        inside = (xx * s.hx - cx) ** 2 + (yy * s.hy - cy) ** 2 <= r * r
        h, x, y = s.max_in_disk(cx, cy, r)
        if not inside.any():
            assert h == float("-inf") and (x, y) == (0.0, 0.0)
            continue
        assert h == H[inside].max()
        # the reported point is a node at that height, or the centre of a block inside the disk
        assert (x - cx) ** 2 + (y - cy) ** 2 <= r * r + 1e-9


def test_lobes_and_tool_model():
    import ctypes as C
    from mcsim._abi import Mode
    g = G.load("lobes")
    for case, ref in zip(g["cases"], g["results"]):
        p = case["params"]
        mx = p.get("modesX") or [dict(k=p["k"], fn=p["fn"], zeta=p["zeta"])]
        my = p.get("modesY") or mx
        MX = (Mode * len(mx))(*[Mode(m["k"], m["fn"], m["zeta"]) for m in mx])
        MY = (Mode * len(my))(*[Mode(m["k"], m["fn"], m["zeta"]) for m in my])
        cap = 200000
        rpm, alim = np.zeros(cap), np.zeros(cap)
        ll = np.zeros(p["lobes"], dtype=np.int32)
        L = oracle.lib()
        tot = L.or_stability_lobes(p["N"], p["Ktc"], p["Krc"], p["phiSt"], p["phiEx"], p["lobes"],
                                   p["rpmMin"], p["rpmMax"], MX, len(mx), MY, len(my),
                                   rpm.ctypes.data, alim.ctypes.data, cap, ll.ctypes.data)
        want_r = np.concatenate([G.f64(l["rpm"]) for l in ref["lobes"]]) if ref["lobes"] else np.zeros(0)
        want_a = np.concatenate([G.f64(l["alim"]) for l in ref["lobes"]]) if ref["lobes"] else np.zeros(0)
        assert tot == len(want_r)
        assert np.array_equal(rpm[:tot], want_r) and np.array_equal(alim[:tot], want_a)
        nl = int(np.count_nonzero(ll))
        for S, v in zip(case["speeds"], ref["alimAt"]):
            a = L.or_alim_at(rpm.ctypes.data, alim.ctypes.data, ll.ctypes.data, nl, float(S))
            assert a == G.num(v)
        dc = np.zeros(4)
        L.or_directional_coeffs(p["Krc"] / p["Ktc"], p["phiSt"], p["phiEx"], dc.ctypes.data)
        assert np.array_equal(dc, np.array(ref["dc"]))
    t = G.load("toolmodel")
    for case, ref in zip(t["cases"], t["results"]):
        k, fn = oracle.tool_model(case["D"], case["stickout"], case["kSpindle"], case.get("mHolder", 0.05))
        assert abs(k / ref["k"] - 1) < 1e-14 and abs(fn / ref["fn"] - 1) < 1e-14


def test_engagement_bit_exact():
    """Simulator.fillEngagement / immersion on the cut golden stocks"""
    g = G.load("engage")
    st = G.load("stock")
    for case, ref in zip(g["cases"], g["results"]):
        sc, sr = st["cases"][case["stock"]], st["results"][case["stock"]]
        s, arr, opt, luts = G.stock_from_case(oracle.OracleStock, sc, sr)
        s.stamps_ex(arr, opt, len(arr), *luts)
        Hb_ref = G.f32(ref["Hb"]).reshape(-1, 180)
        for q, hm, im, hb in zip(case["queries"], ref["hmax"], ref["imm"], Hb_ref):
            h, Hb = oracle.fill_engagement(s, q["cx"], q["cy"], q["ux"], q["uy"], q["R"])
            assert np.array_equal(bits(Hb), bits(hb)) and h == G.num(hm), q
            got = oracle.immersion(Hb, q["tipZ"])
            assert (got is None and im is None) or list(got) == im, q
        assert sum(im is not None for im in ref["imm"]) >= 5
        assert sum(G.num(h) > -1e300 for h in ref["hmax"]) > 50


def test_engagement_orientation():
    """With feed along +x the leading bins (sin phi > 0.02) sample ahead of the cutter and
    phi runs from the +y side (phi -> 0) to the -y side (phi -> pi): a stock cut away on
    the +y side shows the cut height in the first bins only."""
    s = oracle.OracleStock(20.0, 20.0, 5.0, 0.1)
    st = G.stamps_from_json([dict(cx=10.0, cy=20.0 + 7.0, tipZ=-2.0, R=9.0, type="flat", errVal=0, line=1)])
    s.stamps(st, 1)
    h, Hb = oracle.fill_engagement(s, 10.0, 10.0, 1.0, 0.0, 3.0)
    phi = (np.arange(180) + 0.5) / 180 * 2 * np.pi
    lead = np.sin(phi) > 0.02
    assert np.all(np.isneginf(Hb[~lead]))
    # sample points: centre + u*cell + r*(sin(phi) u + cos(phi) v), v = (-uy, ux) = (0, 1)
    y = 10.0 + 2.95 * np.cos(phi)
    cut = lead & (np.hypot(10.1 + 2.95 * np.sin(phi) - 10.0, y - 27.0) < 8.9)
    assert np.all(Hb[cut] == -2.0) and np.all(Hb[lead & (y < 17.9)] == 0.0)
    assert h == 0.0
    assert oracle.immersion(Hb, -1.0) is None or oracle.immersion(Hb, -1.0)[0] > 0


def test_simulator_schedule_bit_exact():
    """The sub-step positions of Simulator.step (src/sim.js:363-377,448) from the path the
    reference's own simulator discretised: the commanded stamps it issued"""
    from mcsim import path
    g = G.load("sim")
    for r in g["results"]:
        P = r["path"]
        px, py, pz, _ = path.substeps(G.f64(P["x"]), G.f64(P["y"]), G.f64(P["z"]), P["rapid"], r["hx"])
        I = r["ideal"]
        assert np.array_equal(px, G.f64(I["cx"])) and np.array_equal(py, G.f64(I["cy"]))
        assert np.array_equal(pz, G.f64(I["tipZ"]))


def test_simulator_run_replayed_bit_exact():
    """A whole simulator run: its stamps replayed on the oracle give the final actual and
    commanded maps bit for bit, each stamp's volume, and every engagement sample it took
    on the stock as it stood at that sub-step"""
    g = G.load("sim")

    def stamp_fn(s, a, o, luts):
        return s.stamps_ex(a, o, len(a), *luts)

    def engage_fn(s, cx, cy, ux, uy, R):
        return oracle.fill_engagement(s, cx, cy, ux, uy, R)[0]

    for c, r in zip(g["cases"], g["results"]):
        s, ideal, vol, vol_ref, ivol, ivol_ref = G.replay_sim(oracle.OracleStock, stamp_fn, engage_fn, c, r)
        assert (s.nx, s.ny, s.hx, s.hy) == (r["nx"], r["ny"], r["hx"], r["hy"])
        for st, ref in ((s, r["stock"]), (ideal, r["commanded"])):
            maps = dict(H=st.H(), err=st.err(), wallD=st.wallD(), markS=st.markS(), markF=st.markF(),
                        lineCut=st.lineCut(), lineWall=st.lineWall())
            G.check_maps(maps, ref, bits)
        assert np.array_equal(vol, vol_ref) and np.array_equal(ivol, ivol_ref)
        assert np.count_nonzero(vol) > 100 and len(r["engage"]) > 500


def test_simulator_engine_traces():
    """The engine as the simulator drives it (any feed direction, a stock map, runout, fnMax,
    a velocity set from outside): per-step forces and displacements"""
    g = G.load("enginex")
    for case, ref in zip(g["cases"], g["results"]):
        cell, spec = G.enginex_case(case)
        assert cell.tool.runout == case["tool"]["runout"]
        fx, fy, x, y = oracle.engine_trace_ex(cell, spec, case["steps"])
        for a, key in ((fx, "Fx"), (fy, "Fy"), (x, "x"), (y, "y")):
            b = G.f64(ref[key])
            if not case["nonlinear"]:
                assert np.array_equal(a, b), key
            else:   # size effect: pow's last bits (see test_oracle_math)
                assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * np.max(np.abs(b))), key
        assert np.any(fx != 0) or case["Hb"][45] != "-inf"


def test_engine_object_replays_simulator_calls():
    """the CuttingEngine object across a simulator's whole call sequence (configure, setWear,
    setWork, setEngagement with flat, ball and profile slices, setEndEngagement on plunges and
    ramps, run, takeStats, chatterIndex; src/cutting.js:21-326 driven by src/sim.js:252-257,
    380-445,645-662): state, stats, chip events and rings after every run.  Bit for bit with
    the linear force law; within 1e-12 with the size effect, whose Math.pow differs from the
    restated fdlibm pow in the last bit"""
    g = G.load("englog")
    walls = ends = 0
    for case, res in zip(g["cases"], g["results"]):
        rtol = 0.0 if case.get("nonlinear") is False else 1e-12
        bad, runs = G.replay_englog(oracle.OracleEngine(), res, rtol=rtol)
        assert not bad, (case["tool"], bad[:3])
        assert runs == sum(1 for op in res["ops"] if op[0] == "run") > 100
        walls += sum(1 for op in res["ops"] if op[0] == "work" and len(op) > 1)
        ends += sum(1 for op in res["ops"] if op[0] == "end" and op[1] > 0)
    assert walls > 0 and ends > 0


def test_engine_object_mistakes_are_caught():
    """a planted mistake in the replayed calls (a wall normal turned, the worn edge
    coefficients changed) is detected by the replay"""
    g = G.load("englog")
    case, res = g["cases"][3], g["results"][3]
    assert case.get("nonlinear") is False
    ops = [list(op) for op in res["ops"]]
    k = next(i for i, op in enumerate(ops) if op[0] == "work" and len(op) > 1)
    ops[k][4], ops[k][5] = ops[k][5], ops[k][4]
    bad, _ = G.replay_englog(oracle.OracleEngine(), dict(res, ops=ops))
    assert bad
    ops = [list(op) for op in res["ops"]]
    for op in ops:
        if op[0] == "wear":
            op[1] *= 1.01
    bad, _ = G.replay_englog(oracle.OracleEngine(), dict(res, ops=ops))
    assert bad


def test_run_max_force_pinned_by_engine_stats():
    """the oracle's per-run largest |(Fx, Fy)| (the sub-step loop's maxF) against the
    reference's statistics over the recorded simulator call sequences: after every run
    stats.maxF = Math.max(its value before the run, the run's maximum) (src/cutting.js:280-281),
    and a run that starts from fresh statistics (after configure or takeStats) sets it alone"""
    import oracle

    class Probe:
        def __init__(self):
            self.e = oracle.OracleEngine()
            self.checked = self.fresh = 0

        def __getattr__(self, k):
            return getattr(self.e, k)

        def run(self, n):
            before = self.e.stats().maxF
            self.e.run(n)
            m, after = self.e.run_max_f(), self.e.stats().maxF
            assert after == max(before, m)
            assert m <= after
            if before == 0:
                assert m == after
                self.fresh += 1
            self.checked += 1

    g = G.load("englog")
    fresh = 0
    for case, res in zip(g["cases"], g["results"]):
        if case.get("nonlinear", True) is not False or case["tool"]["type"] == "bull":
            continue
        p = Probe()
        bad, runs = G.replay_englog(p, res, max_runs=300)
        assert not bad and runs > 0 and p.checked == runs
        fresh += p.fresh
    assert fresh > 0


def test_bracket_limit_pinned_by_reference_brackets():
    """the oracle's bracketed search (no reference counterpart) against timeDomainLimit's
    recorded answers (src/stabtd.js:24-32): handed the reference's own final bracket (its
    low end a depth the reference found stable, its high end one it found chattering) and a
    tolerance the bracket already meets, it confirms both ends (two trials) and answers the
    reference's alim; with a tighter tolerance its bracket nests inside the reference's; a
    hint whose low end chatters, or an invalid hint, falls back to the plain search and
    answers exactly what timeDomainLimit answered"""
    g = G.load("tdl")
    seen = 0
    for case, ref in zip(g["cases"], g["results"]):
        o = case["opts"]
        cell = G.cell_from_json(case["cell"])
        lo, hi = ref["lo"], G.num(ref["hi"])
        kw = dict(aMin=o["aMin"], aMax=o["aMax"], iters=o["iters"])
        plain = oracle.time_domain_limit(cell, o["aMin"], o["aMax"], o["iters"])
        if hi == float("inf") or lo == o["aMin"]:
            continue   # open at aMax, or no stable depth tried: the bracket ends are not both verdicts
        r, fb = oracle.bracket_limit(cell, lo, hi, hi - lo, **kw)
        assert fb == 0 and r.trials == 2 and (r.alim, r.lo, r.hi, r.open) == (ref["alim"], lo, hi, 0)
        r, fb = oracle.bracket_limit(cell, lo, hi, (hi - lo) / 4, **kw)
        assert fb == 0 and lo <= r.lo < r.hi <= hi and r.hi - r.lo <= (hi - lo) / 4 and r.trials > 2
        r, fb = oracle.bracket_limit(cell, hi, min(o["aMax"], 2 * hi), 1e-3, **kw)
        assert fb == 1 and (r.alim, r.lo, r.hi) == (ref["alim"], lo, hi) and r.trials == plain.trials + 1
        r, fb = oracle.bracket_limit(cell, hi, lo, 1e-3, **kw)
        assert fb == 1 and (r.alim, r.lo, r.hi) == (ref["alim"], lo, hi) and r.trials == plain.trials
        seen += 1
        if seen == 6:
            break
    assert seen == 6
