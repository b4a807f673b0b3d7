"""The simulator's sub-step loop on the device (mc_sim_run: engagement, engine run and both
stamps per sub-step, src/sim.js:351-449) against whole runs of the reference Simulator with a
rigid workpiece and the linear force law (golden_simx.json; flat, ball, bull-nose and
imported-edge tools): the deflected stamp positions,
their vibration values, the engine's step counts, and every map of the actual and the
commanded stock bit for bit; the removed volume within 1e-12 (its sum order differs).  The
six programs run in one launch, one workgroup each."""
import numpy as np
import pytest

import golden_util as G

pytestmark = pytest.mark.gpu


def _maps(stock):
    H = stock.H()
    err, wallD, lineCut, lineWall = stock.side()
    markS, markF = stock.marks()
    return dict(H=H, err=err, wallD=wallD, lineCut=lineCut, lineWall=lineWall, markS=markS, markF=markF)


def test_sim_pipeline_matches_reference_runs():
    from mcsim import lib
    g = G.load("simx")
    progs, keep = [], []
    for case, res in zip(g["cases"], g["results"]):
        st = case["stock"]
        act = lib.Stock(st["w"], st["l"], st["h"], case["cell"])
        ide = lib.Stock(st["w"], st["l"], st["h"], case["cell"])
        ide.set_wall_band(2.0)
        assert (act.nx, act.ny) == (res["nx"], res["ny"])
        eng = lib.Engine()
        cell, fnMax = G.simx_cell(res)
        info = eng.configure(cell, fnMax)
        _, kte, kre = res["wears"][0]
        eng.set_wear(kte, kre)
        sub, n = G.simx_program(res)
        typ, lut, rc = G.simx_tool(case, res, lib.tip_lut)
        progs.append(dict(actual=act, ideal=ide, engine=eng, sub=sub, nsub=n, R=res["toolR"], fluteLen=res["fluteLen"],
                          rw=0.0, share=info.stepsPerTooth / 4, type=typ, lut=lut, cornerR=rc, edge=G.simx_edge(case)))
        keep.append((case, res, n))
    print("sim pipeline: %d programs, %s sub-steps" % (len(progs), [k[2] for k in keep]), flush=True)
    outs = lib.sim_run(progs)
    for (case, res, n), out, pg in zip(keep, outs, progs):
        sub = pg["sub"]
        a = res["actual"]
        cx = np.array([sub[k].x + out[k].dx for k in range(n)])
        cy = np.array([sub[k].y + out[k].dy for k in range(n)])
        assert np.array_equal(cx, G.f64(a["cx"])) and np.array_equal(cy, G.f64(a["cy"])), case["tool"]
        assert np.array_equal([o.errVal for o in out], G.f64(a["errVal"])), case["tool"]
        runs = dict((k, m) for k, m in res["runs"])
        assert [o.nsteps for o in out] == [runs.get(k, 0) for k in range(n)], case["tool"]
        assert np.array_equal(G.display_arcs(out), G.f64(res["darc"])), case["tool"]
        assert sum(o.cut for o in out) == len(runs)
        for name, stock, ref in (("actual", pg["actual"], res["stock"]), ("commanded", pg["ideal"], res["commanded"])):
            got = _maps(stock)
            for key in ("H", "err", "wallD", "markS", "markF"):
                assert np.array_equal(got[key], G.zf32(ref[key]), equal_nan=True), (case["tool"], name, key)
            for key in ("lineCut", "lineWall"):
                assert np.array_equal(got[key], G.zi32(ref[key])), (case["tool"], name, key)
        removed = sum(o.vol for o in out)
        assert abs(removed - res["stock"]["removed"]) <= 1e-12 * max(1.0, res["stock"]["removed"]), case["tool"]
        assert any(o.maxF > 0 for o in out)


def test_sim_pipeline_matches_oracle_per_substep():
    """every field the pipeline reports per sub-step against the oracle's loop
    (oracle/simloop.c) on the same recorded programs: deflection, axial depth, vibration
    value, the run's largest |(Fx, Fy)| (maxF), the displayed arc, step count and cut flag
    bit for bit; the stamp's removed volume within 1e-12 relative (the device sums the
    stamp's nodes in another order).  The oracle's maxF is pinned on the reference's
    statistics (test_oracle_golden.py), its volumes on the reference's stamps."""
    import oracle
    from mcsim import lib
    g = G.load("simx")
    progs, orc = [], []
    for case, res in zip(g["cases"], g["results"]):
        st = case["stock"]
        cell, fnMax = G.simx_cell(res)
        _, kte, kre = res["wears"][0]
        sub, n = G.simx_program(res)
        act, ide = lib.Stock(st["w"], st["l"], st["h"], case["cell"]), lib.Stock(st["w"], st["l"], st["h"], case["cell"])
        ide.set_wall_band(2.0)
        eng = lib.Engine()
        info = eng.configure(cell, fnMax)
        eng.set_wear(kte, kre)
        typ, lut, rc = G.simx_tool(case, res, lib.tip_lut)
        progs.append(dict(actual=act, ideal=ide, engine=eng, sub=sub, nsub=n, R=res["toolR"], fluteLen=res["fluteLen"],
                          rw=0.0, share=info.stepsPerTooth / 4, type=typ, lut=lut, cornerR=rc, edge=G.simx_edge(case)))
        oa, oi = oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"]), oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"])
        oi.set_wall_band(2.0)
        oe = oracle.OracleEngine()
        oe.configure(cell, fnMax)
        oe.set_wear(kte, kre)
        typ_o, lut_o, _ = G.simx_tool(case, res, oracle.tip_lut)
        orc.append(oracle.sim_run(oa, oi, oe, sub, n, res["toolR"], res["fluteLen"], 0.0, info.stepsPerTooth / 4,
                                  typ_o, lut_o, rc, edge=G.simx_edge(case)))
    outs = lib.sim_run(progs)
    cuts = 0
    for (case, res), out, want in zip(zip(g["cases"], g["results"]), outs, orc):
        assert len(out) == len(want)
        for k, (a, b) in enumerate(zip(out, want)):
            assert (a.dx, a.dy, a.ap, a.errVal, a.maxF, a.phiSt, a.phiEx, a.nsteps, a.cut, a.arc) == \
                   (b.dx, b.dy, b.ap, b.errVal, b.maxF, b.phiSt, b.phiEx, b.nsteps, b.cut, b.arc), (case["tool"], k)
            assert abs(a.vol - b.vol) <= 1e-12 * max(1.0, abs(b.vol)), (case["tool"], k)
            cuts += a.cut
        assert np.array_equal([o.vol for o in want], G.f64(res["actual"]["vol"]))
    assert cuts > 0
