"""The simulator's sub-step schedule (mcsim/path.sim_substeps, src/sim.js:363-381,389-405)
against whole runs of the reference Simulator with a rigid workpiece: one sub-step per
stamp it issued, the commanded stamp at exactly the sub-step's position, and the feed mark
of every actual stamp equal to the last cutting sub-step's feed per tooth times the flutes."""
import numpy as np

import golden_util as G


def test_sim_substeps_match_reference_runs():
    g = G.load("simx")
    for case, res in zip(g["cases"], g["results"]):
        assert res["frames"] == 1 and len(res["cfgs"]) == 1 and len(res["wears"]) == 1
        sub, n = G.simx_program(res)
        ide = res["ideal"]
        assert n == len(ide["type"]) == len(res["actual"]["type"])
        px = np.array([s.x for s in sub[:n]])
        py = np.array([s.y for s in sub[:n]])
        pz = np.array([s.z for s in sub[:n]])
        assert np.array_equal(px, G.f64(ide["cx"])) and np.array_equal(py, G.f64(ide["cy"]))
        assert np.array_equal(pz, G.f64(ide["tipZ"]))
        assert [s.line for s in sub[:n]] == ide["line"]
        # the actual stamp's radius: no flank wear during the run (rw = 0)
        assert np.all(G.f64(res["actual"]["R"]) == res["toolR"])
        # every engine run happens at a sub-step whose position allows a cut
        runs = {k for k, _ in res["runs"]}
        assert all(0 <= k < n for k in runs)
        # the feed mark f of each actual stamp is the last cutting sub-step's fzEff * N
        f = G.f64(res["actual"]["f"])
        last = 0.0
        for k in range(n):
            if k in runs:
                last = sub[k].fzEff
            assert f[k] == last * res["toolN"], (case["tool"], k)


def test_oracle_sim_loop_matches_reference_runs():
    """the oracle's restatement of the sub-step loop (oracle/simloop.c) against the same
    runs (flat, ball and bull-nose tools): deflected stamp positions, vibration values, step
    counts, per-stamp volumes and every map, bit for bit"""
    import oracle
    g = G.load("simx")
    for case, res in zip(g["cases"], g["results"]):
        st = case["stock"]
        act = oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"])
        ide = oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"])
        ide.set_wall_band(2.0)
        eng = oracle.OracleEngine()
        cell, fnMax = G.simx_cell(res)
        info = eng.configure(cell, fnMax)
        _, kte, kre = res["wears"][0]
        eng.set_wear(kte, kre)
        sub, n = G.simx_program(res)
        typ, lut, rc = G.simx_tool(case, res, oracle.tip_lut)
        out = oracle.sim_run(act, ide, eng, sub, n, res["toolR"], res["fluteLen"], 0.0, info.stepsPerTooth / 4,
                             typ, lut, rc)
        a = res["actual"]
        assert np.array_equal([sub[k].x + out[k].dx for k in range(n)], G.f64(a["cx"])), case["tool"]
        assert np.array_equal([sub[k].y + out[k].dy for k in range(n)], G.f64(a["cy"])), case["tool"]
        assert np.array_equal([o.errVal for o in out], G.f64(a["errVal"]))
        assert np.array_equal([o.vol for o in out], G.f64(a["vol"])), case["tool"]
        runs = dict((k, m) for k, m in res["runs"])
        assert [o.nsteps for o in out] == [runs.get(k, 0) for k in range(n)]
        assert np.array_equal(G.display_arcs(out), G.f64(res["darc"])), case["tool"]
        assert any(o.arc for o in out)
        for name, s_, ref in (("actual", act, res["stock"]), ("commanded", ide, res["commanded"])):
            for key, fn in (("H", s_.H), ("err", s_.err), ("wallD", s_.wallD), ("markS", s_.markS), ("markF", s_.markF)):
                assert np.array_equal(fn(), G.zf32(ref[key]), equal_nan=True), (case["tool"], name, key)
            for key, fn in (("lineCut", s_.lineCut), ("lineWall", s_.lineWall)):
                assert np.array_equal(fn(), G.zi32(ref[key])), (case["tool"], name, key)
