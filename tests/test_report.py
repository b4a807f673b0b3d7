"""The machining report with real-time frames, on the oracle: whole runs of the reference
Simulator stepped by sim.step(1/60) at playback speeds 20-120 under its dynamics budget
(golden_report.json).  The frame schedule (mcsim/path.sim_frames) reproduces every sub-step
the reference took; the oracle's sub-step loop, called once per frame with that frame's
wear, reproduces every stamp, every engine run (the frame budget's share), every frame's
taken stats and chatter index, and the final maps bit for bit; mcsim/report.py turns them
into the reference's per-frame chatter states and its LineReport rows (src/sim.js:644-710,
src/report.js:9-33)."""
import numpy as np

import golden_util as G


def _frame_oracle(act, ide, eng, sub, count, R, fl, rw, typ, lut, rc, dyn, fz0, left):
    import oracle
    return oracle.sim_run(act, ide, eng, sub, count, R, fl, rw, 1.0, typ, lut, rc, dynBudget=dyn, fz0=fz0, nframes=1,
                          budget_left=left)


def _maps(s):
    return dict(H=s.H(), err=s.err(), wallD=s.wallD(), markS=s.markS(), markF=s.markF(), lineCut=s.lineCut(),
                lineWall=s.lineWall())


def test_frame_schedule_matches_reference():
    g = G.load("report")
    for case, res in zip(g["cases"], g["results"]):
        sub, n, frames = G.report_schedule(res, case)
        ide = res["ideal"]
        assert n == len(ide["type"])
        assert np.array_equal([sub[k].x for k in range(n)], G.f64(ide["cx"]))
        assert np.array_equal([sub[k].y for k in range(n)], G.f64(ide["cy"]))
        assert np.array_equal([sub[k].z for k in range(n)], G.f64(ide["tipZ"]))
        assert [sub[k].line for k in range(n)] == ide["line"]
        assert list(np.cumsum([c for _, c in frames])) == [f["end"] for f in res["frames"]]


def test_oracle_report_matches_reference_runs():
    import oracle
    g = G.load("report")
    for case, res in zip(g["cases"], g["results"]):
        outs, fouts, rep, stocks, sub, n, frames = G.run_report(
            case, res, oracle.OracleStock, oracle.OracleEngine, _frame_oracle, oracle.tip_lut,
            lambda x: oracle.math_batch("tan", np.array([x]))[0])
        G.check_report_run(case, res, outs, fouts, rep, stocks, sub, n, frames, _maps)
        assert any(s == "chatter" for s in rep.states) or any(s == "marginal" for s in rep.states)
