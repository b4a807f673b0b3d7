"""The machining report with real-time frames on the device: the reference Simulator's runs
stepped by sim.step(1/60) at playback speeds 20-120 under its dynamics budget
(golden_report.json), replayed one mc_sim_run call per frame.  Every stamp, engine run
(the frame budget's share computed on the device), every frame's taken stats and chatter
index and the final maps are the reference's bit for bit; the per-frame chatter states and
the LineReport rows built from them too (src/sim.js:644-710, src/report.js:9-33).  Then the
same runs with all their frames in one launch (a fixed wear) against the oracle."""
import numpy as np
import pytest

import golden_util as G

pytestmark = pytest.mark.gpu


def _frame_gpu(act, ide, eng, sub, count, R, fl, rw, typ, lut, rc, dyn, fz0, left):
    from mcsim import lib
    outs, frs = lib.sim_run([dict(actual=act, ideal=ide, engine=eng, sub=sub, nsub=count, R=R, fluteLen=fl, rw=rw,
                                  type=typ, lut=lut, cornerR=rc, dynBudget=dyn, fz=fz0, budget_left=left)],
                            frames=True)
    return outs[0], frs[0]


def _maps(s):
    H = s.H()
    err, wallD, lineCut, lineWall = s.side()
    markS, markF = s.marks()
    return dict(H=H, err=err, wallD=wallD, markS=markS, markF=markF, lineCut=lineCut, lineWall=lineWall)


def test_report_runs_frame_by_frame_match_reference():
    from mcsim import lib
    print("device:", lib.device_info(), flush=True)
    g = G.load("report")
    for case, res in zip(g["cases"], g["results"]):
        print("report run: speed %g, %d frames" % (res["speed"], len(res["frames"])), flush=True)
        outs, fouts, rep, stocks, sub, n, frames = G.run_report(
            case, res, lib.Stock, lib.Engine, _frame_gpu, lib.tip_lut, lambda x: lib.js_math("tan", x)[0])
        G.check_report_run(case, res, outs, fouts, rep, stocks, sub, n, frames, _maps)


def test_frames_in_one_launch_match_oracle():
    """all frames of every run in one launch (wear fixed at its first value): sub-steps,
    frame stats and chatter indices identical to the oracle's loop over the same frames"""
    import oracle
    from mcsim import lib
    g = G.load("report")
    progs, refs = [], []
    for case, res in zip(g["cases"], g["results"]):
        st = case["stock"]
        sub, n, frames = G.report_schedule(res, case)
        typ, lut, rc = G.simx_tool(case, res, lib.tip_lut)
        cell, fnMax = G.simx_cell(res)
        _, kte, kre = res["wears"][0]
        nf = sum(1 for _, c in frames if c)
        for mk_s, mk_e, dst in ((lib.Stock, lib.Engine, progs), (oracle.OracleStock, oracle.OracleEngine, refs)):
            a, i = mk_s(st["w"], st["l"], st["h"], case["cell"]), mk_s(st["w"], st["l"], st["h"], case["cell"])
            i.set_wall_band(2.0)
            e = mk_e()
            e.configure(cell, fnMax)
            e.set_wear(kte, kre)
            dst.append(dict(actual=a, ideal=i, engine=e, sub=sub, nsub=n, R=res["toolR"], fluteLen=res["fluteLen"],
                            rw=0.0, type=typ, lut=lut, cornerR=rc, dynBudget=res["dynBudget"], nf=nf))
    outs, frs = lib.sim_run(progs, frames=True)
    for p, q, out, fr in zip(progs, refs, outs, frs):
        o2, f2 = oracle.sim_run(q["actual"], q["ideal"], q["engine"], q["sub"], q["nsub"], q["R"], q["fluteLen"], 0.0,
                                1.0, q["type"], q["lut"], q["cornerR"], dynBudget=q["dynBudget"], nframes=q["nf"])
        assert len(fr) == len(f2) == q["nf"]
        assert [(o.dx, o.dy, o.nsteps, o.errVal, o.arc) for o in out] == [(o.dx, o.dy, o.nsteps, o.errVal, o.arc)
                                                                         for o in o2]
        for a, b in zip(fr, f2):
            assert bytes(a) == bytes(b)
        assert np.array_equal(p["actual"].H(), q["actual"].H())
