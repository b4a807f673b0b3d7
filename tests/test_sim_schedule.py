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
    runs (flat, ball, bull-nose and imported-edge tools): deflected stamp positions, vibration
    values, step counts, per-stamp volumes and every map, bit for bit"""
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
                             typ, lut, rc, edge=G.simx_edge(case))
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


def test_schedule_abi_checks():
    """mc_sim_schedule and mc_path_discretize at the C ABI: size queries, short outputs and
    bad options are reported; a path of one point has no sub-steps; max_frames stops the
    frame loop; frames tile the sub-steps in order"""
    import ctypes as C
    from mcsim import lib, path
    from mcsim._abi import ScheduleOpts, Substep
    L = lib.load()
    pts, n = path.path_points([0.0, 10.0, 10.0], [0.0, 0.0, 5.0], [1.0, -1.0, -1.0], [1, 0, 0], [1, 2, 3],
                              [800.0, 800.0, 600.0], [8000.0, 8000.0, 8000.0])
    ns, nf = C.c_int(-1), C.c_int(-1)
    o = ScheduleOpts(0.1, 1 / 60, 1.0, 1.0, 3, 1000)
    assert L.mc_sim_schedule(pts, n, C.byref(o), None, 0, C.byref(ns), None, 0, C.byref(nf)) == 0
    assert ns.value > 0 and nf.value > 0
    arr = (Substep * ns.value)()
    fr = (C.c_int32 * (2 * nf.value))()
    assert L.mc_sim_schedule(pts, n, C.byref(o), arr, ns.value - 1, C.byref(ns), fr, nf.value, C.byref(nf)) != 0
    assert L.mc_sim_schedule(pts, n, C.byref(o), arr, ns.value, C.byref(ns), fr, nf.value - 1, C.byref(nf)) != 0
    assert L.mc_sim_schedule(pts, n, C.byref(o), arr, ns.value, C.byref(ns), fr, nf.value, C.byref(nf)) == 0
    nxt = 0
    for f in range(nf.value):
        assert fr[2 * f] == nxt
        nxt += fr[2 * f + 1]
    assert nxt == ns.value
    for bad in (ScheduleOpts(0.0, 0.0, 1.0, 1.0, 3, 0), ScheduleOpts(0.1, -1.0, 1.0, 1.0, 3, 0),
                ScheduleOpts(0.1, 0.01, 0.0, 1.0, 3, 10), ScheduleOpts(0.1, 0.0, 1.0, 1.0, 0, 0)):
        assert L.mc_sim_schedule(pts, n, C.byref(bad), None, 0, C.byref(ns), None, 0, C.byref(nf)) != 0
    o2 = ScheduleOpts(0.1, 1 / 60, 1.0, 1.0, 3, 3)
    assert L.mc_sim_schedule(pts, n, C.byref(o2), None, 0, C.byref(ns), None, 0, C.byref(nf)) == 0 and nf.value == 3
    assert L.mc_sim_schedule(pts, 1, C.byref(o), None, 0, C.byref(ns), None, 0, C.byref(nf)) == 0
    assert ns.value == 0 and nf.value == 0
    assert path.sim_substeps([1.0], [2.0], [3.0], [1], [1], [0.0], [0.0], 2, 0.1)[1] == 0
    assert path.discretize([]) == dict(x=[], y=[], z=[], rapid=[], F=[], S=[], line=[])
    mv = dict(to=[1.0, 0.0, 0.0], F=100.0, S=1000.0, rapid=0, line=1, arc=dict(ca=0.0, cb=0.0, cw=0, axes=["x", "x", "z"]))
    mv["from"] = [0.0, 1.0, 0.0]
    arr = path.moves_array([mv])
    assert L.mc_path_discretize(arr, 1, 0.5, None, 0, C.byref(ns)) != 0
    mv["arc"]["axes"] = ["x", "y", "z"]
    arr = path.moves_array([mv])
    assert L.mc_path_discretize(arr, 1, 0.0, None, 0, C.byref(ns)) != 0
    assert L.mc_path_discretize(arr, 1, 0.5, None, 0, C.byref(ns)) == 0 and ns.value > 2


def test_coupled_programs_schedule_with_feedless_rapids():
    """the coupled workload's programs carry no feed on their rapids (None): the schedule
    runs them at the rapid rate, and a feed move without a feed is an error"""
    import pytest
    from mcsim import path, workloads as W
    width = 102.35
    tool, m = W.coupled_plan(0, 4)
    x, y, z, r, ln, F, S = W.coupled_program(1, width, m, speed=W.COUPLED_SPEED)
    assert F[0] is None and r[0]
    sub, n, frames = path.sim_frames(x, y, z, r, ln, F, S, tool["N"], width / 2047, 1e6)
    assert n > 1000 and sum(c for _, c in frames) == n
    sub2, n2 = path.sim_substeps(x, y, z, r, ln, F, S, tool["N"], width / 2047)
    assert n2 == n
    bad = list(F)
    bad[-1] = None
    with pytest.raises(ValueError):
        path.sim_substeps(x, y, z, r, ln, bad, S, tool["N"], width / 2047)
