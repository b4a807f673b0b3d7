"""The path format (src/gcode.js:435-466, the discretizer that feeds the simulator's
schedule): mcsim.path.discretize on the reference parser's own moves (full circles, R arcs,
helical arcs, XZ / YZ plane arcs, a tiny arc) gives the reference's polyline bit for bit at
several chord lengths; and for the simulator's programs, discretize -> sim_substeps gives
every sub-step position of the reference Simulator's runs."""
import numpy as np

import golden_util as G


def _same_pts(got, ref):
    return (np.array_equal(got["x"], G.f64(ref["x"])) and np.array_equal(got["y"], G.f64(ref["y"])) and
            np.array_equal(got["z"], G.f64(ref["z"])) and got["rapid"] == ref["rapid"] and
            got["line"] == ref["line"] and got["F"] == ref["F"] and got["S"] == ref["S"])


def test_discretize_matches_reference():
    from mcsim import path
    g = G.load("paths")
    arcs = 0
    for case, res in zip(g["cases"], g["results"]):
        arcs += sum(1 for m in res["moves"] if m["arc"])
        for ms, ref in zip(case["maxSegs"], res["pts"]):
            got = path.discretize(res["moves"], ms)
            assert _same_pts(got, ref), (case["program"][:40], ms)
    assert arcs >= 10


def test_discretize_then_schedule_matches_simulator_runs():
    from mcsim import path
    gp, gs = G.load("paths"), G.load("simx")
    for pi, si in ((0, 0), (1, 1), (0, 3)):
        res, sres = gp["results"][pi], gs["results"][si]
        assert gp["cases"][pi]["program"] == gs["cases"][si]["program"]
        pts = path.discretize(res["moves"], 0.5)
        P = sres["path"]
        assert np.array_equal(pts["x"], G.f64(P["x"])) and np.array_equal(pts["y"], G.f64(P["y"]))
        assert np.array_equal(pts["z"], G.f64(P["z"])) and pts["line"] == P["line"]
        sub, n = path.sim_substeps(pts["x"], pts["y"], pts["z"], pts["rapid"], pts["line"], pts["F"], pts["S"],
                                   sres["toolN"], sres["hx"])
        ide = sres["ideal"]
        assert n == len(ide["type"])
        assert np.array_equal([s.x for s in sub[:n]], G.f64(ide["cx"]))
        assert np.array_equal([s.y for s in sub[:n]], G.f64(ide["cy"]))
