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
