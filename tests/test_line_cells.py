"""The coupled workload's per-line stability cells (bench.line_cells): the report keeps each
line's heaviest cut with an engagement arc (src/report.js:27-28: ap >= the kept depth and
pex > pst), with the arc the simulator shows after the sub-step (the last one set, [0, pi]
before any; src/sim.js:421-423); lines whose heaviest cut is an end cut alone (ap <= 0.01)
ask no radial limit."""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def _prog(rows):
    from mcsim._abi import Substep, SubstepOut
    n = len(rows)
    sub, out = (Substep * n)(), (SubstepOut * n)()
    for k, (line, cut, ap, arc, st, ex, fz) in enumerate(rows):
        sub[k].line, sub[k].fzEff = line, fz
        out[k].cut, out[k].ap, out[k].arc, out[k].phiSt, out[k].phiEx = cut, ap, arc, st, ex
    return (sub, n), out


def test_heaviest_cut_per_line():
    import bench
    from mcsim import workloads as W
    rows = [
        (1, 1, 0.0, 1, 0.0, math.pi, 0.0),     # plunge: an end cut alone -> no cell
        (2, 1, 1.0, 1, 0.2, 2.0, 0.03),
        (2, 1, 1.5, 0, 9.0, 9.0, 0.031),       # keeps the arc [0.2, 2.0]; deeper -> kept
        (2, 1, 1.5, 1, 0.4, 2.5, 0.032),       # equal depth -> replaces (>=)
        (2, 0, 9.0, 0, 0.0, 0.0, 0.0),         # not cutting -> ignored
        (3, 1, 2.0, 1, 1.0, 1.0, 0.04),        # empty arc (pex == pst) -> not kept
        (3, 1, 0.5, 1, 0.1, 0.9, 0.05),
    ]
    prog, out = _prog(rows)
    tool = W.COUPLED_TOOLS[0]
    cells, keys = bench.line_cells([prog], [out], [tool])
    assert [(p, ln) for p, ln, _ in keys] == [(0, 2), (0, 3)]
    assert [ap for _, _, ap in keys] == [1.5, 0.5]
    c2, c3 = cells
    assert (c2.phiSt, c2.phiEx, c2.fz) == (0.4, 2.5, 0.032)
    assert (c3.phiSt, c3.phiEx, c3.fz) == (0.1, 0.9, 0.05)
    assert c2.S == W.COUPLED_SPEED and c2.tool.N == tool["N"] and np.isclose(c2.tool.R, tool["R"])
