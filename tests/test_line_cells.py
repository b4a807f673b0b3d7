"""The report's per-line stability question on pipeline outputs (mcsim.report): frames are
sampled once each at their end (Simulator.collect, src/sim.js:710) with the line, axial
depth and displayed arc of the frame's last sub-step; each line keeps its heaviest cut with
an engagement arc (src/report.js:27-28: ap >= the kept depth and pex > pst), and every line
that cut with an arc asks for its limit (src/main.js:1033-1038), an end cut alone too (ap 0,
arc [0, pi])."""
import math

from mcsim import report as RP
from mcsim._abi import EngineStats, FrameOut, Substep, SubstepOut


def _prog(rows):
    n = len(rows)
    sub, out = (Substep * n)(), (SubstepOut * n)()
    for k, (line, cut, ap, arc, st, ex, fz) in enumerate(rows):
        sub[k].line, sub[k].fzEff, sub[k].dtm = line, fz, 0.01
        out[k].cut, out[k].ap, out[k].arc, out[k].phiSt, out[k].phiEx = cut, ap, arc, st, ex
    return sub, out


def _frame(n_steps):
    f = FrameOut()
    f.stats = EngineStats()
    f.stats.n = n_steps
    f.stats.sumB = 1.0
    f.stats.sumHb = 0.02
    return f


def test_frames_sample_the_heaviest_cut_per_line():
    rows = [
        (1, 1, 0.0, 1, 0.0, math.pi, 0.02),    # plunge: an end cut alone
        (2, 1, 1.0, 1, 0.2, 2.0, 0.03),        # frame 0 ends here: line 2 at ap 1.0, arc [0.2, 2.0]
        (2, 1, 1.5, 0, 9.0, 9.0, 0.031),       # keeps the arc [0.2, 2.0]
        (2, 1, 1.5, 1, 0.4, 2.5, 0.032),       # frame 1 ends here: equal depth at its end -> replaces
        (2, 0, 9.0, 0, 0.0, 0.0, 0.0),         # frame 2: not cutting
        (3, 1, 2.0, 1, 1.0, 1.0, 0.04),        # frame 3 ends on an empty arc (pex == pst): not kept
        (3, 1, 0.5, 1, 0.1, 0.9, 0.05),        # frame 4
    ]
    sub, out = _prog(rows)
    frames = [(0, 2), (2, 2), (4, 1), (5, 1), (6, 1)]
    fouts = [_frame(10), _frame(10), _frame(0), _frame(10), _frame(10)]
    rep = RP.report_program(sub, out, frames, fouts)
    assert rep.states == ["stable", "stable", "idle", "stable", "stable"]
    lines = RP.line_engagements(rep.report, sub, out)
    assert [ln for ln, *_ in lines] == [1, 2, 3]
    # line 1 was never sampled at a frame's end: the report's defaults, ap 0 over [0, pi]
    assert lines[0][1:] == (0.0, 0.0, math.pi, 0.02)
    assert lines[1][1:] == (1.5, 0.4, 2.5, 0.032)
    assert lines[2][1:] == (0.5, 0.1, 0.9, 0.05)
    r2 = rep.report.lines[2]
    assert r2["t"] == 0.01 * 4 and abs(r2["cut"] - 0.03) < 1e-15


def test_chatter_thresholds():
    assert RP.frame_state(0.56, 3.1, True) == "chatter"
    assert RP.frame_state(0.56, 2.9, True) == "marginal"
    assert RP.frame_state(0.36, 1.6, True) == "marginal"
    assert RP.frame_state(0.36, 1.4, True) == "stable"
    assert RP.frame_state(0.9, 9.0, False) == "idle"
