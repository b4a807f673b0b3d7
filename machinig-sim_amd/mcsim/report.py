"""The machining report on the device pipeline's output: Simulator.collect's per-frame
arithmetic (src/sim.js:644-663,707-710) and the per-line LineReport (src/report.js:9-33).

The device pipeline (mc_sim_run with frames) hands back, per sub-step, its cut flag, axial
depth and the engagement arc it shows, and per frame the engine's taken stats and chatter
index.  This module turns them into what the simulator reports:

- per frame, the display's smoothed chatter index and once-per-tooth scatter (EMA 0.2) and
  the frame's state: chatter when index > 0.55 and once > 3 um, marginal when index > 0.35
  and once > 1.5 um (src/sim.js:663-664,707);
- per line, LineReport.time for every sub-step and LineReport.sample for every frame that
  cut: peak force, vibration, deflection, chatter and marginal time, mean chip thickness and
  the heaviest cut's engagement (src/report.js:18-29).

The edge temperature, its swing and the tool stress of LineReport come from the simulator's
thermal and stress models, which stay with the caller: `extra` supplies them per frame
(0 when absent).
"""
import math

from .path import hypot_js

FIELDS = ("t", "cut", "feedT", "F", "vib", "defl", "temp", "stress", "chatter", "marginal", "swing", "h", "hw",
          "pst", "pex", "ap")


class LineReport:
    """src/report.js:9-33: per program line, machine time, cutting and feed time, and the
    peaks and chatter time of the frames credited to it."""

    def __init__(self):
        self.lines = {}          # insertion-ordered, like the reference's Map

    def rec(self, line):
        r = self.lines.get(line)
        if r is None:
            r = dict(line=line, t=0.0, cut=0.0, feedT=0.0, F=0.0, vib=0.0, defl=0.0, temp=0.0, stress=0.0,
                     chatter=0.0, marginal=0.0, swing=0.0, h=0.0, hw=0.0, pst=0.0, pex=math.pi, ap=0.0)
            self.lines[line] = r
        return r

    def time(self, line, dt, cutting, feed=False):
        r = self.rec(line)
        r["t"] += dt
        if cutting:
            r["cut"] += dt
        if feed:
            r["feedT"] += dt

    def sample(self, line, v, state, dt):
        r = self.rec(line)
        for k in ("F", "vib", "defl", "temp", "stress"):
            r[k] = _jmax(r[k], v[k])
        r["swing"] = _jmax(r["swing"], v.get("swing") or 0)
        if state == "chatter":
            r["chatter"] += dt
        elif state == "marginal":
            r["marginal"] += dt
        if v["h"] > 0 and dt > 0:
            r["hw"] += dt
            r["h"] += (v["h"] - r["h"]) * dt / r["hw"]
        if v["ap"] >= r["ap"] and v["pex"] > v["pst"]:
            r["ap"], r["pst"], r["pex"] = v["ap"], v["pst"], v["pex"]

    def pack(self):
        """rows [line, t, cut, feedT, F, vib, defl, temp, stress, chatter, marginal, swing, h, hw,
        pst, pex, ap] in the order the lines first appeared (LineReport.pack)"""
        return [[r["line"]] + [r[k] for k in FIELDS] for r in self.lines.values()]


def _jmax(a, b):
    """Math.max of two numbers (NaN wins)"""
    if a != a or b != b:
        return math.nan
    return a if a > b else b


def frame_state(chatter_idx, once, cutting):
    """the display state of a frame (src/sim.js:707)"""
    if not cutting:
        return "idle"
    if chatter_idx > 0.55 and once > 3:
        return "chatter"
    if chatter_idx > 0.35 and once > 1.5:
        return "marginal"
    return "stable"


class FrameReporter:
    """Simulator.collect's report part (src/sim.js:644-663,707-710), fed frame by frame with
    the sub-steps the frame took and their device outputs."""

    def __init__(self):
        self.report = LineReport()
        self.chatter_idx = 0.0        # display.chatterIdx
        self.once = 0.0               # display.once
        self.phi_st, self.phi_ex = 0.0, math.pi   # display.phiSt / phiEx (blankDisplay)
        self.line = None              # this.line
        self.last_cut_line = None     # this.lastCutLine
        self.ap = 0.0                 # this.ap
        self.states = []

    def substeps(self, sub, out, first, count):
        """the frame's sub-steps: report.time per sub-step, the simulator's line, last cutting
        line, axial depth and displayed arc (src/sim.js:379,387,421-423,445-447); returns
        the frame's cutting time fs.cutTime"""
        cut_time = 0.0
        for k in range(first, first + count):
            s, o = sub[k], out[k]
            self.line = s.line
            self.ap = o.ap
            if o.arc:
                self.phi_st, self.phi_ex = o.phiSt, o.phiEx
            cut_now = bool(o.cut)
            if cut_now:
                cut_time += s.dtm
            self.report.time(s.line, s.dtm, cut_now, not s.rapid)
            if cut_now:
                self.last_cut_line = s.line
        return cut_time

    def collect(self, stats, ci, cut_time, extra=None):
        """one frame's collect: stats is the frame's taken stats (mc_engine_stats fields),
        ci (index, amp, once) its chatter index; returns the frame's state"""
        n = stats.n
        cutting = n > 0
        F_max = stats.maxF
        defl = hypot_js(stats.sumX / n, stats.sumY / n) * 1e6 if cutting else 0
        vib = _jmax(stats.maxX - stats.minX, stats.maxY - stats.minY) * 1e6 / 2 if cutting else 0
        self.chatter_idx += ((ci[0] if cutting else 0) - self.chatter_idx) * 0.2
        self.once += ((ci[2] or 0 if cutting else 0) - self.once) * 0.2
        state = frame_state(self.chatter_idx, self.once, cutting)
        self.states.append(state)
        if cutting:
            x = extra or {}
            line = self.last_cut_line if self.last_cut_line is not None else self.line
            self.report.sample(line, dict(F=F_max, vib=vib, defl=defl, temp=x.get("temp", 0.0),
                                          stress=x.get("stress", 0.0), swing=x.get("swing", 0.0),
                                          h=stats.sumHb / stats.sumB if stats.sumB > 0 else 0, ap=self.ap or 0,
                                          pst=self.phi_st, pex=self.phi_ex), state, cut_time)
        return state


def line_engagements(report, sub, out):
    """the per-line stability question of the report (src/main.js:1033-1038: asked for every
    line with a positive speed and an engagement arc, pex > pst): (line, ap, phiSt, phiEx, fz)
    of each line's heaviest cut as the report kept it (src/report.js:27-28), fz the feed per
    tooth of the line's last cutting sub-step, in the report's line order"""
    fz = {}
    for k in range(len(out)):
        if out[k].cut:
            fz[sub[k].line] = sub[k].fzEff
    rows = []
    for r in report.lines.values():
        if r["cut"] > 0 and r["pex"] > r["pst"]:
            rows.append((r["line"], r["ap"], r["pst"], r["pex"], fz.get(r["line"], 0.0)))
    return rows


def report_program(sub, out, frames, frame_outs, extra=None):
    """the report of one program run through the device pipeline with frames: frames is the
    schedule's (first, count) list, frame_outs the FrameOut of each non-empty frame, in order"""
    rep = FrameReporter()
    it = iter(frame_outs)
    for f, (first, count) in enumerate(frames):
        if count == 0:
            rep.collect(_EMPTY, (0.0, 0.0, 0.0), 0.0)
            continue
        fo = next(it)
        cut_time = rep.substeps(sub, out, first, count)
        rep.collect(fo.stats, (fo.index, fo.amp, fo.once), cut_time, extra(f) if extra else None)
    return rep


class _Empty:
    n = 0


_EMPTY = _Empty()
