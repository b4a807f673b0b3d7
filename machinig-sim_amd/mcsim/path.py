"""Stamp schedule along a path: the sub-step positions Simulator.step visits
(src/sim.js:351-377,448, with an unbounded frame budget), in the reference's arithmetic.

segLen is Math.hypot of the three deltas (V8: scale by the largest magnitude, Kahan-
compensated sum, sqrt, rescale); the running arc length segS accumulates ds in order;
p = a + (b - a) * (segS / segLen).
"""
import math

import numpy as np

from ._abi import Stamp


def hypot_js(*v):
    a = [abs(float(x)) for x in v]
    if any(math.isinf(x) for x in a):
        return math.inf
    if any(math.isnan(x) for x in a):
        return math.nan
    mx = max(a) if a else 0.0
    if mx == 0:
        return 0.0
    s = 0.0
    comp = 0.0
    for x in a:
        n = x / mx
        summand = n * n - comp
        pre = s + summand
        comp = (pre - s) - summand
        s = pre
    return math.sqrt(s) * mx


def stamp_step(hx):
    """stampStep = min(hx, 0.4) * 0.6 (src/sim.js:355)"""
    return min(hx, 0.4) * 0.6


def substeps(x, y, z, rapid, hx):
    """Sub-step positions of Simulator.step over a discretised path (src/sim.js:363-377,448)
    with an unbounded frame budget: segment k runs from point k-1 to point k in steps of
    ds = min(dsMax, segLen - segS) (dsMax 1 mm on rapids, min(hx, 0.4) * 0.6 otherwise),
    ending when segLen - segS <= 1e-9 before a step or segS >= segLen - 1e-9 after one.
    Returns (px, py, pz, segment index) arrays."""
    step = stamp_step(hx)
    px, py, pz, seg = [], [], [], []
    for k in range(1, len(x)):
        ax, ay, az, bx, by, bz = x[k - 1], y[k - 1], z[k - 1], x[k], y[k], z[k]
        seg_len = hypot_js(bx - ax, by - ay, bz - az)
        ds_max = 1.0 if rapid[k] else step
        s = 0.0
        while not (seg_len - s <= 1e-9):
            s += min(ds_max, seg_len - s)
            f = s / seg_len if seg_len > 0 else 1.0
            px.append(ax + (bx - ax) * f)
            py.append(ay + (by - ay) * f)
            pz.append(az + (bz - az) * f)
            seg.append(k)
            if s >= seg_len - 1e-9:
                break
    return np.array(px), np.array(py), np.array(pz), np.array(seg, dtype=np.int64)


def schedule(segments, hx, err_val=0.0):
    """segments: (x0, y0, z0, x1, y1, z1, R, type, line) -> numpy record arrays of stamps
    (cx, cy, tipZ, R, type, line) and the index of each stamp's segment."""
    step = stamp_step(hx)
    cx, cy, cz, rr, tt, ll, seg = [], [], [], [], [], [], []
    for si, (x0, y0, z0, x1, y1, z1, R, typ, line) in enumerate(segments):
        seg_len = hypot_js(x1 - x0, y1 - y0, z1 - z0)
        s = 0.0
        while not (seg_len - s <= 1e-9):
            ds = min(step, seg_len - s)
            s += ds
            f = s / seg_len if seg_len > 0 else 1.0
            cx.append(x0 + (x1 - x0) * f)
            cy.append(y0 + (y1 - y0) * f)
            cz.append(z0 + (z1 - z0) * f)
            rr.append(R)
            tt.append(typ)
            ll.append(line)
            seg.append(si)
            if s >= seg_len - 1e-9:
                break
    out = np.zeros(len(cx), dtype=[("cx", "<f8"), ("cy", "<f8"), ("tipZ", "<f8"), ("R", "<f8"),
                                   ("type", "<i4"), ("line", "<i4"), ("errVal", "<f8")])
    out["cx"], out["cy"], out["tipZ"], out["R"] = cx, cy, cz, rr
    out["type"], out["line"], out["errVal"] = tt, ll, err_val
    return out, np.asarray(seg, dtype=np.int64)


def as_stamp_array(rec):
    """numpy stamp records -> ctypes Stamp array sharing the same bytes layout"""
    assert rec.dtype.itemsize == np.dtype([("a", "<f8", 4), ("b", "<i4", 2), ("c", "<f8")]).itemsize
    rec = np.ascontiguousarray(rec)
    arr = (Stamp * len(rec)).from_buffer_copy(rec.tobytes())
    return arr


def path_points(x, y, z, rapid, line, F, S):
    """the polyline as an mc_path_point array (point k carries the segment ending at it)"""
    from ._abi import PathPoint
    n = len(x)
    pts = (PathPoint * max(1, n))()
    for k in range(n):
        # a rapid runs at the machine's rapid rate: its feed may be absent (None)
        if F[k] is None and not rapid[k]:
            raise ValueError("path point %d: a feed move needs a feed rate" % k)
        f = 0.0 if F[k] is None else float(F[k])
        pts[k] = PathPoint(float(x[k]), float(y[k]), float(z[k]), f, float(S[k]), 1 if rapid[k] else 0,
                           int(line[k]), k, 0)
    return pts, n


def _schedule(pts, n, N, hx, dt_real=0.0, speed=1.0, feed_override=1.0, max_frames=1 << 30):
    from . import lib
    from ._abi import ScheduleOpts, Substep
    import ctypes as C
    o = ScheduleOpts(float(hx), float(dt_real), float(speed), float(feed_override), int(N),
                     int(min(max_frames, (1 << 31) - 1)))
    ns, nf = C.c_int(0), C.c_int(0)
    L = lib.load()
    lib._check(L.mc_sim_schedule(pts, n, C.byref(o), None, 0, C.byref(ns), None, 0, C.byref(nf)))
    arr = (Substep * max(1, ns.value))()
    fr = (C.c_int32 * max(2, 2 * nf.value))()
    lib._check(L.mc_sim_schedule(pts, n, C.byref(o), arr, ns.value, C.byref(ns), fr if dt_real > 0 else None,
                                 nf.value if dt_real > 0 else 0, C.byref(nf)))
    frames = [(fr[2 * f], fr[2 * f + 1]) for f in range(nf.value)]
    return arr, ns.value, frames


def sim_substeps(x, y, z, rapid, line, F, S, N, hx):
    """The sub-steps Simulator.step visits over a discretised path with an unbounded frame
    budget (src/sim.js:363-381,389-405), as mc_substep records (mc_sim_schedule with dt_real
    0): position, the feed direction the simulator holds (this.u, starting at [1, 0]),
    machine time ds / v, feed per tooth along the xy path, axial feed per tooth while
    descending, the move's xy length, rapid flag and program line.  F and S are the path
    points' feed [mm/min] and speed [rpm] (overrides at 1); N the tool's flutes.
    Returns (Substep array, count)."""
    pts, n = path_points(x, y, z, rapid, line, F, S)
    arr, ns, _ = _schedule(pts, n, N, hx)
    return arr, ns


GUARD = 4000   # loop passes per frame at most (src/sim.js:362)


def sim_frames(x, y, z, rapid, line, F, S, N, hx, dt_real, speed=1.0, feed_override=1.0, max_frames=1 << 30):
    """The sub-steps of successive Simulator.step(dt_real) frames at playback speed `speed`
    (src/sim.js:351-377,448), as mc_substep records (mc_sim_schedule): each frame covers
    dt_real * speed of machine time, its last sub-step shortened to what is left of it (ds =
    min(dsMax, segLen - segS, v budgetT)); the frame's sub-step estimate nSubEst = max(1,
    ceil(budgetT rate / 60 / stampStep)) from the segment it starts on, and each sub-step's
    guard (the loop's count in its frame, segment switches included, at most GUARD) feed the
    dynamics budget's share on the device (src/sim.js:358-362,410).  The first sub-step of a
    frame carries MC_FRAME_FIRST, the last MC_FRAME_LAST.  F, S: the path points' feed
    [mm/min] and speed [rpm] (spindle override 1); N the tool's flutes.  Returns (Substep
    array, count, frames), frames a list of (first sub-step, sub-step count) per frame."""
    if not dt_real > 0:
        raise ValueError("dt_real must be > 0")
    pts, n = path_points(x, y, z, rapid, line, F, S)
    return _schedule(pts, n, N, hx, dt_real, speed, feed_override, max_frames)


CHORD_TOL = 0.001   # largest chord error of an arc [mm] (src/gcode.js:433)
_AX = {"x": 0, "y": 1, "z": 2}


def moves_array(moves):
    """reference parser moves (dicts with from, to ([x, y, z]), rapid, F, S, line and arc None
    or {ca, cb, cw, axes ([a, b, n]), sweep}) as an mc_move array"""
    from ._abi import Move
    arr = (Move * max(1, len(moves)))()
    for i, mv in enumerate(moves):
        m = arr[i]
        m.from_[:] = [float(v) for v in mv["from"]]
        m.to[:] = [float(v) for v in mv["to"]]
        m.F, m.S, m.rapid, m.line = float(mv["F"]), float(mv["S"]), 1 if mv["rapid"] else 0, int(mv["line"])
        arc = mv.get("arc")
        if arc:
            m.arc, m.cw = 1, 1 if arc["cw"] else 0
            m.axes[:] = [_AX[k] for k in (arc.get("axes") or ["x", "y", "z"])]
            m.ca, m.cb = float(arc["ca"]), float(arc["cb"])
            if arc.get("sweep") is not None:
                m.has_sweep, m.sweep = 1, float(arc["sweep"])
    return arr


def discretize(moves, max_seg=0.5):
    """The path discretizer (src/gcode.js:435-466, mc_path_discretize): moves -> polyline
    points, each carrying the attributes of the segment ending at it.  An arc is cut into
    segs = max(2 (1 for a null arc), ceil(|sweep| r / max_seg), ceil(|sweep| / dAng)) chords,
    dAng the angle whose chord lies CHORD_TOL inside the arc; its last point snaps to the
    move's end.  Returns a dict of lists x, y, z, rapid, F, S, line."""
    import ctypes as C
    from . import lib
    from ._abi import PathPoint
    out = dict(x=[], y=[], z=[], rapid=[], F=[], S=[], line=[])
    if not moves:
        return out
    arr = moves_array(moves)
    n = C.c_int(0)
    L = lib.load()
    lib._check(L.mc_path_discretize(arr, len(moves), float(max_seg), None, 0, C.byref(n)))
    pts = (PathPoint * n.value)()
    lib._check(L.mc_path_discretize(arr, len(moves), float(max_seg), pts, n.value, C.byref(n)))
    for p in pts:
        for k in ("x", "y", "z", "F", "S", "rapid", "line"):
            out[k].append(getattr(p, k))
    return out
