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


def sim_substeps(x, y, z, rapid, line, F, S, N, hx):
    """The sub-steps Simulator.step visits over a discretised path with an unbounded frame
    budget (src/sim.js:363-381,389-405), as mc_substep records: position, the feed direction
    the simulator holds (this.u, starting at [1, 0]), machine time ds / v, feed per tooth
    along the xy path, axial feed per tooth while descending, the move's xy length, rapid
    flag and program line.  F and S are the path points' feed [mm/min] and speed [rpm]
    (overrides at 1); N the tool's flutes."""
    from ._abi import Substep
    step = stamp_step(hx)
    rows = []
    u = (1.0, 0.0)
    for k in range(1, len(x)):
        ax, ay, az, bx, by, bz = x[k - 1], y[k - 1], z[k - 1], x[k], y[k], z[k]
        seg_len = hypot_js(bx - ax, by - ay, bz - az)
        v = (15000 if rapid[k] else F[k]) / 60
        ds_max = 1.0 if rapid[k] else step
        s = 0.0
        while True:
            ds = min(ds_max, seg_len - s)
            if seg_len - s <= 1e-9:
                break
            dtm = ds / v
            s += ds
            f = s / seg_len if seg_len > 0 else 1.0
            px, py, pz = ax + (bx - ax) * f, ay + (by - ay) * f, az + (bz - az) * f
            dxy = hypot_js(bx - ax, by - ay)
            if dxy > 1e-6:
                u = ((bx - ax) / dxy, (by - ay) / dxy)
            fza = fz_eff = 0.0
            if not rapid[k]:   # a rapid never cuts, so its feeds are not used
                Sk, Fk = S[k], F[k]
                fza = (Fk / (Sk * N)) * ((az - bz) / seg_len) if (bz < az and Sk >= 1) else 0.0
                fz_eff = (Fk / (Sk * N)) * (dxy / seg_len) if dxy > 1e-6 else 0.0
            rows.append(Substep(px, py, pz, u[0], u[1], dtm, fz_eff, fza, dxy, 1 if rapid[k] else 0, int(line[k])))
            if s >= seg_len - 1e-9:
                break
    arr = (Substep * max(1, len(rows)))(*rows)
    return arr, len(rows)


GUARD = 4000   # loop passes per frame at most (src/sim.js:362)


def sim_frames(x, y, z, rapid, line, F, S, N, hx, dt_real, speed=1.0, feed_override=1.0, max_frames=1 << 30):
    """The sub-steps of successive Simulator.step(dt_real) frames at playback speed `speed`
    (src/sim.js:351-377,448), as mc_substep records: each frame covers dt_real * speed of
    machine time, its last sub-step shortened to what is left of it (ds = min(dsMax, segLen -
    segS, v budgetT)); the frame's sub-step estimate nSubEst = max(1, ceil(budgetT rate / 60 /
    stampStep)) from the segment it starts on, and each sub-step's guard (the loop's count in
    its frame, segment switches included) feed the dynamics budget's share on the device
    (src/sim.js:358-362,410).  The first sub-step of a frame carries MC_FRAME_FIRST, the
    last MC_FRAME_LAST.  F, S: the path points' feed [mm/min] and speed [rpm] (spindle
    override 1); N the tool's flutes.  Returns (Substep array, count, frames), frames a list of
    (first sub-step, sub-step count) per frame."""
    from ._abi import Substep, MC_FRAME_FIRST, MC_FRAME_LAST
    step = stamp_step(hx)
    npts = len(x)
    rows, frames = [], []
    u = (1.0, 0.0)
    seg, seg_s = 1, 0.0
    done = npts < 2
    while not done and len(frames) < max_frames:
        budget = dt_real * speed
        c = min(seg, npts - 1)
        rate = 15000 if rapid[c] else F[c] * feed_override
        n_sub_est = max(1, math.ceil(budget * rate / 60 / step))
        guard = 0
        first = len(rows)
        while budget > 1e-9 and not done:
            if not guard < GUARD:
                break
            guard += 1
            ax, ay, az, bx, by, bz = x[seg - 1], y[seg - 1], z[seg - 1], x[seg], y[seg], z[seg]
            seg_len = hypot_js(bx - ax, by - ay, bz - az)
            v = (15000 if rapid[seg] else F[seg] * feed_override) / 60
            ds_max = 1.0 if rapid[seg] else step
            ds = min(ds_max, seg_len - seg_s, v * budget)
            if seg_len - seg_s <= 1e-9:
                seg, seg_s = seg + 1, 0.0
                if seg >= npts:
                    done, seg = True, npts - 1
                continue
            dtm = ds / v
            budget -= dtm
            seg_s += ds
            f = seg_s / seg_len if seg_len > 0 else 1.0
            px, py, pz = ax + (bx - ax) * f, ay + (by - ay) * f, az + (bz - az) * f
            dxy = hypot_js(bx - ax, by - ay)
            if dxy > 1e-6:
                u = ((bx - ax) / dxy, (by - ay) / dxy)
            fza = fz_eff = 0.0
            if not rapid[seg]:
                Sk, Fk = S[seg], F[seg] * feed_override
                fza = (Fk / (Sk * N)) * ((az - bz) / seg_len) if (bz < az and Sk >= 1) else 0.0
                fz_eff = (Fk / (Sk * N)) * (dxy / seg_len) if dxy > 1e-6 else 0.0
            rows.append(Substep(px, py, pz, u[0], u[1], dtm, fz_eff, fza, dxy, 1 if rapid[seg] else 0, int(line[seg]),
                                float(n_sub_est), guard, 0))
            if seg_s >= seg_len - 1e-9:
                seg, seg_s = seg + 1, 0.0
                if seg >= npts:
                    done, seg = True, npts - 1
        if len(rows) > first:
            rows[first].frame |= MC_FRAME_FIRST
            rows[-1].frame |= MC_FRAME_LAST
        frames.append((first, len(rows) - first))
    arr = (Substep * max(1, len(rows)))(*rows)
    return arr, len(rows), frames


CHORD_TOL = 0.001   # largest chord error of an arc [mm] (src/gcode.js:433)


def discretize(moves, max_seg=0.5):
    """The path discretizer (src/gcode.js:435-466) in the reference's arithmetic: moves ->
    polyline points, each carrying the attributes of the segment ending at it.  An arc is
    cut into segs = max(2 (1 for a null arc), ceil(|sweep| r / max_seg), ceil(|sweep| /
    dAng)) chords, dAng the angle whose chord lies CHORD_TOL inside the arc; its last point
    snaps to the move's end.  atan2 / acos / cos / sin are V8's (the library's host math).

    moves: dicts with from, to ([x, y, z]), rapid, F, S, line, and arc None or {ca, cb, cw,
    axes ([a, b, n]), sweep (None unless set by cutter compensation)}.
    Returns a dict of lists x, y, z, rapid, F, S, line."""
    from . import lib
    out = dict(x=[], y=[], z=[], rapid=[], F=[], S=[], line=[])

    def put(p, rapid, F, S, line):
        out["x"].append(p[0]); out["y"].append(p[1]); out["z"].append(p[2])
        out["rapid"].append(1 if rapid else 0); out["F"].append(F); out["S"].append(S); out["line"].append(line)

    if not moves:
        return out
    m0 = moves[0]
    put(list(m0["from"]), True, m0["F"], m0["S"], m0["line"])
    AX = {"x": 0, "y": 1, "z": 2}
    two_pi = 2 * math.pi
    for mv in moves:
        frm, to, arc = mv["from"], mv["to"], mv.get("arc")
        if not arc:
            put(list(to), mv["rapid"], mv["F"], mv["S"], mv["line"])
            continue
        a, b, n = (AX[k] for k in (arc.get("axes") or ["x", "y", "z"]))
        ca, cb = arc["ca"], arc["cb"]
        r = hypot_js(frm[a] - ca, frm[b] - cb)
        a0, a1 = lib.js_math("atan2", [frm[b] - cb, to[b] - cb], [frm[a] - ca, to[a] - ca])
        a0, a1 = float(a0), float(a1)
        sweep = a1 - a0
        if arc.get("sweep") is not None:
            sweep = arc["sweep"]
        elif arc["cw"]:
            if sweep >= -1e-9:
                sweep -= two_pi
        elif sweep <= 1e-9:
            sweep += two_pi
        d_ang = 2 * float(lib.js_math("acos", [max(-1.0, 1 - CHORD_TOL / max(r, CHORD_TOL))])[0])
        segs = max(1 if abs(sweep) * r < 1e-9 else 2, math.ceil(abs(sweep) * r / max_seg),
                   math.ceil(abs(sweep) / d_ang))
        ang = [a0 + sweep * i / segs for i in range(1, segs + 1)]
        cs, sn = lib.js_math("cos", ang), lib.js_math("sin", ang)
        for i in range(1, segs + 1):
            p = [0.0, 0.0, 0.0]
            p[a] = ca + r * float(cs[i - 1])
            p[b] = cb + r * float(sn[i - 1])
            p[n] = frm[n] + (to[n] - frm[n]) * i / segs
            if i == segs:
                p = list(to)   # snap to the exact end point
            put(p, mv["rapid"], mv["F"], mv["S"], mv["line"])
    return out
