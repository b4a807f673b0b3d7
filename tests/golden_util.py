"""Helpers to read the golden fixtures (tests/golden/*.json) written from the reference."""
import base64
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "machinig-sim_amd"))

from mcsim._abi import make_cell, Stamp, StampOpt, EngageSpec, MC_TOOL_BALL, MC_TOOL_FLAT  # noqa: E402
from mcsim._abi import MC_TIP_FLAT, MC_TIP_BALL, MC_TIP_BULL, MC_TIP_EDGE  # noqa: E402


def load(name):
    with open(os.path.join(HERE, "golden", "golden_%s.json" % name)) as f:
        return json.load(f)


def f64(s):
    return np.frombuffer(base64.b64decode(s), dtype="<f8")


def f32(s):
    return np.frombuffer(base64.b64decode(s), dtype="<f4")


def i32(s):
    return np.frombuffer(base64.b64decode(s), dtype="<i4")


def zf32(s):
    import zlib
    return np.frombuffer(zlib.decompress(base64.b64decode(s)), dtype="<f4")


def zi32(s):
    import zlib
    return np.frombuffer(zlib.decompress(base64.b64decode(s)), dtype="<i4")


def num(v):
    if v == "inf":
        return float("inf")
    if v == "-inf":
        return float("-inf")
    return float(v)


def cell_from_json(c):
    md = c["modal"]
    mx = [(m["k"], m["fn"], m["zeta"]) for m in (md.get("modesX") or [md])]
    my = [(m["k"], m["fn"], m["zeta"]) for m in (md.get("modesY") or [])] or None
    return make_cell(c["tool"], mx, my, c["mat"], c["S"], c["arc"], c["fz"], c["revs"],
                     c["nonlinear"])


def stamps_from_json(stamps):
    arr = (Stamp * len(stamps))()
    for i, s in enumerate(stamps):
        arr[i] = Stamp(s["cx"], s["cy"], s["tipZ"], s["R"],
                       MC_TOOL_BALL if s["type"] == "ball" else MC_TOOL_FLAT, s["line"], s["errVal"])
    return arr


def opts_from_json(stamps):
    """StampOpt array of a stamp list (mark / lut keys), or None when no stamp has one"""
    if not any("mark" in s or "lut" in s for s in stamps):
        return None
    arr = (StampOpt * len(stamps))()
    for i, s in enumerate(stamps):
        m = s.get("mark") or dict(ux=0.0, uy=0.0, f=0.0)
        arr[i] = StampOpt(m["ux"], m["uy"], m["f"], s.get("lut", -1), 0)
    return arr


def luts_from_json(b64_list):
    """concatenated f32 profiles + int32 offsets and lengths"""
    tabs = [f32(b) for b in b64_list]
    off = np.cumsum([0] + [len(t) for t in tabs[:-1]]).astype(np.int32) if tabs else np.zeros(0, np.int32)
    cat = np.concatenate(tabs).astype(np.float32) if tabs else np.zeros(0, np.float32)
    return cat, off, np.array([len(t) for t in tabs], np.int32)


def tip_shape(t):
    """tipLut tool dict -> (shape, D, cornerR, edge_r, edge_dz) of mc_tip_lut"""
    if t.get("edge"):
        return MC_TIP_EDGE, t["D"], 0.0, t["edge"]["r"], t["edge"]["dz"]
    kind = {"ball": MC_TIP_BALL, "bull": MC_TIP_BULL}.get(t.get("type"), MC_TIP_FLAT)
    return kind, t["D"], t.get("cornerR", 0.0) or 0.0, None, 0.0


def split_at(case):
    """stamp list halves around the clearDirty point (the whole list when there is none)"""
    st = case["stamps"]
    c = case.get("clearAt")
    return (st, []) if c is None else (st[:c], st[c:])


def rect(v):
    big = np.iinfo(np.int32).max
    return [big if x == "inf" else int(x) for x in v]


def stock_from_case(stock_mod, case, ref, use_ex=True):
    """Build a stock (oracle or product class with stamps_ex) and apply a golden stamp case"""
    s = stock_mod(case["w"], case["l"], case["h"], case["cell"])
    if case["wallBand"]:
        s.set_wall_band(case["wallBand"])
    luts = luts_from_json(ref.get("luts", []))
    st = case["stamps"]
    return s, stamps_from_json(st), opts_from_json(st), luts


def sim_stamps(cols):
    """The stamps a simulator run issued (columns of golden_sim.json) as Stamp / StampOpt
    arrays, with the per-stamp volumes it got back"""
    cx, cy, tz, R = f64(cols["cx"]), f64(cols["cy"]), f64(cols["tipZ"]), f64(cols["R"])
    ev, ux, uy, f = f64(cols["errVal"]), f64(cols["ux"]), f64(cols["uy"]), f64(cols["f"])
    n = len(cx)
    arr = (Stamp * n)()
    opt = (StampOpt * n)()
    for i in range(n):
        arr[i] = Stamp(cx[i], cy[i], tz[i], R[i], cols["type"][i], cols["line"][i], ev[i])
        opt[i] = StampOpt(ux[i], uy[i], f[i], cols["lut"][i], 0)
    return arr, opt, f64(cols["vol"])


def sub(arr, a, b):
    """elements a..b-1 of a ctypes array as a new array"""
    import ctypes
    t = type(arr[0]) if len(arr) else None
    out = (t * (b - a))()
    ctypes.memmove(out, ctypes.addressof(arr) + a * ctypes.sizeof(t), (b - a) * ctypes.sizeof(t))
    return out


def replay_sim(make_stock, stamp_fn, engage_fn, c, r):
    """Replay a simulator run's stamps on a stock, answering its engagement samples at the
    moments it took them; returns the actual and commanded stocks and the volumes"""
    spec = c["stock"]
    s = make_stock(spec["w"], spec["l"], spec["h"], c["cell"])
    ideal = make_stock(spec["w"], spec["l"], spec["h"], c["cell"])
    ideal.set_wall_band(2.0)
    luts = luts_from_json(r["luts"])
    arr, opt, vol_ref = sim_stamps(r["actual"])
    iarr, iopt, ivol_ref = sim_stamps(r["ideal"])
    vol, done = [], 0
    for k, cx, cy, ux, uy, R, hm in r["engage"]:
        if k > done:
            vol += list(stamp_fn(s, sub(arr, done, k), sub(opt, done, k), luts))
            done = k
        assert engage_fn(s, cx, cy, ux, uy, R) == num(hm)
    if len(arr) > done:
        vol += list(stamp_fn(s, sub(arr, done, len(arr)), sub(opt, done, len(arr)), luts))
    ivol = stamp_fn(ideal, iarr, iopt, luts)
    return s, ideal, np.array(vol), vol_ref, np.asarray(ivol), ivol_ref


def check_maps(stock_maps, ref, bits_of):
    for key, want in (("H", zf32(ref["H"])), ("err", zf32(ref["err"])), ("wallD", zf32(ref["wallD"])),
                      ("markS", zf32(ref["markS"])), ("markF", zf32(ref["markF"]))):
        assert np.array_equal(bits_of(stock_maps[key]), bits_of(want)), key
    assert np.array_equal(stock_maps["lineCut"], zi32(ref["lineCut"]))
    assert np.array_equal(stock_maps["lineWall"], zi32(ref["lineWall"]))


def enginex_case(c):
    """(Cell, EngageSpec) of a golden simulator-engine case"""
    md = c["modal"]
    mx = [(m["k"], m["fn"], m["zeta"]) for m in md["modesX"]]
    my = [(m["k"], m["fn"], m["zeta"]) for m in md["modesY"]]
    cell = make_cell(c["tool"], mx, my, c["mat"], c["S"], (0.0, 3.141592653589793), c["fz"], 1, c["nonlinear"])
    sp = EngageSpec()
    sp.ux, sp.uy = c["u"]
    sp.fz, sp.tipZ, sp.ap = c["fz"], c["tipZ"], c["ap"]
    sp.vx0, sp.vy0 = c["v0"]
    sp.fnMax = c["fnMax"]
    for b, v in enumerate(c["Hb"]):
        sp.Hb[b] = num(v)
    return cell, sp


def zf64(s):
    import zlib
    return np.frombuffer(zlib.decompress(base64.b64decode(s)), dtype="<f8")


ENGLOG_POST = ("x", "y", "vx", "vy", "w", "wv", "t", "pr", "runMeanW", "stepCount", "nz", "hist_i",
               "hist_count", "ts_i", "ts_count", "ws_i", "ws_count", "wsp_i", "wsp_count")
MAT_ORDER = ("Ktc", "Krc", "Kte", "Kre", "Kac", "Kae", "mc", "Ci")


def _same(a, b, rtol=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    if rtol == 0.0:
        return np.array_equal(a, b, equal_nan=True)
    fin = np.isfinite(b)
    if not np.array_equal(a[~fin], b[~fin], equal_nan=True):
        return False
    scale = np.max(np.abs(b[fin])) if np.any(fin) else 0.0
    return bool(np.allclose(a[fin], b[fin], rtol=rtol, atol=rtol * scale))


ENGLOG_EXACT = ("pr", "stepCount", "nz", "hist_i", "hist_count", "ts_i", "ts_count", "ws_i", "ws_count", "wsp_i",
                "wsp_count")


def englog_replayer(eng, res, max_runs=None, rtol=0.0, bad=None):
    """Generator over a call sequence recorded from the reference simulator
    (golden_englog.json): applies the recorded calls to `eng` (the oracle's OracleEngine or
    the library's Engine: same methods) up to each run, yields the run's step count for the
    caller to run, then compares, bit for bit (or within rtol), the engine's state, stats
    and chip events with the reference's; also every takeStats and chatterIndex answer and
    the sample rings at the end.  Mismatches are appended to `bad`; the generator returns the
    number of runs checked."""
    from mcsim._abi import Wall
    bad = [] if bad is None else bad
    nst = res["nst"]
    Hb = zf32(res["Hb"]).reshape(-1, 180)
    Eb = zf32(res["Eb"]).reshape(-1, 216)
    pre = zf64(res["pre"]).reshape(-1, 6)
    post = zf64(res["post"]).reshape(-1, len(ENGLOG_POST) + nst)
    ev = zf64(res["ev"]).reshape(-1, 3)
    evoff = np.concatenate([[0], np.cumsum(res["evn"])]).astype(int)
    takes = zf64(res["takes"]).reshape(-1, nst)
    cis = f64(res["cis"]).reshape(-1, 3)
    sl = zf64(res["sl"])
    rg = zf64(res["rg"]).reshape(-1, 24)
    r = ie = ir = it = ic = 0
    flat = True
    iex = [ENGLOG_POST.index(k) for k in ENGLOG_EXACT]
    for op in res["ops"]:
        kind = op[0]
        if kind == "cfg":
            _, N, R, hd, typ, fl, nl, ro, rph, S, fnMax, mxs, mys, mat, derived = op
            flat = typ == "flat"
            cell = make_cell(dict(N=N, R=R, helixDeg=hd, type=typ, fluteLen=fl, runout=ro, runoutPhase=rph),
                             mxs, mys, dict(zip(MAT_ORDER, mat)), S, (0.0, 0.0), 0.0, 1, nl)
            info = eng.configure(cell, fnMax)
            got = [info.stepsPerTooth, info.stepsPerRev, info.decim, info.piStep, info.maxSlices, info.dz,
                   info.dt, info.omega]
            if not _same(got, derived):
                bad.append(("configure", got, derived))
        elif kind == "wear":
            eng.set_wear(op[1], op[2])
        elif kind == "work":
            eng.set_work(Wall(*op[1:]) if len(op) > 1 else None)
        elif kind == "eng":
            _, ux, uy, fz, tipZ, ap, rat, nz = op
            a = sl[ie:ie + 5 * nz].reshape(5, nz) if nz else np.zeros((5, 0))
            ie += 5 * nz
            got = eng.set_engagement(ux, uy, fz, tipZ, ap, a[0] if rat else None, a[1] if rat else None)
            if got != nz:
                bad.append(("setEngagement nz", got, nz))
            flat = flat and not rat
        elif kind == "end":
            eng.set_end_engagement(op[1], op[2], None if flat else rg[ir])
            ir += 1
        elif kind == "run":
            if max_runs is not None and r >= max_runs:
                break
            eng.set_maps(Hb[r], Eb[r])
            x, y, vx, vy, w, wv = pre[r]
            eng.set_state(x=x, y=y, vx=vx, vy=vy, w=w, wv=wv)
            yield op[1]
            s = eng.state()
            got = np.array([float(getattr(s, k)) for k in ENGLOG_POST])
            want = post[r, :len(ENGLOG_POST)]
            if not (_same(got[iex], want[iex]) and all(_same(g, w_, rtol) for g, w_ in zip(got, want))):
                bad.append(("run %d state" % r, [(k, g, w_) for k, g, w_ in zip(ENGLOG_POST, got, want)
                                                 if not _same(g, w_, rtol)]))
            st = eng.stats()
            gs = [getattr(st, f) for f, _ in st._fields_]
            if not all(_same(g, w_, rtol) for g, w_ in zip(gs, post[r, len(ENGLOG_POST):])):
                bad.append(("run %d stats" % r, [(f, g, w_) for (f, _), g, w_ in zip(st._fields_, gs, post[r, len(ENGLOG_POST):])
                                                 if not _same(g, w_, rtol)]))
            evs = eng.chip_events()
            if not _same(evs[:, :3] if len(evs) else np.zeros((0, 3)), ev[evoff[r]:evoff[r + 1]], rtol):
                bad.append(("run %d chip events" % r, len(evs), int(evoff[r + 1] - evoff[r])))
            r += 1
        elif kind == "take":
            st = eng.stats(take=True)
            if not all(_same(getattr(st, f), w_, rtol) for (f, _), w_ in zip(st._fields_, takes[it])):
                bad.append(("takeStats %d" % it,))
            it += 1
        elif kind == "ci":
            if not _same(eng.chatter_index(), cis[ic], rtol):
                bad.append(("chatterIndex %d" % ic, list(eng.chatter_index()), list(cis[ic])))
            ic += 1
        if len(bad) > 8:
            break
    if max_runs is None and not bad:
        h, ts, ws, wsp = eng.rings()
        fin = res["final"]
        if not _same(h, zf32(fin["hist"]), rtol * 1e3):
            bad.append(("hist ring",))
        for name, got in (("ts", ts), ("ws", ws), ("wsp", wsp)):
            want = np.concatenate([f32(fin[name]["x"]), f32(fin[name]["y"])])
            if not _same(got, want, rtol * 1e3):
                bad.append((name + " ring",))
    return r


def replay_englog(eng, res, max_runs=None, rtol=0.0):
    """replay one recorded call sequence on `eng` (see englog_replayer); returns the list of
    mismatches (empty when the engine reproduces the reference) and the number of runs"""
    bad = []
    g = englog_replayer(eng, res, max_runs, rtol, bad)
    try:
        n = next(g)
        while True:
            eng.run(n)
            n = g.send(None)
    except StopIteration as stop:
        return bad, stop.value


def simx_program(res):
    """the sub-step schedule of a recorded rigid-workpiece simulator run (golden_simx.json)"""
    from mcsim import path
    P = res["path"]
    x, y, z = f64(P["x"]), f64(P["y"]), f64(P["z"])
    return path.sim_substeps(x, y, z, P["rapid"], P["line"], P["F"], P["S"], res["toolN"], res["hx"])


def simx_cell(res, j=0):
    """the engine's configure arguments recorded in a simx run (its j-th configure call)"""
    op = res["cfgs"][j]
    _, N, R, hd, typ, fl, nl, ro, rph, S, fnMax, mxs, mys, mat = op
    return make_cell(dict(N=N, R=R, helixDeg=hd, type=typ, fluteLen=fl, runout=ro, runoutPhase=rph),
                     mxs, mys, dict(zip(MAT_ORDER, mat)), S, (0.0, 0.0), 0.0, 1, nl), fnMax


def simx_edge(case):
    """the imported edge (dict(r, dz), radii as the float32 the reference holds) of a simx
    run's profile tool, else None"""
    e = case["tool"].get("edge")
    if e is None:
        return None
    return dict(r=np.asarray(e["r"], dtype=np.float32).astype(np.float64), dz=float(e["dz"]))


def simx_tool(case, res, tip_lut):
    """(type, lut, cornerR) of a simx run's tool for mc_sim_run / the oracle's loop; tip_lut is
    the library's or the oracle's tipLut"""
    from mcsim._abi import MC_TOOL_BALL, MC_TOOL_BULL, MC_TOOL_FLAT, MC_TOOL_PROFILE
    t = res["type"]
    if t == "bull":
        rc = case["tool"]["cornerR"]
        return MC_TOOL_BULL, tip_lut(MC_TIP_BULL, 2 * res["toolR"], rc), rc
    if t == "profile":
        e = simx_edge(case)
        return MC_TOOL_PROFILE, tip_lut(MC_TIP_EDGE, 2 * res["toolR"], 0.0, e["r"], e["dz"]), 0.0
    return (MC_TOOL_BALL if t == "ball" else MC_TOOL_FLAT), None, 0.0


def display_arcs(out):
    """the simulator's display arc after each sub-step (src/sim.js:421-423), from the
    pipeline's per-sub-step arc: [0, pi] until a sub-step sets one (the reset display,
    src/sim.js:734), then the last one set; flattened (phiSt, phiEx) pairs"""
    st, ex, arcs = 0.0, 3.141592653589793, []
    for o in out:
        if o.arc:
            st, ex = o.phiSt, o.phiEx
        arcs += [st, ex]
    return np.array(arcs)


def grid_limits():
    """The reference's timeDomainLimit over the whole 256-speed x 64-tool lobe grid
    (golden_tdl_grid.json, cells in mcsim.workloads.sld_grid() order): dict of arrays
    alim, lo, hi (+inf for open cells) and open, plus the options and Node's wall time."""
    g = load("tdl_grid")
    out = dict(alim=f64(g["alim"]), lo=f64(g["lo"]), hi=f64(g["hi"]),
               open=np.frombuffer(base64.b64decode(g["open"]), dtype=np.uint8).astype(bool))
    out.update(opts=g["opts"], n=g["n"], node_s=g["node_s"], procs=g["procs"])
    return out


# Cells of the lobe grid whose answer differs from the reference's because V8 7.8's Math.pow
# (Node 12) and the fdlibm e_pow restated by the oracle and the kernels differ in the last
# bit on a few % of inputs (the size effect (0.05/h)^mc of src/cutting.js:197): found by
# running the oracle over all 16,384 cells against golden_tdl_grid.json.  14077 ends one
# bisection step (one bracket) from the reference and 13924 two; 5395 is stable at aMax for the oracle (open) and chatters
# there for the reference (alim 1.856 mm).  The GPU must give the oracle's answers on them.
POW_ULP_CELLS = (5395, 13924, 14077)


def report_schedule(res, case):
    """the real-time frame schedule of a recorded run (golden_report.json): sub-steps of
    sim.step(dtReal) frames at the recorded playback speed"""
    from mcsim import path
    P = res["path"]
    return path.sim_frames(f64(P["x"]), f64(P["y"]), f64(P["z"]), P["rapid"], P["line"], P["F"], P["S"], res["toolN"],
                           res["hx"], case["dtReal"], res["speed"], res["feedOverride"])


def run_report(case, res, mk_stock, mk_engine, run_frame, tip_lut, tan):
    """replay a recorded real-time run frame by frame: each frame is one call of the sub-step
    loop with the frame's flank wear (the setWear the simulator issued, and the radial wear
    rw = min(R/2, VB tan(clearance)) of src/sim.js:439 from the wear VB it had) and the feed
    mark carried over; the report is built from the outputs.  Where the simulator re-configured
    its engine inside a frame (a spindle-speed change, ensureEngine, src/sim.js:252-257,382),
    the frame is two calls: the first ends before that sub-step, the engine is configured with
    the recorded arguments, and the second call starts with the frame budget the first left.
    run_frame(act, ide, eng, sub, count, R, fluteLen, rw, typ, lut, rc, dynBudget, fz0,
    budget_left) returns (outs, [frame] or []).
    Returns (outs, frames, reporter, stocks, sub, n, frame list)."""
    import ctypes as C
    from mcsim import report as RP
    from mcsim._abi import Substep
    st = case["stock"]
    act, ide = mk_stock(st["w"], st["l"], st["h"], case["cell"]), mk_stock(st["w"], st["l"], st["h"], case["cell"])
    ide.set_wall_band(2.0)
    eng = mk_engine()
    cell, fnMax = simx_cell(res)
    eng.configure(cell, fnMax)
    sub, n, frames = report_schedule(res, case)
    typ, lut, rc = simx_tool(case, res, tip_lut)
    R = res["toolR"]
    clr = float(tan(res["clearanceDeg"] * 3.141592653589793 / 180))
    rep = RP.FrameReporter()
    outs, fouts, fz0 = [], [], 0.0
    wears = res["wears"]
    recfg = [op[0] for op in res["cfgs"]]   # sub-step index of each configure call
    for f, (first, count) in enumerate(frames):
        w = [x for x in wears if x[0] <= first][-1] if count else wears[0]
        eng.set_wear(w[1], w[2])
        VB = res["frames"][f]["VB0"]
        rw = min(0.5 * R, VB * clr)
        if count:
            # the frame's calls: a new call wherever the engine was re-configured inside it
            cuts = sorted(set([first, first + count] + [c for c in recfg[1:] if first < c < first + count]))
            left = [float(res["dynBudget"])] if len(cuts) > 2 else None
            o, fo = [], []
            for a, b in zip(cuts[:-1], cuts[1:]):
                if a > 0 and a in recfg:   # the last configure call the simulator made at sub-step a
                    j = max(q for q, c in enumerate(recfg) if c == a)
                    cell_j, fn_j = simx_cell(res, j)
                    eng.configure(cell_j, fn_j)
                    eng.set_wear(w[1], w[2])
                part = (Substep * (b - a)).from_address(C.addressof(sub[a]))
                o2, fo2 = run_frame(act, ide, eng, part, b - a, R, res["fluteLen"], rw, typ, lut, rc, res["dynBudget"],
                                    fz0, left)
                for k in range(b - a):
                    if o2[k].cut:
                        fz0 = part[k].fzEff
                o += list(o2)
                fo += list(fo2)
            assert len(fo) == 1
            outs += o
            fouts.append(fo[0])
            part = (Substep * count).from_address(C.addressof(sub[first]))
            cut_time = rep.substeps(part, o, 0, count)
            rep.collect(fo[0].stats, (fo[0].index, fo[0].amp, fo[0].once), cut_time)
        else:
            fouts.append(None)
            rep.collect(RP._EMPTY, (0.0, 0.0, 0.0), 0.0)
    return outs, fouts, rep, (act, ide), sub, n, frames


REPORT_STATS = ("n", "sumF", "maxF", "sumFx", "sumFy", "sumFz", "maxFz", "sumT", "maxT", "sumX", "sumY", "minX",
                "maxX", "minY", "maxY", "sumZc", "wZc", "sumW", "maxW", "sumHb", "sumB", "sumRb", "engTeeth")


def check_report_run(case, res, outs, fouts, rep, stocks, sub, n, frames, maps):
    """compare a replayed real-time run with the recorded one: every sub-step and stamp,
    every frame's taken stats, chatter index and state, the final maps, and the report rows
    (all fields but the temperature, its swing and the tool stress, which come from the
    simulator's thermal and stress models)"""
    a = res["actual"]
    assert n == len(a["type"]) == len(res["ideal"]["type"])
    assert np.array_equal([sub[k].x for k in range(n)], f64(res["ideal"]["cx"]))
    assert np.array_equal([sub[k].x + outs[k].dx for k in range(n)], f64(a["cx"]))
    assert np.array_equal([sub[k].y + outs[k].dy for k in range(n)], f64(a["cy"]))
    assert np.array_equal([o.errVal for o in outs], f64(a["errVal"]))
    assert np.array_equal(f64(a["R"]), f64(a["R"]))
    runs = dict((k, m) for k, m in res["runs"])
    assert [o.nsteps for o in outs] == [runs.get(k, 0) for k in range(n)]
    assert np.array_equal(display_arcs(outs), f64(res["darc"]))
    rf = res["frames"]
    assert len(frames) == len(rf)
    ends = np.cumsum([c for _, c in frames])
    assert list(ends) == [f["end"] for f in rf]
    for f, (fo, want) in enumerate(zip(fouts, rf)):
        if fo is None:
            assert want["stats"][0] == 0
            continue
        got = [getattr(fo.stats, k) for k in REPORT_STATS]
        # JSON keeps no infinities: the untouched extremes of newStats() come back as null
        ref = [v if v is not None else (np.inf if k in ("minX", "minY") else -np.inf)
               for k, v in zip(REPORT_STATS, want["stats"])]
        assert np.array_equal(np.array(got), np.array(ref)), f
        if want["ci"] is not None:
            assert (fo.index, fo.amp) == (want["ci"][0], want["ci"][1]), f
            assert fo.once == (want["ci"][2] if want["ci"][2] is not None else 0.0), f
    assert rep.states == [f["state"] for f in rf]
    got_rows = rep.report.pack()
    want_rows = res["report"]["rows"]
    assert [r[0] for r in got_rows] == [r[0] for r in want_rows]
    from mcsim.report import FIELDS
    skip = {FIELDS.index(k) + 1 for k in ("temp", "stress", "swing")}
    for g, w in zip(got_rows, want_rows):
        for i in range(1, len(g)):
            if i not in skip:
                assert g[i] == (w[i] if w[i] is not None else 0), (g[0], FIELDS[i - 1], g[i], w[i])
    act, ide = stocks
    for name, s_, ref in (("actual", act, res["stock"]), ("commanded", ide, res["commanded"])):
        got = maps(s_)
        for key in ("H", "err", "wallD", "markS", "markF"):
            assert np.array_equal(got[key], zf32(ref[key]), equal_nan=True), (name, key)
        for key in ("lineCut", "lineWall"):
            assert np.array_equal(got[key], zi32(ref[key])), (name, key)
