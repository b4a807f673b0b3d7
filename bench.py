#!/usr/bin/env python3
"""Benchmark of the two hot paths on one MI355X.

A step is one pass of the time-domain stability search (timeDomainLimit,
{aMin 0.05, aMax 10, iters 10}) over every cell of the 256-speed x 64-tool lobe grid
(16,384 cells, 2-4 flutes, 1-4 modes per axis, 20-80 revolutions per trial, process
damping and size effect on).  It is timed twice: plain, and guided by the learned
surrogate (trained first, untimed, on 262,144 cells of other seeded tools; its prediction
only steers which bisection nodes are speculated, so the results are the plain search's
own, checked cell by cell).  value = SLD cells resolved per second by the faster of the
two; both are in config.  The +-15 % bracket form of the surrogate is reported with its
error against plain bisection.  The heightmap path (512 x 512 at 0.1 mm, 5,000 seeded
zigzag segments = 2.41 M stamps) is reported beside it in GB/s, and the simulator's coupled
sub-step loop (16 seeded raster programs, each on its own 2048 x 2048 stock) in sub-steps/s.

Prints ONE JSON line.  usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "machinig-sim_amd"))

import numpy as np  # noqa: E402

from mcsim import lib, workloads as W, path  # noqa: E402
from mcsim._abi import cell_array, MC_TOOL_BALL, MC_FRAME_LAST  # noqa: E402

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector, vendor figure
DYN_BUDGET = 220000.0     # the simulator's default dynamics budget per frame (src/sim.js:50)
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, vendor figure


def flop_alg(cells, res):
    """FP64 flops of the dynamic (non-revolution-periodic) part of the search:
    6 per visit (edge position, chip thickness), 18 per engaged visit (forces and their
    accumulation), +5 with process damping, +1 with the size effect, 11 per mode and 2
    per step (modal update)."""
    tot = 0.0
    for c, r in zip(cells, res):
        per_eng = 18 + (5 if c.nonlinear and c.mat.Ci != 0 else 0) + (1 if c.nonlinear and c.mat.mc != 0 else 0)
        tot += 6.0 * r.visits + per_eng * r.engaged + (11.0 * (c.nmx + c.nmy) + 2) * r.steps
    return tot


def reference_limits(workload):
    """The reference's own answers on the default lobe grid (tests/golden/golden_tdl_grid.json,
    every cell resolved by its timeDomainLimit under Node in the build container), or None
    for another workload."""
    if workload != "sld_grid_256speeds_x_64tools":
        return None
    import base64
    try:
        with open(os.path.join(ROOT, "tests", "golden", "golden_tdl_grid.json")) as f:
            g = json.load(f)
    except OSError:
        return None
    arr = lambda k: np.frombuffer(base64.b64decode(g[k]), dtype="<f8")
    return dict(alim=arr("alim"), lo=arr("lo"), hi=arr("hi"), node_s=g["node_s"], procs=g["procs"], node=g["node"])


def same_limits(res, ref):
    """cells whose (alim, lo, hi) differ from the reference's"""
    alim = np.array([r.alim for r in res]); lo = np.array([r.lo for r in res]); hi = np.array([r.hi for r in res])
    return int(np.count_nonzero((alim != ref["alim"]) | (lo != ref["lo"]) | (hi != ref["hi"])))


def reference_differences(res, ref):
    """the cells whose limits differ from the reference's: the cell, both limits and brackets,
    and the distance between the limits in widths of the reference's final bracket (one
    bisection step apart reads about 1)"""
    out = []
    for i, r in enumerate(res):
        if (r.alim, r.lo, r.hi) != (ref["alim"][i], ref["lo"][i], ref["hi"][i]):
            w = ref["hi"][i] - ref["lo"][i]
            out.append(dict(cell=i, alim=r.alim, lo=r.lo, hi=r.hi, ref_alim=float(ref["alim"][i]),
                            ref_lo=float(ref["lo"][i]), ref_hi=float(ref["hi"][i]),
                            brackets_apart=float(abs(r.alim - ref["alim"][i]) / w) if w > 0 else None))
    return out


def cpu_baseline(cells, opts, target_s, gpu_res=None):
    """The oracle (FP64 CPU restatement) on a bounded sample of the same cells: a seeded
    shuffle of the grid, run in chunks of 8 cells per thread (one cell per thread at a
    time) until target_s seconds have elapsed.  With gpu_res, each sampled cell's answer is
    compared with the GPU's (alim, bracket, trials, steps, visits)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 64))
    order = np.random.RandomState(12345).permutation(len(cells))
    done, dt, pos, diff = 0, 0.0, 0, 0
    chunk = 8 * threads
    while dt < target_s and pos < len(order):
        idx = order[pos:pos + chunk]
        pos += len(idx)
        arr = cell_array([cells[i] for i in idx])
        t0 = time.perf_counter()
        got = oracle.bisect_batch(arr, len(idx), threads=threads, **opts)
        dt += time.perf_counter() - t0
        done += len(idx)
        if gpu_res is not None:
            diff += sum(1 for i, o in zip(idx, got)
                        if (o.alim, o.lo, o.hi, o.open, o.trials, o.steps, o.visits) !=
                        (gpu_res[i].alim, gpu_res[i].lo, gpu_res[i].hi, gpu_res[i].open, gpu_res[i].trials,
                         gpu_res[i].steps, gpu_res[i].visits))
    out = dict(value=done / dt, unit="SLD cells/s", cores=threads, kind="port",
               sample="%d of %d cells (seeded shuffle), %.1f s on %d threads" % (done, len(cells), dt, threads))
    if gpu_res is not None:
        out["cells_differing_from_gpu"] = diff
    return out


def calibrate_copy(nbytes):
    """A plain device-to-device copy of nbytes (for scaling rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters to bytes)."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))


def profiled_traffic(kernel):
    """HBM bytes per launch of a kernel, from the committed rocprofv3 counter summary."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)[kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def map_digests(stock):
    """SHA-256 of each map of a library stock (the form tests/golden/golden_zigzag.json holds)"""
    import hashlib
    err, wallD, lineCut, lineWall = stock.side()
    markS, markF = stock.marks()
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    return dict(H=sha(stock.H()), err=sha(err), wallD=sha(wallD), lineCut=sha(lineCut), lineWall=sha(lineWall),
                markS=sha(markS), markF=sha(markF))


def heightmap(steps, cpu=True, cpu_seconds=10.0):
    """Config 1: the seeded 5,000-segment zigzag (flat R5 / ball R3 layers) on the 512 x 512
    stock at 0.1 mm, stamped every min(hx, 0.4) * 0.6 mm along the path (2.41 M stamps), in
    one batch with the config's outputs: the final maps and the removed volume per segment
    (mc_stamp_batch_lines).  Checked against the reference's own answer for the whole batch
    (the SHA-256 of every final map and the total removed volume, tests/golden/golden_zigzag.json).

    Roofline (HBM): the algorithmic bytes are those any implementation must move once: the
    stamp records in (48 B each), the five touched maps in and out (H, err, wallD, lineCut,
    lineWall: 40 B per node) and the per-segment volumes out.  The reference's own loop reads
    and writes 8 B per bounding-box node and stamp (bbox_bytes); the tiles' exact culling
    skips most of those, so that figure is reported beside the roofline, not as it.
    cpu_baseline: the oracle (oracle/stock.c) stamping a prefix of the same batch on one host
    thread for about cpu_seconds, its maps compared with the GPU's on the same prefix; the
    reference's Node time for the whole batch (measured where the fixture was written) is
    reported beside it."""
    import json as _json
    segs = W.zigzag_segments(5000)
    rec, _ = path.schedule(segs, 0.1)
    arr = path.as_stamp_array(rec)
    n = len(rec)
    nlines = int(rec["line"].max()) + 1
    nodes = 512 * 512
    alg_bytes = 48.0 * n + 40.0 * nodes + 8.0 * nlines
    hx = 0.1
    Rw = rec["R"] + 2 * hx
    i0 = np.maximum(0, np.floor((rec["cx"] - Rw) / hx)); i1 = np.minimum(511, np.ceil((rec["cx"] + Rw) / hx))
    j0 = np.maximum(0, np.floor((rec["cy"] - Rw) / hx)); j1 = np.minimum(511, np.ceil((rec["cy"] + Rw) / hx))
    bbox_bytes = 8.0 * np.where((i1 >= i0) & (j1 >= j0), (i1 - i0 + 1) * (j1 - j0 + 1), 0).sum()
    times, kms = [], []
    lv = dg = None
    for _ in range(steps + 1):
        s = lib.Stock(51.1, 51.1, 20, 0.1)
        t0 = time.perf_counter()
        lv = s.stamps_lines(arr, nlines)
        times.append(time.perf_counter() - t0)
        kms.append(lib.last_kernel_ms())
        dg = map_digests(s)
        s.close()
    wall = float(np.mean(times[1:]))
    k = float(np.mean(kms[1:]))
    # the drop-in path: the same batch with a volume per stamp (what Stock.stamp returns)
    s = lib.Stock(51.1, 51.1, 20, 0.1)
    t0 = time.perf_counter()
    vs = s.stamps(arr)
    t_stamp = time.perf_counter() - t0
    k_stamp = lib.last_kernel_ms()
    dg_stamp = map_digests(s)
    s.close()
    per_line = np.bincount(rec["line"], weights=vs, minlength=nlines)
    line_rel = float(np.max(np.abs(lv - per_line) / np.maximum(1e-30, np.abs(per_line))))
    with open(os.path.join(ROOT, "tests", "golden", "golden_zigzag.json")) as f:
        g = _json.load(f)
    ident_ref = all(dg[key] == g[key] and dg_stamp[key] == g[key] for key in dg)
    vol_rel = abs(float(lv.sum()) - g["removed"]) / g["removed"]
    gbps = alg_bytes / (k * 1e-3) / 1e9
    out = dict(workload="zigzag_512x512_5000seg", stamps=n, segments=nlines - 1, ms_per_batch=wall * 1e3,
               kernel_ms=k, stamps_per_s=n / wall, gb_per_s=gbps,
               per_stamp_volumes={"ms_per_batch": t_stamp * 1e3, "kernel_ms": k_stamp, "stamps_per_s": n / t_stamp},
               roofline={"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbps / HBM_PEAK_GBS, "traffic": profiled_traffic("stamp_tiles"),
                         "kernel": "stamp_tiles", "alg_bytes": alg_bytes,
                         "alg_def": "48 B per stamp record in + 40 B per node (H, err, wallD, lineCut, lineWall "
                                    "in and out) + 8 B per segment volume out",
                         "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)"},
               bbox_bytes=bbox_bytes, bbox_gb_per_s=bbox_bytes / (k * 1e-3) / 1e9,
               identical_to_reference=ident_ref, removed_rel_diff_vs_reference=vol_rel,
               segment_volumes_rel_diff_vs_per_stamp=line_rel,
               reference_node={"stamps_per_s": n / (g["node_ms"] * 1e-3), "seconds": g["node_ms"] * 1e-3,
                               "sample": "the whole batch by the reference's Stock.stamp under Node 12, one thread, "
                                         "in the build container (tests/golden/golden_zigzag.json)"})
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        o = oracle.OracleStock(51.1, 51.1, 20, 0.1)
        done, step, t0 = 0, 20000, time.perf_counter()
        while done < n and time.perf_counter() - t0 < cpu_seconds:
            b = min(n, done + step)
            o.stamps(path.as_stamp_array(rec[done:b]), b - done)
            done = b
        cw = time.perf_counter() - t0
        s = lib.Stock(51.1, 51.1, 20, 0.1)
        s.stamps_lines(path.as_stamp_array(rec[:done]), nlines)
        same = bool(np.array_equal(s.H(), o.H()) and np.array_equal(s.side()[0], o.err()) and
                    np.array_equal(s.side()[2], o.lineCut()))
        s.close()
        rate = done / cw
        out["cpu_baseline"] = dict(value=rate, unit="stamps/s", cores=1, kind="port",
                                   gb_per_s=rate * alg_bytes / n / 1e9,
                                   sample="the first %d of the %d stamps on one host thread (oracle/stock.c), %.1f s"
                                          % (done, n, cw), identical_to_gpu=same)
        out["speedup_vs_cpu"] = out["stamps_per_s"] / rate
    return out


def zigzag_forces(n_sub, cpu):
    """The zigzag workload's forces: its first n_sub sub-steps (the first layer with the flat
    10 mm end mill and the start of the ball 6 mm layer, 512 x 512 stock at 0.1 mm) through
    the simulator's sub-step loop (mc_sim_run: engagement, engine run, both stamps), one
    program, the rigid mode of the reference's target test; per-segment maximum force
    |(Fx, Fy)| from each sub-step's run statistics (src/cutting.js:280-281).  One program is one
    serial chain (each sub-step reads the stock the previous one cut), so the rate is the
    chain's, set beside the oracle's loop on one host thread, whose results the device's
    equal (checked here)."""
    from mcsim._abi import make_cell
    layers = W.zigzag_layers()
    cell_of = lambda t: make_cell(t, W.ZIGZAG_MODES, None, W.MAT_A5052, W.ZIGZAG_SPEED, (0.0, 0.0), 0.0, 1, True)
    sub_of = lambda L: path.sim_substeps(L["x"], L["y"], L["z"], L["rapid"], L["line"], L["F"], L["S"],
                                         L["tool"]["N"], 0.1)
    typ = lambda tool: W.MC_TOOL_FLAT if tool["type"] == "flat" else W.MC_TOOL_BALL
    kms = []

    def run_dev(a, i, e, sub, n, tool, share):
        r = lib.sim_run([dict(actual=a, ideal=i, engine=e, sub=sub, nsub=n, R=tool["R"], fluteLen=tool["fluteLen"],
                              rw=0.0, share=share, type=typ(tool))])[0]
        kms.append(lib.last_kernel_ms())
        return r
    a, i, e = lib.Stock(51.1, 51.1, 20, 0.1), lib.Stock(51.1, 51.1, 20, 0.1), lib.Engine()
    i.set_wall_band(2.0)
    t0 = time.perf_counter()
    outs, lines = W.zigzag_coupled(layers, (a, i), e, run_dev, cell_of, sub_of, max_substeps=n_sub)
    wall = time.perf_counter() - t0
    seg = W.segment_max_force(outs, lines)
    vals = np.array(list(seg.values()))
    res = dict(workload="zigzag_512x512_coupled_first_%d_substeps" % len(outs), substeps=len(outs),
               segments=len(seg), engine_steps=int(sum(o.nsteps for o in outs)), wall_s=wall,
               kernel_ms=float(sum(kms)), substeps_per_s=len(outs) / wall,
               segment_max_force_N=dict(max=float(vals.max()), median=float(np.median(vals)),
                                        cutting_segments=int((vals > 0).sum())))
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        def run_cpu(a, i, e, sub, n, tool, share):
            return oracle.sim_run(a, i, e, sub, n, tool["R"], tool["fluteLen"], 0.0, share, typ(tool))
        ao, io = oracle.OracleStock(51.1, 51.1, 20, 0.1), oracle.OracleStock(51.1, 51.1, 20, 0.1)
        io.set_wall_band(2.0)
        t0 = time.perf_counter()
        o_outs, o_lines = W.zigzag_coupled(layers, (ao, io), oracle.OracleEngine(), run_cpu, cell_of, sub_of,
                                           max_substeps=n_sub)
        cw = time.perf_counter() - t0
        same = (o_lines == lines and W.segment_max_force(o_outs, o_lines) == seg and
                all((p.dx, p.dy, p.maxF, p.nsteps, p.ap) == (q.dx, q.dy, q.maxF, q.nsteps, q.ap)
                    for p, q in zip(outs, o_outs)) and np.array_equal(a.H(), ao.H()))
        res["cpu_baseline"] = dict(value=len(o_outs) / cw, unit="sub-steps/s", cores=1, kind="port",
                                   sample="the same %d sub-steps, oracle/simloop.c on one host thread, %.1f s"
                                          % (len(o_outs), cw))
        res["identical_to_oracle"] = bool(same)
    a.close(); i.close(); e.close()
    return res


def sld_40rpm(cpu):
    """Config 0, the reference's own CPU-runnable case: the 40-speed diagram of one tool
    (timeDomainLimit at 40 speeds from 3000 to 15000 rpm, src/stabtd.js:24-32) as one batched
    search, checked against the reference's limits (golden_tdl.json) and timed beside the
    oracle on the host (one thread per cell, and one thread for all)."""
    cells, opts = W.sld_40rpm()
    arr = cell_array(cells)
    res, kms, ws = None, [], []
    for _ in range(4):                    # the first is a warm-up
        t0 = time.perf_counter()
        res = lib.bisect_batch(arr, **opts)
        ws.append(time.perf_counter() - t0)
        kms.append(lib.last_kernel_ms())
    wall = float(np.mean(ws[1:]))
    out = dict(workload="sld_40rpm_3000_15000", cells=len(cells), wall_ms=1e3 * wall,
               kernel_ms=float(np.mean(kms[1:])), cells_per_s=len(cells) / wall)
    try:
        with open(os.path.join(ROOT, "tests", "golden", "golden_tdl.json")) as f:
            g = json.load(f)
        ref = [g["results"][i] for i in g["meta"]["sld40"]]
        out["identical_to_reference"] = all((r.alim, r.lo) == (q["alim"], q["lo"]) and
                                            (r.hi == q["hi"] or (q["hi"] == "inf" and r.open))
                                            for r, q in zip(res, ref))
    except OSError:
        pass
    # the reference's own timeDomainLimit on these 40 cells, one thread (Node 12, build container)
    out["reference_node"] = dict(seconds=6.7, cells_per_s=40 / 6.7, cores=1,
                                 sample="tests/golden/time_reference.py, in the build container")
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        for k in (1, threads):
            t0 = time.perf_counter()
            got = oracle.bisect_batch(arr, len(cells), threads=k, **opts)
            dt = time.perf_counter() - t0
            out["cpu_%d_threads" % k] = dict(value=len(cells) / dt, unit="SLD cells/s", cores=k, kind="port",
                                             identical_to_gpu=all((a.alim, a.lo, a.hi) == (b.alim, b.lo, b.hi)
                                                                  for a, b in zip(got, res)))
    return out


def dropin_substeps(cpu):
    """The drop-in path's cost per simulator sub-step: the JS modules called as src/sim.js
    calls them (machinig-sim_amd/js/dropin_substeps.mjs: one device call each for the
    engagement, the engine run and the two stamps, per sub-step) on the first recorded
    simulator program of golden_simx.json, beside the oracle's fused loop on one host thread
    on the same sub-steps, and the reference's own Simulator timed in the build container
    (tests/golden/time_reference.py sim)."""
    import subprocess
    js = os.path.join(ROOT, "machinig-sim_amd", "js", "dropin_substeps.mjs")
    golden = os.path.join(ROOT, "tests", "golden", "golden_simx.json")
    out = dict(workload="golden_simx_program0_dropin", reference_node=dict(
        us_per_substep=[196.1, 176.7, 126.7], cores=1,
        sample="the reference's Simulator on the same program under Node 12 in the build container, "
               "three warm runs (tests/golden/time_reference.py sim); the oracle's loop there: 45 us"))
    try:
        r = subprocess.run(["node", js, golden, "0", "3"], capture_output=True, text=True, timeout=120)
        if r.returncode == 0:
            out.update(json.loads(r.stdout.strip().splitlines()[-1]))
        else:
            out["error"] = r.stderr[-400:]
    except (OSError, subprocess.TimeoutExpired) as ex:
        out["error"] = str(ex)
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        from mcsim._abi import make_cell, MC_TOOL_BALL, MC_TOOL_FLAT
        with open(golden) as f:
            g = json.load(f)
        case, res = g["cases"][0], g["results"][0]
        P = res["path"]
        dec = lambda k: np.frombuffer(__import__("base64").b64decode(P[k]), dtype="<f8")
        sub, n = path.sim_substeps(dec("x"), dec("y"), dec("z"), P["rapid"], P["line"], P["F"], P["S"], res["toolN"],
                                   res["hx"])
        c = res["cfgs"][0]
        mat = dict(zip(("Ktc", "Krc", "Kte", "Kre", "Kac", "Kae", "mc", "Ci"), c[13]))
        cell = make_cell(dict(N=c[1], R=c[2], helixDeg=c[3], type=c[4], fluteLen=c[5], runout=c[7], runoutPhase=c[8]),
                         c[11], c[12], mat, c[9], (0.0, 0.0), 0.0, 1, c[6] != 0)
        st = case["stock"]
        ts = []
        for _ in range(3):
            a, i = oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"]), oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"])
            i.set_wall_band(2.0)
            e = oracle.OracleEngine()
            info = e.configure(cell, c[10])
            e.set_wear(res["wears"][0][1], res["wears"][0][2])
            t0 = time.perf_counter()
            oracle.sim_run(a, i, e, sub, n, res["toolR"], res["fluteLen"], 0.0, info.stepsPerTooth / 4,
                           MC_TOOL_BALL if res["type"] == "ball" else MC_TOOL_FLAT)
            ts.append(1e6 * (time.perf_counter() - t0) / n)
        out["cpu_oracle_us_per_substep"] = ts
    return out


def line_cells(progs, outs, frs, tools, revs=40, ckeys=None):
    """The machining report's per-line stability question: each program's report is built
    from the pipeline's frames (mcsim.report: Simulator.collect and LineReport,
    src/sim.js:644-710, src/report.js:9-33), and every line with an engagement arc asks for
    its limit at its heaviest cut's arc and the line's speed (src/main.js:1033-1038), here
    by the time-domain search.  Returns the cells, (program, line, ap) keys and the per-frame
    chatter states of every program; with a list ckeys, each cell's (tool, arc, feed) identity
    is appended to it (cells of equal identity are the same search)."""
    from mcsim._abi import make_cell
    from mcsim import report as RP
    cells, keys, states = [], [], []
    for p, ((sub, n, frames), out, fr, tool) in enumerate(zip(progs, outs, frs, tools)):
        rep = RP.report_program(sub, out, frames, fr)
        states.append(rep.states)
        if ckeys is not None:
            print("coupled: report of program %d built" % p, file=sys.stderr, flush=True)
        for ln, ap, st, ex, fz in RP.line_engagements(rep.report, sub, out):
            cells.append(make_cell(tool, W.COUPLED_MODES, None, W.MAT_A5052, W.COUPLED_SPEED, (st, ex), fz, revs, True))
            keys.append((p, ln, ap))
            if ckeys is not None:
                ckeys.append((tool["type"], st, ex, fz))
    return cells, keys, states


def coupled(n_prog, n_lines, cpu_threads, cpu, opts, surrogate=None, full_segments=None, cpu_prefix=None):
    """The simulator's coupled sub-step loop (engagement -> engine run -> commanded and
    deflected stamps, src/sim.js:351-449) for n_prog seeded raster programs, each on its own
    2048 x 2048 stock at 0.05 mm (102.35 mm square), flat 12 mm and ball 8 mm tools taking
    turns, two modes per axis, 3 to 6 passes (workloads.coupled_plan), one launch of
    mc_sim_run (one workgroup per program), in the simulator's frames of 4000 loop passes
    (sim.step with an unbounded machine-time budget), each frame's engine stats taken and its
    chatter index computed on the device.  Reported in sub-steps per second; the oracle's
    restated loop runs the same programs on host threads as the CPU baseline and checks the
    device's deflections, step counts, depths, forces, frames and final heights.  Then the
    report (mcsim.report) is built from the frames and every line's stability check
    (line_cells) runs as one batched time-domain search, guided by the surrogate when one was
    trained (checked identical to the plain search).  full_segments: config 4 at full size
    instead, programs of that many moves (workloads.coupled_full_program: a zigzag pocket, then
    chorded contour passes); the lines' checks then run once per distinct (tool, arc, feed),
    and the oracle's CPU baseline runs the first cpu_prefix sub-steps of every program."""
    from mcsim._abi import make_cell, MC_TOOL_FLAT, MC_TOOL_BALL
    width, cell = 102.35, 0.05
    hx = width / 2047
    progs, tools, passes, nseg = [], [], [], []
    for p in range(n_prog):
        tool, m = W.coupled_plan(p, n_lines)
        if full_segments:
            x, y, z, r, ln, F, S = W.coupled_full_program(p + 1, width, full_segments, speed=W.COUPLED_SPEED)
        else:
            x, y, z, r, ln, F, S = W.coupled_program(p + 1, width, m, speed=W.COUPLED_SPEED)
        nseg.append(len(x) - 1)
        # frames of 4000 loop passes with an unbounded machine-time budget (sim.step(1e6)): the
        # frame budget's share is then stepsPerTooth / 4 for every cutting sub-step
        progs.append(path.sim_frames(x, y, z, r, ln, F, S, tool["N"], hx, 1e6))
        tools.append(tool)
        passes.append(m)
    ecells = [make_cell(t, W.COUPLED_MODES, None, W.MAT_A5052, W.COUPLED_SPEED, (0.0, 0.0), 0.0, 1, True) for t in tools]
    ttype = [MC_TOOL_BALL if t["type"] == "ball" else MC_TOOL_FLAT for t in tools]
    kte, kre = W.MAT_A5052["Kte"], W.MAT_A5052["Kre"]

    def setup(limit=None):
        P = []
        for (sub, n, _), t, ec, ty in zip(progs, tools, ecells, ttype):
            act, ide = lib.Stock(width, width, 20, cell), lib.Stock(width, width, 20, cell)
            ide.set_wall_band(2.0)
            eng = lib.Engine()
            eng.configure(ec)
            eng.set_wear(kte, kre)
            P.append(dict(actual=act, ideal=ide, engine=eng, sub=sub, nsub=n if limit is None else min(n, limit),
                          R=t["R"], fluteLen=t["fluteLen"], rw=0.0, type=ty, dynBudget=DYN_BUDGET))
        return P

    def release(P):
        for g in P:
            g["actual"].close(); g["ideal"].close(); g["engine"].close()

    W0 = setup(400)
    lib.sim_run(W0)          # warm-up launch (code objects, allocations) on the programs' first sub-steps
    release(W0)
    P = setup()
    print("coupled: %d programs, %d sub-steps, one launch" % (n_prog, sum(n for _, n, _ in progs)), file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    outs, frs = lib.sim_run(P, frames=True)
    wall = time.perf_counter() - t0
    kms = lib.last_kernel_ms()
    nsub = sum(n for _, n, _ in progs)
    steps = sum(o.nsteps for out in outs for o in out)
    cuts = sum(o.cut for out in outs for o in out)
    # algorithmic bytes: both stamps of every sub-step read and write the f32 height of each
    # node of their bounding boxes (the commanded stock's wall band is 2 mm)
    nodes = 0
    for (sub, n, _), t in zip(progs, tools):
        sa = np.ctypeslib.as_array(sub)[:n]
        cxs, cys = sa["x"].astype(np.float64), sa["y"].astype(np.float64)
        for Rw in (t["R"] + 2 * hx, t["R"] + 2.0):
            i0 = np.maximum(0, np.floor((cxs - Rw) / hx)); i1 = np.minimum(2047, np.ceil((cxs + Rw) / hx))
            j0 = np.maximum(0, np.floor((cys - Rw) / hx)); j1 = np.minimum(2047, np.ceil((cys + Rw) / hx))
            nodes += np.where((i1 >= i0) & (j1 >= j0), (i1 - i0 + 1) * (j1 - j0 + 1), 0).sum()
    wl = ("coupled_%dprog_2048x2048_%dseg_zigzag_contour_flat12_ball8" % (n_prog, full_segments) if full_segments
          else "coupled_%dprog_2048x2048_%dto%dpasses_flat12_ball8" % (n_prog, min(passes), max(passes)))
    res = dict(workload=wl, programs=n_prog, passes=None if full_segments else passes,
               segments=nseg if full_segments else None,
               substeps=nsub, cutting_substeps=cuts, engine_steps=steps, wall_s=wall, kernel_ms=kms,
               substeps_per_s=nsub / wall, engine_steps_per_s=steps / wall,
               roofline={"bound": "hbm", "achieved": 8.0 * nodes / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": 8.0 * nodes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "kernel": "sim_run", "traffic": None if full_segments else profiled_traffic("sim_run")},
               note="algorithmic bytes = 8 B x bbox nodes of both stamps per sub-step; one workgroup per "
                    "program runs its sub-steps in order (each reads the stock the previous one cut), so "
                    "the kernel is bound by the latency of that chain, far below the HBM roof")
    print("coupled: kernel %.1f s; building the reports" % (kms / 1e3), file=sys.stderr, flush=True)
    ck = [] if full_segments else None
    lc, keys, states = line_cells(progs, outs, frs, tools, ckeys=ck)
    which = list(range(len(lc)))
    if full_segments:   # one search per distinct cell; every line reads its cell's result
        index, uniq = {}, []
        for i, key in enumerate(ck):
            if key not in index:
                index[key] = len(uniq)
                uniq.append(lc[i])
            which[i] = index[key]
        lines_total, lc = len(keys), uniq
    res["frames"] = sum(len(f) for f in frs)
    res["frame_states"] = {k: sum(st.count(k) for st in states) for k in ("stable", "marginal", "chatter", "idle")}
    t0 = time.perf_counter()
    plain = lib.bisect_batch(lc, **opts)
    t_plain = time.perf_counter() - t0
    checks = dict(lines=len(keys), plain_s=t_plain, plain_kernel_ms=lib.last_kernel_ms(),
                  lines_over_limit=sum(1 for (p, ln, ap), w in zip(keys, which) if ap > plain[w].alim),
                  alim_mm_min=min(r.alim for r in plain), alim_mm_max=max(r.alim for r in plain))
    if full_segments:
        checks["distinct_cells"] = len(lc)
    if surrogate is not None:
        from mcsim import surrogate as S
        model, sigma = surrogate
        t0 = time.perf_counter()
        guided = S.guided_search(model, lc, sigma, **opts)
        checks.update(guided_s=time.perf_counter() - t0, guided_kernel_ms=lib.last_kernel_ms(),
                      identical_to_plain=all((g.alim, g.lo, g.hi, g.open) == (q.alim, q.lo, q.hi, q.open)
                                             for g, q in zip(guided, plain)))
    res["line_checks"] = checks
    res["pipeline_s"] = wall + checks.get("guided_s", t_plain)
    if cpu:
        import threading
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        k = min(cpu_threads, n_prog)
        lim = [progs[i][1] if not cpu_prefix else min(progs[i][1], cpu_prefix) for i in range(k)]
        outs_c = [None] * k
        frs_c = [None] * k
        H_c = [None] * k

        def work(i):
            sub, n, frames = progs[i]
            act, ide = oracle.OracleStock(width, width, 20, cell), oracle.OracleStock(width, width, 20, cell)
            ide.set_wall_band(2.0)
            eng = oracle.OracleEngine()
            eng.configure(ecells[i])
            eng.set_wear(kte, kre)
            nf = sum(1 for q in range(lim[i]) if sub[q].frame & MC_FRAME_LAST)
            r = oracle.sim_run(act, ide, eng, sub, lim[i], tools[i]["R"], tools[i]["fluteLen"], 0.0, 1.0, ttype[i],
                               dynBudget=DYN_BUDGET, nframes=nf)
            outs_c[i], frs_c[i] = r if nf else (r, [])
            H_c[i] = act.H()
        th = [threading.Thread(target=work, args=(i,)) for i in range(k)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        cw = time.perf_counter() - t0
        same = all(all(a.dx == b.dx and a.dy == b.dy and a.nsteps == b.nsteps and a.arc == b.arc and
                       a.phiSt == b.phiSt and a.phiEx == b.phiEx and a.ap == b.ap and a.errVal == b.errVal and
                       a.maxF == b.maxF and a.cut == b.cut and abs(a.vol - b.vol) <= 1e-12 * max(1.0, abs(b.vol))
                       for a, b in zip(outs[i][:lim[i]], outs_c[i]))
                   and len(outs_c[i]) == lim[i]
                   and all(bytes(a) == bytes(b) for a, b in zip(frs[i], frs_c[i])) and len(frs_c[i]) <= len(frs[i])
                   and (lim[i] < progs[i][1] or len(frs[i]) == len(frs_c[i]))
                   and (lim[i] < progs[i][1] or np.array_equal(P[i]["actual"].H(), H_c[i])) for i in range(k))
        csub = sum(lim)
        res["cpu_baseline"] = dict(value=csub / cw, unit="sub-steps/s", cores=k, kind="port",
                                   sample="%d of the %d programs on %d host threads (oracle/simloop.c), %s%.1f s" %
                                   (k, n_prog, k, "the first %d sub-steps of each, " % cpu_prefix if cpu_prefix else "",
                                    cw))
        res["identical_to_oracle"] = bool(same)
    release(P)
    return res


def surrogate_leg(cells, opts, plain, steps, n_train, epochs):
    """Train the bracket surrogate on cells the GPU search resolves (random tools and speeds
    of the lobe-grid distributions, not the grid's own), then time the surrogate-guided
    search of the grid (identical results to the plain search) and report the +-15 %
    bracket form against plain bisection."""
    import math
    from mcsim import surrogate as S
    t0 = time.perf_counter()
    Xtr, ytr, _ = S.make_dataset(S.sample_cells(n_train, seed=2), opts)
    Xva, yva, _ = S.make_dataset(S.sample_cells(4096, seed=3), opts)
    t_data = time.perf_counter() - t0
    print("surrogate data: %d + 4096 cells in %.1f s" % (n_train, t_data), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    model, loss = S.train(Xtr, ytr, epochs=epochs, batch=1024, lr=2e-3, lr_final=4e-5)
    t_train = time.perf_counter() - t0
    train_ms = lib.last_kernel_ms()
    sigma = S.residual_sigma(model, Xva, yva)
    print("surrogate trained: %d epochs in %.2f s, held-out sigma(ln alim) %.3f" % (epochs, t_train, sigma),
          file=sys.stderr, flush=True)
    prior_sigma = sigma
    times, kms = [], []
    res = None
    for s in range(steps + 1):                  # one untimed warm-up
        t0 = time.perf_counter()
        res = S.guided_search(model, cells, prior_sigma, **opts)
        dt = time.perf_counter() - t0
        if s:
            times.append(dt)
            kms.append(lib.last_kernel_ms())
            print("guided step %d: %.3f s (kernel %.3f s)" % (s - 1, dt, kms[-1] / 1e3), file=sys.stderr, flush=True)
    identical = all((res[i].alim, res[i].lo, res[i].hi, res[i].open, res[i].trials) ==
                    (plain[i].alim, plain[i].lo, plain[i].hi, plain[i].open, plain[i].trials)
                    for i in range(len(cells)))
    t0 = time.perf_counter()
    br, fb = S.bracketed_search(model, cells, tol=0.05, width=0.15, **opts)
    t_br = time.perf_counter() - t0
    hit = [i for i in range(len(cells)) if not fb[i]]
    agree = sum(1 for i in hit if br[i].lo - (plain[i].hi - plain[i].lo) - 1e-9 <= plain[i].alim
                <= br[i].hi + (plain[i].hi - plain[i].lo) + 1e-9)
    err = np.array([abs(br[i].alim - plain[i].alim) for i in range(len(cells))])
    pred = model.predict(lib.surrogate_features(cells))
    lnp = np.array([math.log(r.alim) for r in plain])
    return res, (model, prior_sigma), dict(
        guided_cells_per_s=len(cells) / float(np.mean(times)), guided_ms_per_step=1e3 * float(np.mean(times)),
        guided_kernel_ms=float(np.mean(kms)), identical_to_plain=identical, prior_sigma=prior_sigma,
        training_cells=n_train, heldout_sigma_ln_alim=sigma, grid_sigma_ln_alim=float(np.sqrt(np.mean((pred - lnp) ** 2))),
        grid_within_15pct=float(np.mean(np.abs(pred - lnp) < math.log(1.15))),
        data_generation_s=t_data, train_s=t_train, train_kernel_ms=train_ms, epochs=epochs,
        final_train_loss=float(loss[-1]),
        bracket_15pct={"cells_per_s": len(cells) / t_br, "bracket_hit_rate": len(hit) / len(cells),
                       "hits_agreeing_with_plain": agree / max(1, len(hit)),
                       "abs_alim_err_vs_plain_p95_mm": float(np.percentile(err, 95)),
                       "abs_alim_err_vs_plain_max_mm": float(err.max()), "tol_mm": 0.05}), float(np.mean(kms))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tools", type=int, default=64)
    ap.add_argument("--speeds", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-heightmap", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-coupled", action="store_true")
    ap.add_argument("--coupled-programs", type=int, default=16)
    ap.add_argument("--coupled-lines", type=int, default=4)
    ap.add_argument("--zigzag-substeps", type=int, default=25000,
                    help="sub-steps of the zigzag program run through the coupled loop (its forces)")
    ap.add_argument("--train-cells", type=int, default=262144)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--coupled-full", type=int, default=0, metavar="SEGMENTS",
                    help="only the coupled leg at config 4's full size: programs of SEGMENTS moves (1e5), the "
                         "oracle on the first --cpu-prefix sub-steps of each: one JSON object")
    ap.add_argument("--cpu-prefix", type=int, default=20000)
    ap.add_argument("--heightmap-only", action="store_true",
                    help="only the heightmap leg (for its profiles): one JSON object")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (counter calibration)")
    a = ap.parse_args()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if a.heightmap_only:
        print(json.dumps(heightmap(a.steps, not a.no_cpu, a.cpu_seconds)))
        return
    if a.coupled_full:
        _, opts = W.sld_grid(n_speeds=2, n_tools=1)
        print(json.dumps(coupled(a.coupled_programs, a.coupled_lines, 16, not a.no_cpu, opts,
                                 full_segments=a.coupled_full, cpu_prefix=a.cpu_prefix)))
        return
    cells, opts = W.sld_grid(n_speeds=a.speeds, n_tools=a.tools)
    arr = cell_array(cells)
    print("device:", lib.device_info(), file=sys.stderr, flush=True)
    tuning = None
    if not a.no_autotune:
        from mcsim import autotune
        t0 = time.perf_counter()
        best, best_ms, table = autotune.tune_search(arr, opts)
        tuning = dict(best=best, kernel_ms=best_ms, seconds=time.perf_counter() - t0, candidates=table)
        print("autotuned schedule %r (%.0f ms) in %.1f s" % (best, best_ms, tuning["seconds"]), file=sys.stderr, flush=True)
    for w in range(a.warmup):
        lib.bisect_batch(arr, **opts)
        print("warmup %d done" % w, file=sys.stderr, flush=True)
    times, kms = [], []
    res = None
    for s in range(a.steps):
        t0 = time.perf_counter()
        res = lib.bisect_batch(arr, **opts)
        times.append(time.perf_counter() - t0)
        kms.append(lib.last_kernel_ms())
        print("step %d: %.3f s (kernel %.3f s)" % (s, times[-1], kms[-1] / 1e3), file=sys.stderr, flush=True)
    ms = 1e3 * float(np.mean(times))
    kms_mean = float(np.mean(kms))
    n = len(cells)
    plain_value = n / (ms / 1e3)
    value, ms_step, kernel_ms, path = plain_value, ms, kms_mean, "plain"
    surr = model = None
    if not a.no_surrogate:
        gres, model, surr, gk = surrogate_leg(cells, opts, res, a.steps, a.train_cells, a.epochs)
        if surr["identical_to_plain"] and surr["guided_cells_per_s"] > plain_value:
            value, ms_step, kernel_ms, path = surr["guided_cells_per_s"], surr["guided_ms_per_step"], gk, "surrogate-guided"
    # the roofline is that of the plain launch of cell_search (the one the rocprofv3
    # summaries in profiles/ cover); the guided launch is the same kernel with another
    # speculation order, and its time is in config.surrogate
    ref = reference_limits(out_workload := "sld_grid_%dspeeds_x_%dtools" % (a.speeds, a.tools))
    ref_check = None
    if ref is not None:
        ref_check = dict(cells=n, plain_cells_differing=same_limits(res, ref),
                         plain_differences=reference_differences(res, ref),
                         reference="timeDomainLimit of the reference (Node %s, %d processes, %.0f s for the grid: "
                                   "%.1f cells/s) from tests/golden/golden_tdl_grid.json" %
                                   (ref["node"], ref["procs"], ref["node_s"], n / ref["node_s"]),
                         reference_cells_per_s=n / ref["node_s"])
        if surr is not None:
            ref_check["guided_cells_differing"] = same_limits(gres, ref)
    flops = flop_alg(cells, res)
    achieved = flops / (kms_mean * 1e-3) / 1e12
    trials = sum(r.trials for r in res)
    visits = sum(r.visits for r in res)
    out = {
        "metric": "SLD cells/sec at \u22640.05 mm b_lim error (plain vs surrogate); heightmap GB/s",
        "value": value,
        "unit": "SLD cells/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Park-Miller tool/mode grid; surrogate trained on other seeded tools)",
        "config": {"workload": out_workload, "cells": n,
                   "bisection": opts, "value_path": path, "plain_cells_per_s": plain_value,
                   "plain_ms_per_step": ms, "plain_kernel_ms": kms_mean, "trials": trials,
                   "slice_visits": visits, "slice_visits_per_s": visits / (ms_step / 1e3),
                   "surrogate": surr, "schedule": tuning or lib.search_tuning(),
                   "reference_check": ref_check,
                   "identical_to_reference": None if ref_check is None else
                   (ref_check["plain_cells_differing"] == 0 and ref_check.get("guided_cells_differing", 0) == 0)},
        "roofline": {"bound": "mfma", "pipe": "FP64 VALU", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": profiled_traffic("cell_search"),
                     "kernel": "cell_search", "launch": "plain", "kernel_ms": kms_mean, "flop_alg": flops,
                     "roof": "compute bound, on the FP64 vector (VALU) pipe: MI355X's FP64 vector rate "
                             "78.6 TFLOP/s (equal to its FP64 matrix rate); the kernel issues no matrix "
                             "instructions: each trial is a serial recurrence with no matrix shape",
                     "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                       "copy-calibrated)"},
    }
    out["sld_40rpm"] = sld_40rpm(not a.no_cpu)
    out["dropin_substeps"] = dropin_substeps(not a.no_cpu)
    if not a.no_heightmap:
        out["heightmap"] = heightmap(2, not a.no_cpu)
        out["zigzag_forces"] = zigzag_forces(a.zigzag_substeps, not a.no_cpu)
    if not a.no_coupled:
        out["coupled"] = coupled(a.coupled_programs, a.coupled_lines, 16, not a.no_cpu, opts, model)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cells, opts, a.cpu_seconds, res)
        out["config"]["identical_to_oracle_sample"] = out["cpu_baseline"]["cells_differing_from_gpu"] == 0
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
