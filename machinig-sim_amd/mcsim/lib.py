"""ctypes bindings of libmcsim.so (include/mcsim.h).

The library must be built in-tree (``python machinig-sim_amd/build.py``); loading fails
loudly when it is missing — there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from ._abi import (Cell, Trial, BisectOpts, Limit, Stamp, StampOpt, TrainOpts, LobeArgs, Mode,
                   SearchTuning, EngageQuery, EngageResult, EngageSpec, cell_array, MC_OK, MC_SURR_IN,
                   MC_STOCK_B, EngineInfo, Wall, EngineState, EngineStats, ChipEvent, MC_ENGINE_HIST,
                   MC_ENGINE_RING, Substep, SubstepOut, SimProgram, WallOut, MC_MAX_MODES, FrameOut,
                   MC_FRAME_LAST)

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libmcsim.so")

EXPORTS = [
    "mc_init", "mc_last_error", "mc_device_info", "mc_last_kernel_time", "mc_chatter_batch", "mc_bisect_batch",
    "mc_bisect_bracket_batch", "mc_bisect_prior_batch", "mc_engine_trace", "mc_stock_create", "mc_stock_destroy",
    "mc_stock_dims", "mc_stock_set_wall_band", "mc_stock_read_h", "mc_stock_write_h",
    "mc_stock_read_side", "mc_stamp_batch", "mc_stock_sample", "mc_surrogate_features",
    "mc_mlp_create", "mc_mlp_destroy", "mc_mlp_nparams", "mc_mlp_params", "mc_mlp_set_scaling",
    "mc_mlp_get_scaling", "mc_mlp_grads", "mc_mlp_train", "mc_mlp_predict",
    "mc_directional_coeffs", "mc_stability_lobes", "mc_alim_at", "mc_analytic_limits", "mc_composite_lobes",
    "mc_set_search_tuning", "mc_get_search_tuning", "mc_stamp_batch_ex", "mc_stock_read_marks",
    "mc_stock_dirty", "mc_stock_clear_dirty", "mc_stock_block_max", "mc_stock_max_in_disk", "mc_tip_lut",
    "mc_stock_engagement", "mc_engine_trace_ex", "mc_engine_create", "mc_engine_destroy", "mc_engine_configure",
    "mc_engine_set_wear", "mc_engine_set_work", "mc_engine_set_engagement", "mc_engine_set_end_engagement",
    "mc_engine_set_maps", "mc_engine_get_state", "mc_engine_set_state", "mc_engine_run", "mc_engine_run_batch",
    "mc_engine_read_stats", "mc_engine_chip_events", "mc_engine_chatter_index", "mc_engine_read_rings",
    "mc_sim_run", "mc_js_math", "mc_stamp_batch_noalloc", "mc_path_discretize", "mc_sim_schedule",
    "mc_last_search_trials", "mc_stamp_batch_lines",
]

_LIB = None


class McError(RuntimeError):
    pass


def load():
    """Load libmcsim.so and declare its signatures (no GPU call is made)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise McError("libmcsim.so not built: run python machinig-sim_amd/build.py")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    L.mc_init.argtypes = [C.c_int]
    L.mc_last_error.restype = C.c_char_p
    L.mc_device_info.argtypes = [C.c_char_p, C.c_int]
    L.mc_last_kernel_time.argtypes = [C.POINTER(C.c_double)]
    L.mc_chatter_batch.argtypes = [C.POINTER(Cell), P, C.c_int, C.POINTER(Trial)]
    L.mc_bisect_batch.argtypes = [C.POINTER(Cell), C.c_int, C.POINTER(BisectOpts), C.POINTER(Limit)]
    L.mc_bisect_bracket_batch.argtypes = [C.POINTER(Cell), C.c_int, C.POINTER(BisectOpts), P, P,
                                          C.c_double, C.POINTER(Limit), P]
    L.mc_bisect_prior_batch.argtypes = [C.POINTER(Cell), C.c_int, C.POINTER(BisectOpts), P, P, C.POINTER(Limit)]
    L.mc_engine_trace.argtypes = [C.POINTER(Cell), C.c_double, C.c_int64, P, P, P, P,
                                  C.POINTER(C.c_int64)]
    L.mc_stock_create.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(P)]
    L.mc_stock_destroy.argtypes = [P]
    L.mc_stock_dims.argtypes = [P, P, P, P, P]
    L.mc_stock_set_wall_band.argtypes = [P, C.c_double]
    L.mc_stock_read_h.argtypes = [P, P]
    L.mc_stock_write_h.argtypes = [P, P]
    L.mc_stock_read_side.argtypes = [P, P, P, P, P]
    L.mc_stamp_batch.argtypes = [P, C.POINTER(Stamp), C.c_int, P]
    L.mc_stamp_batch_noalloc.argtypes = [P, C.POINTER(Stamp), C.c_int, P]
    L.mc_stamp_batch_lines.argtypes = [P, C.POINTER(Stamp), C.c_int, P, C.c_int]
    L.mc_stock_sample.argtypes = [P, P, C.c_int, P]
    L.mc_stamp_batch_ex.argtypes = [P, C.POINTER(Stamp), P, C.c_int, P, P, P, C.c_int, P]
    L.mc_stock_read_marks.argtypes = [P, P, P]
    L.mc_stock_dirty.argtypes = [P, P]
    L.mc_stock_clear_dirty.argtypes = [P]
    L.mc_stock_block_max.argtypes = [P, P]
    L.mc_stock_max_in_disk.argtypes = [P, P, C.c_int, P]
    L.mc_stock_engagement.argtypes = [P, P, C.c_int, P, P]
    L.mc_engine_trace_ex.argtypes = [C.POINTER(Cell), C.POINTER(EngageSpec), C.c_int64, P, P, P, P,
                                     C.POINTER(C.c_int64)]
    L.mc_tip_lut.argtypes = [C.c_int32, C.c_double, C.c_double, P, C.c_int32, C.c_double, P, C.c_int32,
                             C.POINTER(C.c_int32)]
    L.mc_surrogate_features.argtypes = [C.POINTER(Cell), C.c_int, P]
    L.mc_mlp_create.argtypes = [C.c_uint64, C.POINTER(P)]
    L.mc_mlp_destroy.argtypes = [P]
    L.mc_mlp_nparams.argtypes = []
    L.mc_mlp_params.argtypes = [P, P, C.c_int]
    L.mc_mlp_set_scaling.argtypes = [P, P, P, C.c_float, C.c_float]
    L.mc_mlp_get_scaling.argtypes = [P, P, P, P, P]
    L.mc_mlp_grads.argtypes = [P, P, P, C.c_int, P, P]
    L.mc_mlp_train.argtypes = [P, P, P, C.c_int, C.POINTER(TrainOpts), P]
    L.mc_mlp_predict.argtypes = [P, P, C.c_int, P]
    L.mc_directional_coeffs.argtypes = [C.c_double, C.c_double, C.c_double, P]
    L.mc_stability_lobes.argtypes = [C.POINTER(LobeArgs), P, P, C.c_int, P, P]
    L.mc_alim_at.argtypes = [P, P, P, C.c_int, C.c_double, P]
    L.mc_analytic_limits.argtypes = [C.POINTER(Cell), C.c_int, P]
    L.mc_composite_lobes.argtypes = [C.POINTER(LobeArgs), C.c_int, C.c_double, C.c_double, P, P, C.c_int, P]
    L.mc_engine_create.argtypes = [C.POINTER(P)]
    L.mc_engine_destroy.argtypes = [P]
    L.mc_engine_configure.argtypes = [P, C.POINTER(Cell), C.c_double, C.POINTER(EngineInfo)]
    L.mc_engine_set_wear.argtypes = [P, C.c_double, C.c_double]
    L.mc_engine_set_work.argtypes = [P, P]
    L.mc_engine_set_engagement.argtypes = [P] + [C.c_double] * 5 + [P, P, P]
    L.mc_engine_set_end_engagement.argtypes = [P, C.c_double, C.c_double, P]
    L.mc_engine_set_maps.argtypes = [P, P, P]
    L.mc_engine_get_state.argtypes = [P, C.POINTER(EngineState)]
    L.mc_engine_set_state.argtypes = [P, C.POINTER(EngineState)]
    L.mc_engine_run.argtypes = [P, C.c_int64]
    L.mc_engine_run_batch.argtypes = [P, P, C.c_int]
    L.mc_engine_read_stats.argtypes = [P, C.POINTER(EngineStats), C.c_int]
    L.mc_engine_chip_events.argtypes = [P, P, C.c_int64, P]
    L.mc_engine_chatter_index.argtypes = [P, P]
    L.mc_engine_read_rings.argtypes = [P, P, P, P, P]
    L.mc_sim_run.argtypes = [C.POINTER(SimProgram), C.c_int]
    L.mc_js_math.argtypes = [C.c_int32, P, P, P, C.c_int64]
    L.mc_last_search_trials.argtypes = [P, P]
    L.mc_path_discretize.argtypes = [P, C.c_int, C.c_double, P, C.c_int, P]
    L.mc_sim_schedule.argtypes = [P, C.c_int, P, P, C.c_int, P, P, C.c_int, P]
    L.mc_set_search_tuning.argtypes = [C.POINTER(SearchTuning)]
    L.mc_get_search_tuning.argtypes = [C.POINTER(SearchTuning)]
    _LIB = L
    return L


def _check(st):
    if st != MC_OK:
        raise McError("libmcsim error %d: %s" % (st, load().mc_last_error().decode()))


def device_info():
    buf = C.create_string_buffer(256)
    _check(load().mc_device_info(buf, 256))
    return buf.value.decode()


def last_kernel_ms():
    ms = C.c_double()
    _check(load().mc_last_kernel_time(C.byref(ms)))
    return ms.value


def last_search_trials():
    """(executed, serial) trials of the last search call: every trial the speculative kernel
    ran, and those on the serial bisection path"""
    ex, se = C.c_int64(), C.c_int64()
    _check(load().mc_last_search_trials(C.byref(ex), C.byref(se)))
    return ex.value, se.value


def _cells(cells):
    if isinstance(cells, C.Array):
        return cells, len(cells)
    arr = cell_array(cells)
    return arr, len(cells)


def chatter_batch(cells, ap):
    """chatterAt for (cells[i], ap[i]); returns the Trial array"""
    arr, n = _cells(cells)
    ap = np.ascontiguousarray(ap, dtype=np.float64)
    assert len(ap) == n
    out = (Trial * n)()
    _check(load().mc_chatter_batch(arr, ap.ctypes.data, n, out))
    return out


def bisect_batch(cells, aMin=0.05, aMax=10.0, iters=10):
    """timeDomainLimit for every cell; returns the Limit array"""
    arr, n = _cells(cells)
    o = BisectOpts(aMin, aMax, iters, 0)
    out = (Limit * n)()
    _check(load().mc_bisect_batch(arr, n, C.byref(o), out))
    return out


def bracket_batch(cells, lo, hi, tol, aMin=0.05, aMax=10.0, iters=10):
    arr, n = _cells(cells)
    o = BisectOpts(aMin, aMax, iters, 0)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    out = (Limit * n)()
    fb = np.zeros(n, dtype=np.int32)
    _check(load().mc_bisect_bracket_batch(arr, n, C.byref(o), lo.ctypes.data, hi.ctypes.data,
                                          float(tol), out, fb.ctypes.data))
    return out, fb


def prior_batch(cells, mu, sig, aMin=0.05, aMax=10.0, iters=10):
    """timeDomainLimit guided by a prior ln(alim) ~ N(mu, sig^2); same results as bisect_batch"""
    arr, n = _cells(cells)
    o = BisectOpts(aMin, aMax, iters, 0)
    mu = np.ascontiguousarray(np.broadcast_to(mu, (n,)), dtype=np.float64)
    sig = np.ascontiguousarray(np.broadcast_to(sig, (n,)), dtype=np.float64)
    out = (Limit * n)()
    _check(load().mc_bisect_prior_batch(arr, n, C.byref(o), mu.ctypes.data, sig.ctypes.data, out))
    return out


def engine_trace(cell, ap, nsteps):
    a = [np.zeros(max(1, nsteps)) for _ in range(4)]
    n = C.c_int64(0)
    _check(load().mc_engine_trace(C.byref(cell), float(ap), int(nsteps), *[x.ctypes.data for x in a],
                                  C.byref(n)))
    return [x[:n.value] for x in a]


def engine_trace_ex(cell, spec, nsteps):
    """per-step Fx, Fy, x, y of the engine as the simulator drives it (mc_engine_trace_ex)"""
    a = [np.zeros(max(1, nsteps)) for _ in range(4)]
    n = C.c_int64(0)
    _check(load().mc_engine_trace_ex(C.byref(cell), C.byref(spec), int(nsteps), *[x.ctypes.data for x in a],
                                     C.byref(n)))
    return [x[:n.value] for x in a]


class Engine:
    """Device-resident CuttingEngine (new CuttingEngine(), src/cutting.js:22-35): the state
    survives between calls, as the simulator keeps its engine across sub-steps."""

    def __init__(self):
        self.L = load()
        self.p = C.c_void_p()
        _check(self.L.mc_engine_create(C.byref(self.p)))

    def close(self):
        if self.p:
            self.L.mc_engine_destroy(self.p)
            self.p = None

    __del__ = close

    def configure(self, cell, fnMax=0.0):
        info = EngineInfo()
        _check(self.L.mc_engine_configure(self.p, C.byref(cell), float(fnMax), C.byref(info)))
        return info

    def set_wear(self, Kte, Kre):
        _check(self.L.mc_engine_set_wear(self.p, float(Kte), float(Kre)))

    def set_work(self, wall):
        _check(self.L.mc_engine_set_work(self.p, C.byref(wall) if wall is not None else None))

    def set_engagement(self, ux, uy, fz, tipZ, ap, sliceR=None, sliceKap=None):
        nz = C.c_int32()
        r = np.ascontiguousarray(sliceR, dtype=np.float64) if sliceR is not None else None
        k = np.ascontiguousarray(sliceKap, dtype=np.float64) if sliceKap is not None else None
        _check(self.L.mc_engine_set_engagement(self.p, ux, uy, fz, tipZ, ap,
                                               r.ctypes.data if r is not None else None,
                                               k.ctypes.data if k is not None else None, C.byref(nz)))
        return nz.value

    def set_end_engagement(self, fza, rOut, rings=None):
        g = np.ascontiguousarray(rings, dtype=np.float64) if rings is not None else None
        _check(self.L.mc_engine_set_end_engagement(self.p, float(fza), float(rOut),
                                                   g.ctypes.data if g is not None else None))

    def set_maps(self, Hb=None, Eb=None):
        hb = np.ascontiguousarray(Hb, dtype=np.float32) if Hb is not None else None
        eb = np.ascontiguousarray(Eb, dtype=np.float32) if Eb is not None else None
        _check(self.L.mc_engine_set_maps(self.p, hb.ctypes.data if hb is not None else None,
                                         eb.ctypes.data if eb is not None else None))

    def state(self):
        s = EngineState()
        _check(self.L.mc_engine_get_state(self.p, C.byref(s)))
        return s

    def set_state(self, **kw):
        s = self.state()
        for k, v in kw.items():
            setattr(s, k, v)
        _check(self.L.mc_engine_set_state(self.p, C.byref(s)))

    def run(self, n):
        _check(self.L.mc_engine_run(self.p, int(n)))

    def stats(self, take=False):
        s = EngineStats()
        _check(self.L.mc_engine_read_stats(self.p, C.byref(s), 1 if take else 0))
        return s

    def chip_events(self):
        n = C.c_int64()
        _check(self.L.mc_engine_chip_events(self.p, None, 0, C.byref(n)))
        out = (ChipEvent * max(1, n.value))()
        _check(self.L.mc_engine_chip_events(self.p, out, n.value, C.byref(n)))
        return np.array([[e.vol, e.phi, e.z, e.ux, e.uy, e.vx, e.vy] for e in out[:n.value]]).reshape(n.value, 7)

    def chatter_index(self):
        out = np.zeros(3)
        _check(self.L.mc_engine_chatter_index(self.p, out.ctypes.data))
        return out

    def rings(self):
        h = np.zeros(5 * MC_ENGINE_HIST, np.float32)
        r = [np.zeros(2 * MC_ENGINE_RING, np.float32) for _ in range(3)]
        _check(self.L.mc_engine_read_rings(self.p, h.ctypes.data, *[x.ctypes.data for x in r]))
        return h, r[0], r[1], r[2]


def engine_run_batch(engines, nsteps):
    """run engine i for nsteps[i] steps, all in one launch"""
    n = len(engines)
    ptrs = (C.c_void_p * max(1, n))(*[e.p.value for e in engines])
    steps = np.ascontiguousarray(nsteps, dtype=np.int64)
    _check(load().mc_engine_run_batch(ptrs, steps.ctypes.data, n))


JS_MATH = dict(sin=0, cos=1, tan=2, atan=3, atan2=4, acos=5, pow=6, hypot=7)


def js_math(fn, x, y=None):
    """V8's Math.<fn> on the host (mc_js_math): elementwise over x (and y); no GPU call"""
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    yy = np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64) if y is not None else None
    out = np.empty_like(x)
    _check(load().mc_js_math(JS_MATH[fn], x.ctypes.data, yy.ctypes.data if yy is not None else None,
                             out.ctypes.data, x.size))
    return out


def sim_run(programs, frames=False, walls=False):
    """the simulator's sub-step loop on the device for several programs in one launch.
    programs: dicts with actual, ideal (Stock), engine (Engine), sub (Substep array) and
    optionally nsub (its length), R, fluteLen, rw, share, type (MC_TOOL_*), for a bull-nose
    tool lut (its tipLut) and cornerR, dynBudget (> 0: the frame budget's share per
    sub-step instead of the fixed share), fz (this.fz at the start), edge (dict(r, dz): the
    imported edge of an MC_TOOL_PROFILE tool, whose lut is its mc_tip_lut table) and wall (dict(E, rho,
    kTool): the thin-wall model, workpieceFlex; the engine configured with fnMax
    MC_WALL_FN_CAP) and budget_left (a one-element list: the frame budget left when the call
    starts inside a frame, replaced by the budget left after its last sub-step); returns a
    SubstepOut list per program, with frames=True also a FrameOut
    list per program (one per sub-step flagged MC_FRAME_LAST), and with walls=True also a
    WallOut list per program (None for a rigid workpiece), in that order"""
    n = len(programs)
    arr = (SimProgram * max(1, n))()
    outs, keep, frs, wls, bls = [], [], [], [], []
    for i, g in enumerate(programs):
        sub = g["sub"]
        ns = int(g.get("nsub", len(sub)))
        if not 0 <= ns <= len(sub):
            raise ValueError("program %d: nsub %d outside [0, %d] (its sub-step array)" % (i, ns, len(sub)))
        out = (SubstepOut * max(1, ns))()
        outs.append((out, ns))
        lut = g.get("lut")
        lut_arr = np.ascontiguousarray(lut, dtype=np.float32) if lut is not None else None
        keep.append(lut_arr)
        nf = sum(1 for k in range(ns) if sub[k].frame & MC_FRAME_LAST) if frames else 0
        fr = (FrameOut * max(1, nf))()
        frs.append((fr, nf))
        w = g.get("wall")
        wo = (WallOut * max(1, ns))() if w else None
        wls.append((wo, ns))
        e = g.get("edge")
        er = np.ascontiguousarray(e["r"], dtype=np.float64) if e is not None else None
        keep.append(er)
        bl = g.get("budget_left")
        bcell = C.c_double(float(bl[0])) if bl is not None else None
        bls.append((bl, bcell))
        arr[i] = SimProgram(g["actual"].p.value, g["ideal"].p.value, g["engine"].p.value,
                            C.cast(sub, C.POINTER(Substep)), C.cast(out, C.POINTER(SubstepOut)),
                            float(g["R"]), float(g["fluteLen"]), float(g.get("rw", 0.0)), float(g.get("share", 1.0)),
                            ns, int(g["type"]),
                            lut_arr.ctypes.data_as(C.POINTER(C.c_float)) if lut_arr is not None else None,
                            len(lut_arr) if lut_arr is not None else 0, nf, float(g.get("cornerR", 0.0)),
                            float(g.get("dynBudget", 0.0)), float(g.get("fz", 0.0)),
                            C.cast(fr, C.c_void_p) if nf else None, 1 if w else 0, 0,
                            float(w["E"]) if w else 0.0, float(w["rho"]) if w else 0.0,
                            float(w["kTool"]) if w else 0.0, C.cast(wo, C.POINTER(WallOut)) if w else None,
                            er.ctypes.data_as(C.POINTER(C.c_double)) if er is not None else None,
                            float(e["dz"]) if e is not None else 0.0, len(er) if er is not None else 0, 0,
                            C.pointer(bcell) if bcell is not None else None)
    _check(load().mc_sim_run(arr, n))
    for bl, bcell in bls:
        if bl is not None:
            bl[0] = bcell.value
    res = [[o[:ns] for o, ns in outs]]
    if frames:
        res.append([f[:nf] for f, nf in frs])
    if walls:
        res.append([wo[:ns] if wo is not None else None for wo, ns in wls])
    return res[0] if len(res) == 1 else tuple(res)


class Stock:
    """Device-resident heightfield (new Stock(w, l, h, cell), src/stock.js:9)."""

    def __init__(self, w, l, h, cell=0.3):
        self.L = load()
        self.p = C.c_void_p()
        _check(self.L.mc_stock_create(float(w), float(l), float(h), float(cell), C.byref(self.p)))
        nx, ny = C.c_int32(), C.c_int32()
        hx, hy = C.c_double(), C.c_double()
        _check(self.L.mc_stock_dims(self.p, C.byref(nx), C.byref(ny), C.byref(hx), C.byref(hy)))
        self.nx, self.ny, self.hx, self.hy = nx.value, ny.value, hx.value, hy.value

    def close(self):
        if self.p:
            self.L.mc_stock_destroy(self.p)
            self.p = None

    __del__ = close

    def set_wall_band(self, b):
        _check(self.L.mc_stock_set_wall_band(self.p, float(b)))

    def H(self):
        out = np.empty(self.nx * self.ny, dtype=np.float32)
        _check(self.L.mc_stock_read_h(self.p, out.ctypes.data))
        return out

    def write_H(self, H):
        H = np.ascontiguousarray(H, dtype=np.float32)
        assert H.size == self.nx * self.ny
        _check(self.L.mc_stock_write_h(self.p, H.ctypes.data))

    def side(self):
        n = self.nx * self.ny
        e, w = np.empty(n, np.float32), np.empty(n, np.float32)
        lc, lw = np.empty(n, np.int32), np.empty(n, np.int32)
        _check(self.L.mc_stock_read_side(self.p, e.ctypes.data, w.ctypes.data, lc.ctypes.data,
                                         lw.ctypes.data))
        return e, w, lc, lw

    def stamps(self, stamps, n=None, want_vol=True):
        if not isinstance(stamps, C.Array):
            arr = (Stamp * len(stamps))()
            for i, s in enumerate(stamps):
                arr[i] = s
            stamps = arr
        n = len(stamps) if n is None else n
        vol = np.zeros(n) if want_vol else None
        _check(self.L.mc_stamp_batch(self.p, stamps, n, vol.ctypes.data if want_vol else None))
        return vol

    def stamps_lines(self, stamps, nlines, n=None):
        """the stamps with their removed volume summed per program line (mc_stamp_batch_lines):
        returns the nlines per-line volumes"""
        n = len(stamps) if n is None else n
        vol = np.zeros(int(nlines))
        _check(self.L.mc_stamp_batch_lines(self.p, stamps, n, vol.ctypes.data, int(nlines)))
        return vol

    def sample(self, xy):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        out = np.empty(len(xy))
        _check(self.L.mc_stock_sample(self.p, xy.ctypes.data, len(xy), out.ctypes.data))
        return out

    def stamps_ex(self, stamps, opts, n=None, luts=None, lut_off=None, lut_len=None, want_vol=True):
        """Stamps with feed marks and tip profiles: opts is a StampOpt array (or None);
        luts the concatenated f32 profiles with int32 offsets / lengths per profile."""
        n = len(stamps) if n is None else n
        vol = np.zeros(n) if want_vol else None
        keep = [np.ascontiguousarray(a, dt) if a is not None else None
                for a, dt in ((luts, np.float32), (lut_off, np.int32), (lut_len, np.int32))]
        nl = 0 if keep[2] is None else len(keep[2])
        ptr = [a.ctypes.data if a is not None and a.size else None for a in keep]
        _check(self.L.mc_stamp_batch_ex(self.p, stamps, opts, n, ptr[0], ptr[1], ptr[2], nl,
                                        vol.ctypes.data if want_vol else None))
        return vol

    def marks(self):
        n = self.nx * self.ny
        ms, mf = np.empty(n, np.float32), np.empty(n, np.float32)
        _check(self.L.mc_stock_read_marks(self.p, ms.ctypes.data, mf.ctypes.data))
        return ms, mf

    def dirty(self):
        r = (C.c_int32 * 4)()
        _check(self.L.mc_stock_dirty(self.p, r))
        return list(r)

    def clear_dirty(self):
        _check(self.L.mc_stock_clear_dirty(self.p))

    def block_maxima(self):
        nb = ((self.nx + MC_STOCK_B - 1) // MC_STOCK_B) * ((self.ny + MC_STOCK_B - 1) // MC_STOCK_B)
        out = np.empty(nb, np.float32)
        _check(self.L.mc_stock_block_max(self.p, out.ctypes.data))
        return out

    def max_in_disk(self, q):
        """q: (n, 4) rows {cx, cy, r, above} -> (n, 3) rows {h, x, y}"""
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 4)
        out = np.empty((len(q), 3))
        _check(self.L.mc_stock_max_in_disk(self.p, q.ctypes.data, len(q), out.ctypes.data))
        return out

    def engagement(self, queries, want_hb=True):
        """Simulator.fillEngagement + immersion for a list of dicts (cx, cy, ux, uy, R, tipZ,
        fluteLen=inf): (results EngageResult array, Hb f32 (n, 180) or None)"""
        n = len(queries)
        qa = (EngageQuery * max(1, n))()
        for i, q in enumerate(queries):
            qa[i] = EngageQuery(q["cx"], q["cy"], q["ux"], q["uy"], q["R"], q["tipZ"],
                                q.get("fluteLen", float("inf")))
        res = (EngageResult * max(1, n))()
        Hb = np.empty((n, 180), np.float32) if want_hb else None
        _check(self.L.mc_stock_engagement(self.p, qa, n, Hb.ctypes.data if want_hb else None, res))
        return res, Hb


def tip_lut(shape, D, corner_r=0.0, edge_r=None, edge_dz=0.0):
    """tipLut (src/assembly.js:38-54): f32 table of heights above the tip"""
    L = load()
    er = np.ascontiguousarray(edge_r if edge_r is not None else [], np.float64)
    n = C.c_int32()
    args = (int(shape), float(D), float(corner_r), er.ctypes.data if er.size else None, len(er),
            float(edge_dz))
    _check(L.mc_tip_lut(*args, None, 0, C.byref(n)))
    out = np.zeros(n.value, np.float32)
    _check(L.mc_tip_lut(*args, out.ctypes.data, n.value, C.byref(n)))
    return out


def surrogate_features(cells):
    """the 24 raw surrogate inputs of each cell (n x 24 float32)"""
    arr, n = _cells(cells)
    out = np.zeros((n, MC_SURR_IN), dtype=np.float32)
    _check(load().mc_surrogate_features(arr, n, out.ctypes.data))
    return out


class Mlp:
    """The bracket surrogate (24 -> 256 -> 256 -> 256 -> 1), trained on the device."""

    def __init__(self, seed=1):
        self.L = load()
        self.p = C.c_void_p()
        _check(self.L.mc_mlp_create(int(seed), C.byref(self.p)))
        self.nparams = self.L.mc_mlp_nparams()

    def close(self):
        if self.p:
            self.L.mc_mlp_destroy(self.p)
            self.p = None

    __del__ = close

    def params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        _check(self.L.mc_mlp_params(self.p, out.ctypes.data, 0))
        return out

    def set_params(self, v):
        v = np.ascontiguousarray(v, dtype=np.float32)
        assert v.size == self.nparams
        _check(self.L.mc_mlp_params(self.p, v.ctypes.data, 1))

    def set_scaling(self, xm, xs, ym, ys):
        xm = np.ascontiguousarray(xm, dtype=np.float32)
        xs = np.ascontiguousarray(xs, dtype=np.float32)
        _check(self.L.mc_mlp_set_scaling(self.p, xm.ctypes.data, xs.ctypes.data, float(ym), float(ys)))

    def scaling(self):
        xm, xs = np.empty(MC_SURR_IN, np.float32), np.empty(MC_SURR_IN, np.float32)
        ym, ys = C.c_float(), C.c_float()
        _check(self.L.mc_mlp_get_scaling(self.p, xm.ctypes.data, xs.ctypes.data, C.byref(ym), C.byref(ys)))
        return xm, xs, ym.value, ys.value

    def grads(self, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_float()
        _check(self.L.mc_mlp_grads(self.p, X.ctypes.data, t.ctypes.data, len(t), g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def train(self, X, t, epochs=100, batch=4096, lr=2e-3, lr_final=5e-5, beta1=0.9, beta2=0.999,
              eps=1e-8, fit_scaling=True, seed=1):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        o = TrainOpts(int(epochs), int(batch), lr, lr_final, beta1, beta2, eps, int(fit_scaling), int(seed))
        loss = np.zeros(epochs, dtype=np.float32)
        _check(self.L.mc_mlp_train(self.p, X.ctypes.data, t.ctypes.data, len(t), C.byref(o), loss.ctypes.data))
        return loss

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.empty(len(X), dtype=np.float32)
        _check(self.L.mc_mlp_predict(self.p, X.ctypes.data, len(X), out.ctypes.data))
        return out

    MAGIC = b"MCSURR01"

    def save(self, path, sigma=0.0):
        """the model file the JS Surrogate reads: MAGIC, int32 nparams, int32 inputs, float32
        y_mean, y_std, sigma (the held-out spread of ln alim), 4 zero bytes, then float32
        x_mean[inputs], x_std[inputs], params[nparams], little-endian"""
        import struct
        xm, xs, ym, ys = self.scaling()
        with open(path, "wb") as f:
            f.write(self.MAGIC + struct.pack("<iifffi", self.nparams, MC_SURR_IN, ym, ys, float(sigma), 0))
            for a in (xm, xs, self.params()):
                f.write(np.ascontiguousarray(a, dtype="<f4").tobytes())

    @classmethod
    def load(cls, path):
        """(model, sigma) from a file save() wrote"""
        import struct
        raw = open(path, "rb").read()
        if raw[:8] != cls.MAGIC:
            raise ValueError("not a surrogate model file")
        npar, nin, ym, ys, sigma, _ = struct.unpack("<iifffi", raw[8:32])
        m = cls()
        if npar != m.nparams or nin != MC_SURR_IN:
            raise ValueError("model file for another network shape")
        v = np.frombuffer(raw, dtype="<f4", offset=32)
        m.set_scaling(v[:nin], v[nin:2 * nin], ym, ys)
        m.set_params(v[2 * nin:2 * nin + npar])
        return m, sigma


def directional_coeffs(Kr, phiSt, phiEx):
    out = np.zeros(4)
    _check(load().mc_directional_coeffs(float(Kr), float(phiSt), float(phiEx), out.ctypes.data))
    return out


def stability_lobes(N, Ktc, Krc, phiSt, phiEx, lobes, rpmMin, rpmMax, modes_x, modes_y=None, cap=400000):
    """stabilityLobes (src/sld.js:23): list of (rpm array, alim array) per non-empty lobe"""
    a = lobe_args(N, Ktc, Krc, phiSt, phiEx, lobes, rpmMin, rpmMax, modes_x, modes_y)
    rpm, alim = np.zeros(cap), np.zeros(cap)
    ll = np.zeros(max(1, lobes), dtype=np.int32)
    n = C.c_int(0)
    _check(load().mc_stability_lobes(C.byref(a), rpm.ctypes.data, alim.ctypes.data, cap, ll.ctypes.data, C.byref(n)))
    out, off = [], 0
    for k in ll[ll > 0]:
        out.append((rpm[off:off + k].copy(), alim[off:off + k].copy()))
        off += k
    return out


def alim_at(lobes, S):
    """alimAt (src/sld.js:64) over the lobes stability_lobes returns"""
    if not lobes:
        return float("inf")
    rpm = np.concatenate([l[0] for l in lobes])
    alim = np.concatenate([l[1] for l in lobes])
    ll = np.array([len(l[0]) for l in lobes], dtype=np.int32)
    out = C.c_double()
    _check(load().mc_alim_at(rpm.ctypes.data, alim.ctypes.data, ll.ctypes.data, len(ll), float(S), C.byref(out)))
    return out.value


def lobe_args(N, Ktc, Krc, phiSt, phiEx, lobes, rpmMin, rpmMax, modes_x, modes_y=None):
    """one mc_lobe_args record (stabilityLobes' parameters, src/sld.js:23)"""
    modes_y = modes_y or modes_x
    if not (1 <= len(modes_x) <= MC_MAX_MODES and 1 <= len(modes_y) <= MC_MAX_MODES):
        raise ValueError("1..%d modes per axis" % MC_MAX_MODES)
    a = LobeArgs()
    a.N, a.lobes, a.Ktc, a.Krc, a.phiSt, a.phiEx = int(N), int(lobes), Ktc, Krc, phiSt, phiEx
    a.rpmMin, a.rpmMax, a.nmx, a.nmy = rpmMin, rpmMax, len(modes_x), len(modes_y)
    for i, m in enumerate(modes_x):
        a.mx[i] = Mode(*map(float, m))
    for i, m in enumerate(modes_y):
        a.my[i] = Mode(*map(float, m))
    return a


def composite_lobes(nodes, step=25.0, hi=44000.0):
    """Simulator.lobes' composite curve (src/sim.js:270-281) from its speed nodes (LobeArgs
    records, see lobe_args; each node's window [rpmMin, rpmMax) is read every step below hi):
    (rpm, alim) arrays, node after node; 1e3 where no lobe covers a speed"""
    n = len(nodes)
    arr = (LobeArgs * n)(*nodes)
    cnt = C.c_int(0)
    _check(load().mc_composite_lobes(arr, n, float(step), float(hi), None, None, 0, C.byref(cnt)))
    rpm, alim = np.zeros(cnt.value), np.zeros(cnt.value)
    _check(load().mc_composite_lobes(arr, n, float(step), float(hi), rpm.ctypes.data, alim.ctypes.data, cnt.value,
                                     C.byref(cnt)))
    return rpm, alim


def analytic_limits(cells):
    arr, n = _cells(cells)
    out = np.zeros(n)
    _check(load().mc_analytic_limits(arr, n, out.ctypes.data))
    return out


def search_tuning():
    t = SearchTuning()
    _check(load().mc_get_search_tuning(C.byref(t)))
    return dict(waves_per_cu=t.waves_per_cu, wide_max=t.wide_max, wide_frac=t.wide_frac, use_graph=t.use_graph)


def set_search_tuning(waves_per_cu=8, wide_max=256, wide_frac=0.5, use_graph=1):
    t = SearchTuning(int(waves_per_cu), int(wide_max), float(wide_frac), int(use_graph))
    _check(load().mc_set_search_tuning(C.byref(t)))
