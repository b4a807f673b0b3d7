"""ctypes bindings of liboracle.so — TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference's hot paths (see engine.c, sld.c,
stock.c, fdlibm.c).  Only tests/, the smoke check and bench.py's CPU-baseline leg load
it, as the checker; the product never does.
"""
import ctypes as C
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "machinig-sim_amd"))

from mcsim._abi import (Cell, Trial, BisectOpts, Limit, Mode, Stamp, StampOpt, EngageSpec, WallOut,  # noqa: E402
                        EngineInfo, Wall, EngineState, EngineStats, ChipEvent, MC_ENGINE_HIST, MC_ENGINE_RING,
                        Substep, SubstepOut, FrameOut)

_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        L.or_math_batch.argtypes = [C.c_int, P, P, P, C.c_int64]
        L.or_chatter_at.argtypes = [C.POINTER(Cell), C.c_double, C.POINTER(Trial), P]
        L.or_time_domain_limit.argtypes = [C.POINTER(Cell), C.POINTER(BisectOpts), C.POINTER(Limit)]
        L.or_bracket_limit.argtypes = [C.POINTER(Cell), C.POINTER(BisectOpts), C.c_double, C.c_double,
                                       C.c_double, C.POINTER(Limit), C.POINTER(C.c_int32)]
        L.or_engine_trace.argtypes = [C.POINTER(Cell), C.c_double, C.c_int64, P, P, P, P]
        L.or_engine_trace.restype = C.c_int64
        L.or_engine_trace_ex.argtypes = [C.POINTER(Cell), C.POINTER(EngageSpec), C.c_int64, P, P, P, P]
        L.or_engine_trace_ex.restype = C.c_int64
        L.or_bisect_batch_mt.argtypes = [C.POINTER(Cell), C.c_int, C.POINTER(BisectOpts), C.c_int,
                                         C.POINTER(Limit)]
        L.or_chatter_batch_mt.argtypes = [C.POINTER(Cell), P, C.c_int, C.c_int, C.POINTER(Trial)]
        L.or_engine_create.restype = P
        L.or_engine_destroy.argtypes = [P]
        L.or_engine_configure.argtypes = [P, C.POINTER(Cell), C.c_double, C.POINTER(EngineInfo)]
        L.or_engine_set_wear.argtypes = [P, C.c_double, C.c_double]
        L.or_engine_set_work.argtypes = [P, P]
        L.or_engine_set_engagement.argtypes = [P] + [C.c_double] * 5 + [P, P, P]
        L.or_engine_set_end_engagement.argtypes = [P, C.c_double, C.c_double, P]
        L.or_engine_set_maps.argtypes = [P, P, P]
        L.or_engine_get_state.argtypes = [P, C.POINTER(EngineState)]
        L.or_engine_set_state.argtypes = [P, C.POINTER(EngineState)]
        L.or_engine_run.argtypes = [P, C.c_int64]
        L.or_engine_stats.argtypes = [P, C.POINTER(EngineStats), C.c_int]
        L.or_engine_run_max_f.argtypes = [P]
        L.or_engine_run_max_f.restype = C.c_double
        L.or_engine_chip_events.argtypes = [P, P, C.c_int64]
        L.or_engine_chip_events.restype = C.c_int64
        L.or_engine_chatter_index.argtypes = [P, P]
        L.or_engine_read_rings.argtypes = [P, P, P, P, P]
        L.or_sim_run.argtypes = [P, P, P, C.POINTER(Substep), C.c_int] + [C.c_double] * 4 + \
            [C.c_int, P, C.c_int, C.c_double, C.POINTER(SubstepOut), C.c_double, C.c_double, P, C.c_int,
             C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(WallOut), P, C.c_int, C.c_double, P]
        L.or_wall_at.argtypes = [P] + [C.c_double] * 4 + [P] + [C.c_double] * 6 + [C.POINTER(WallOut)]
        L.or_wall_at.restype = C.c_int
        L.or_stock_zbottom.argtypes = [P]
        L.or_stock_zbottom.restype = C.c_double
        L.or_directional_coeffs.argtypes = [C.c_double, C.c_double, C.c_double, P]
        L.or_stability_lobes.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.c_double, C.c_double, C.POINTER(Mode), C.c_int,
                                         C.POINTER(Mode), C.c_int, P, P, C.c_int, P]
        L.or_alim_at.argtypes = [P, P, P, C.c_int, C.c_double]
        L.or_alim_at.restype = C.c_double
        L.or_composite_lobes.argtypes = [C.c_int] + [P] * 12 + [C.c_int, C.c_double, C.c_double, P, P]
        L.or_composite_lobes.restype = C.c_int
        L.or_stock_new.argtypes = [C.c_double] * 4
        L.or_stock_new.restype = P
        L.or_stock_free.argtypes = [P]
        L.or_stock_dims.argtypes = [P, P, P, P, P]
        L.or_stock_set_wall_band.argtypes = [P, C.c_double]
        for f in ("or_stock_h", "or_stock_err", "or_stock_walld", "or_stock_linecut",
                  "or_stock_linewall"):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = P
        L.or_stamp_batch.argtypes = [P, C.POINTER(Stamp), C.c_int, P]
        L.or_stamp_batch_ex.argtypes = [P, C.POINTER(Stamp), P, C.c_int, P, P, P, P]
        L.or_stock_mark_wall.argtypes = [P] + [C.c_double] * 6
        for f in ("or_stock_marks", "or_stock_markf", "or_stock_wallerr"):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = P
        L.or_stock_dirty.argtypes = [P, P]
        L.or_stock_clear_dirty.argtypes = [P]
        L.or_stock_block_max.argtypes = [P, C.c_int32, C.c_int32]
        L.or_stock_block_max.restype = C.c_float
        L.or_stock_max_in_disk.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double, P]
        L.or_tip_lut.argtypes = [C.c_int, C.c_double, C.c_double, P, C.c_int, C.c_double, P]
        L.or_tip_lut.restype = C.c_int
        L.or_fill_engagement.argtypes = [P] + [C.c_double] * 5 + [P]
        L.or_fill_engagement.restype = C.c_double
        L.or_immersion.argtypes = [P, C.c_double, P, P]
        L.or_immersion.restype = C.c_int
        L.or_stock_sample.argtypes = [P, C.c_double, C.c_double]
        L.or_stock_sample.restype = C.c_double
        L.or_tool_model.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, P, P]
        for f in ("or_sin", "or_cos", "or_tan", "or_atan", "or_js_round"):
            getattr(L, f).argtypes = [C.c_double]
            getattr(L, f).restype = C.c_double
        for f in ("or_pow", "or_hypot"):
            getattr(L, f).argtypes = [C.c_double, C.c_double]
            getattr(L, f).restype = C.c_double
        _LIB = L
    return _LIB


MATH_FN = dict(sin=0, cos=1, tan=2, pow=3, hypot=4, atan=5, atan2=6)


def math_batch(fn, x, y=None):
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = x if y is None else np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(x)
    lib().or_math_batch(MATH_FN[fn], x.ctypes.data, y.ctypes.data, out.ctypes.data, len(x))
    return out


def chatter_at(cell, ap):
    t = Trial()
    st = lib().or_chatter_at(C.byref(cell), float(ap), C.byref(t), None)
    if st:
        raise RuntimeError("oracle chatterAt failed: %d" % st)
    return t


def time_domain_limit(cell, aMin=0.05, aMax=10.0, iters=10):
    o = BisectOpts(aMin, aMax, iters, 0)
    r = Limit()
    st = lib().or_time_domain_limit(C.byref(cell), C.byref(o), C.byref(r))
    if st:
        raise RuntimeError("oracle timeDomainLimit failed: %d" % st)
    return r


def bracket_limit(cell, lo, hi, tol, aMin=0.05, aMax=10.0, iters=10):
    o = BisectOpts(aMin, aMax, iters, 0)
    r = Limit()
    fb = C.c_int32(0)
    st = lib().or_bracket_limit(C.byref(cell), C.byref(o), float(lo), float(hi), float(tol),
                                C.byref(r), C.byref(fb))
    if st:
        raise RuntimeError("oracle bracket search failed: %d" % st)
    return r, fb.value


def bisect_batch(cells_arr, n, aMin=0.05, aMax=10.0, iters=10, threads=1):
    o = BisectOpts(aMin, aMax, iters, 0)
    out = (Limit * n)()
    lib().or_bisect_batch_mt(cells_arr, n, C.byref(o), threads, out)
    return out


def chatter_batch(cells_arr, ap, threads=1):
    import numpy as np
    ap = np.ascontiguousarray(ap, dtype=np.float64)
    out = (Trial * len(ap))()
    lib().or_chatter_batch_mt(cells_arr, ap.ctypes.data, len(ap), threads, out)
    return out


def engine_trace(cell, ap, nsteps):
    import numpy as np
    a = [np.zeros(nsteps) for _ in range(4)]
    n = lib().or_engine_trace(C.byref(cell), float(ap), int(nsteps),
                              *[x.ctypes.data for x in a])
    return [x[:n] for x in a]


def engine_trace_ex(cell, spec, nsteps):
    """the engine of one simulator sub-step (src/sim.js:252-257,403-404) restated"""
    import numpy as np
    a = [np.zeros(max(1, nsteps)) for _ in range(4)]
    n = lib().or_engine_trace_ex(C.byref(cell), C.byref(spec), int(nsteps), *[x.ctypes.data for x in a])
    assert n == nsteps
    return [x[:n] for x in a]


class OracleEngine:
    """the CuttingEngine object restated (oracle/cengine.c), with the same methods as the
    device engine's binding so one replay drives either"""

    def __init__(self):
        self.L = lib()
        self.h = self.L.or_engine_create()
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_engine_destroy(self.h)
            self.h = None

    def configure(self, cell, fnMax=0.0):
        info = EngineInfo()
        st = self.L.or_engine_configure(self.h, C.byref(cell), float(fnMax), C.byref(info))
        if st:
            raise RuntimeError("or_engine_configure: %d" % st)
        return info

    def set_wear(self, Kte, Kre):
        self.L.or_engine_set_wear(self.h, float(Kte), float(Kre))

    def set_work(self, wall):
        self.L.or_engine_set_work(self.h, C.byref(wall) if wall is not None else None)

    def set_engagement(self, ux, uy, fz, tipZ, ap, sliceR=None, sliceKap=None):
        import numpy as np
        nz = C.c_int32()
        r = np.ascontiguousarray(sliceR, dtype=np.float64) if sliceR is not None else None
        k = np.ascontiguousarray(sliceKap, dtype=np.float64) if sliceKap is not None else None
        st = self.L.or_engine_set_engagement(self.h, ux, uy, fz, tipZ, ap, r.ctypes.data if r is not None else None,
                                             k.ctypes.data if k is not None else None, C.byref(nz))
        if st:
            raise RuntimeError("or_engine_set_engagement: %d" % st)
        return nz.value

    def set_end_engagement(self, fza, rOut, rings=None):
        import numpy as np
        g = np.ascontiguousarray(rings, dtype=np.float64) if rings is not None else None
        self.L.or_engine_set_end_engagement(self.h, float(fza), float(rOut), g.ctypes.data if g is not None else None)

    def set_maps(self, Hb=None, Eb=None):
        import numpy as np
        hb = np.ascontiguousarray(Hb, dtype=np.float32) if Hb is not None else None
        eb = np.ascontiguousarray(Eb, dtype=np.float32) if Eb is not None else None
        self.L.or_engine_set_maps(self.h, hb.ctypes.data if hb is not None else None,
                                  eb.ctypes.data if eb is not None else None)

    def state(self):
        s = EngineState()
        self.L.or_engine_get_state(self.h, C.byref(s))
        return s

    def set_state(self, **kw):
        s = self.state()
        for k, v in kw.items():
            setattr(s, k, v)
        self.L.or_engine_set_state(self.h, C.byref(s))

    def run(self, n):
        st = self.L.or_engine_run(self.h, int(n))
        if st:
            raise RuntimeError("or_engine_run: %d" % st)

    def stats(self, take=False):
        s = EngineStats()
        self.L.or_engine_stats(self.h, C.byref(s), 1 if take else 0)
        return s

    def run_max_f(self):
        """largest |(Fx, Fy)| over the last run (the sub-step loop's maxF)"""
        return self.L.or_engine_run_max_f(self.h)

    def chip_events(self):
        import numpy as np
        n = self.L.or_engine_chip_events(self.h, None, 0)
        out = (ChipEvent * max(1, n))()
        self.L.or_engine_chip_events(self.h, out, n)
        return np.array([[e.vol, e.phi, e.z, e.ux, e.uy, e.vx, e.vy] for e in out[:n]]).reshape(n, 7)

    def chatter_index(self):
        import numpy as np
        out = np.zeros(3)
        self.L.or_engine_chatter_index(self.h, out.ctypes.data)
        return out

    def rings(self):
        import numpy as np
        h = np.zeros(5 * MC_ENGINE_HIST, np.float32)
        r = [np.zeros(2 * MC_ENGINE_RING, np.float32) for _ in range(3)]
        self.L.or_engine_read_rings(self.h, h.ctypes.data, *[x.ctypes.data for x in r])
        return h, r[0], r[1], r[2]


def sim_run(actual, ideal, engine, sub, n, R, fluteLen, rw, share, typ, lut=None, cornerR=0.0, dynBudget=0.0,
            fz0=0.0, nframes=0, wall=None, edge=None, budget_left=None):
    """the simulator's sub-step loop restated (oracle/simloop.c); returns SubstepOut records,
    and with nframes > 0 also the FrameOut records of the frames that ended.  wall: None for a
    rigid workpiece, or dict(E, rho, kTool) for the thin-wall model (workpieceFlex); then the
    WallOut records of every sub-step come last in the returned tuple.  edge: dict(r, dz) of
    an MC_TOOL_PROFILE tool (its imported edge).  budget_left: None, or a one-element list
    holding the frame budget left when the call starts (a call that begins inside a frame),
    replaced by the budget left after its last sub-step"""
    import numpy as np
    out = (SubstepOut * max(1, n))()
    fr = (FrameOut * max(1, nframes))()
    wo = (WallOut * max(1, n))()
    la = np.ascontiguousarray(lut, dtype=np.float32) if lut is not None else None
    w = wall or {}
    er = np.ascontiguousarray(edge["r"], dtype=np.float64) if edge is not None else None
    bl = C.c_double(float(budget_left[0]) if budget_left is not None else 0.0)
    st = lib().or_sim_run(actual.p, ideal.p, engine.h, C.cast(sub, C.POINTER(Substep)), int(n), float(R),
                          float(fluteLen), float(rw), float(share), int(typ),
                          la.ctypes.data if la is not None else None, len(la) if la is not None else 0,
                          float(cornerR), out, float(dynBudget), float(fz0), fr if nframes else None, int(nframes),
                          1 if wall else 0, float(w.get("E", 0.0)), float(w.get("rho", 0.0)),
                          float(w.get("kTool", 0.0)), wo if wall else None,
                          er.ctypes.data if er is not None else None, len(er) if er is not None else 0,
                          float(edge["dz"]) if edge is not None else 0.0,
                          C.byref(bl) if budget_left is not None else None)
    if st:
        raise RuntimeError("or_sim_run: %d" % st)
    if budget_left is not None:
        budget_left[0] = bl.value
    res = [out[:n]]
    if nframes:
        ended = min(nframes, sum(1 for k in range(n) if sub[k].frame & 2))
        res.append(fr[:ended])
    if wall:
        res.append(wo[:n])
    return res[0] if len(res) == 1 else tuple(res)


def wall_at(stock, px, py, pz, R, Hb, ux, uy, zLoad, E, rho, kTool):
    """wallAt restated (oracle/wall.c): a WallOut with found 0 (null) or the wall's fields"""
    import numpy as np
    hb = np.ascontiguousarray(Hb, dtype=np.float32)
    w = WallOut()
    w.found = lib().or_wall_at(stock.p, float(px), float(py), float(pz), float(R), hb.ctypes.data, float(ux),
                               float(uy), float(zLoad), float(E), float(rho), float(kTool), C.byref(w))
    return w


def composite_lobes(nodes, step, hi):
    """Simulator.lobes' composite curve (src/sim.js:270-281) over speed nodes, each a dict of
    N, Ktc, Krc, phiSt, phiEx, lobes, rpmMin, rpmMax, mx, my (lists of (k, fn, zeta)):
    (rpm, alim) arrays"""
    import numpy as np
    n = len(nodes)
    I = lambda k: np.array([d[k] for d in nodes], dtype=np.int32)
    F = lambda k: np.array([d[k] for d in nodes], dtype=np.float64)
    mx = np.zeros((n, 4, 3))
    my = np.zeros((n, 4, 3))
    for i, d in enumerate(nodes):
        mx[i, :len(d["mx"])] = d["mx"]
        my[i, :len(d["my"])] = d["my"]
    nmx = np.array([len(d["mx"]) for d in nodes], dtype=np.int32)
    nmy = np.array([len(d["my"]) for d in nodes], dtype=np.int32)
    args = [I("N"), F("Ktc"), F("Krc"), F("phiSt"), F("phiEx"), I("lobes"), F("rpmMin"), F("rpmMax"),
            mx, nmx, my, nmy]
    ptrs = [a.ctypes.data for a in args]
    tot = lib().or_composite_lobes(n, *ptrs, 4, float(step), float(hi), None, None)
    rpm, alim = np.zeros(tot), np.zeros(tot)
    lib().or_composite_lobes(n, *ptrs, 4, float(step), float(hi), rpm.ctypes.data, alim.ctypes.data)
    return rpm, alim


class OracleStock:
    def __init__(self, w, l, h, cell):
        self.L = lib()
        self.p = self.L.or_stock_new(w, l, h, cell)
        nx, ny = C.c_int32(), C.c_int32()
        hx, hy = C.c_double(), C.c_double()
        self.L.or_stock_dims(self.p, C.byref(nx), C.byref(ny), C.byref(hx), C.byref(hy))
        self.nx, self.ny, self.hx, self.hy = nx.value, ny.value, hx.value, hy.value

    def __del__(self):
        if getattr(self, "p", None):
            self.L.or_stock_free(self.p)
            self.p = None

    def set_wall_band(self, b):
        self.L.or_stock_set_wall_band(self.p, float(b))

    def _arr(self, fn, ctype):
        import numpy as np
        n = self.nx * self.ny
        ptr = C.cast(getattr(self.L, fn)(self.p), C.POINTER(ctype * n))
        return np.frombuffer(ptr.contents, dtype=np.float32 if ctype is C.c_float else np.int32).copy()

    def H(self):
        return self._arr("or_stock_h", C.c_float)

    def set_H(self, H):
        """overwrite the heights (for reading a recorded state: sample, floor and the wall model
        read H only; the block maxima are not refreshed)"""
        import numpy as np
        h = np.ascontiguousarray(H, dtype=np.float32)
        assert h.size == self.nx * self.ny
        C.memmove(self.L.or_stock_h(self.p), h.ctypes.data, h.nbytes)

    def err(self):
        return self._arr("or_stock_err", C.c_float)

    def wallD(self):
        return self._arr("or_stock_walld", C.c_float)

    def lineCut(self):
        return self._arr("or_stock_linecut", C.c_int32)

    def lineWall(self):
        return self._arr("or_stock_linewall", C.c_int32)

    def stamps(self, stamps_arr, n):
        import numpy as np
        vol = np.zeros(n)
        self.L.or_stamp_batch(self.p, stamps_arr, n, vol.ctypes.data)
        return vol

    def sample(self, x, y):
        return self.L.or_stock_sample(self.p, float(x), float(y))

    def markS(self):
        return self._arr("or_stock_marks", C.c_float)

    def markF(self):
        return self._arr("or_stock_markf", C.c_float)

    def wallErr(self):
        return self._arr("or_stock_wallerr", C.c_float)

    def mark_wall(self, cx, cy, R, nx, ny, val):
        self.L.or_stock_mark_wall(self.p, *[float(v) for v in (cx, cy, R, nx, ny, val)])

    def stamps_ex(self, stamps_arr, opt_arr, n, luts=None, lut_off=None, lut_len=None):
        """Stamps with feed marks / profiles: opt_arr is a StampOpt array (or None); luts
        the concatenated f32 profiles, lut_off / lut_len int32 per profile."""
        import numpy as np
        vol = np.zeros(n)
        keep = [np.ascontiguousarray(a, dt) if a is not None else None
                for a, dt in ((luts, np.float32), (lut_off, np.int32), (lut_len, np.int32))]
        ptr = [a.ctypes.data if a is not None else None for a in keep]
        self.L.or_stamp_batch_ex(self.p, stamps_arr, opt_arr, n, ptr[0], ptr[1], ptr[2],
                                 vol.ctypes.data)
        return vol

    def dirty(self):
        r = (C.c_int32 * 4)()
        self.L.or_stock_dirty(self.p, r)
        return list(r)

    def clear_dirty(self):
        self.L.or_stock_clear_dirty(self.p)

    def block_max(self, bi, bj):
        return self.L.or_stock_block_max(self.p, int(bi), int(bj))

    def block_maxima(self):
        import numpy as np
        bnx, bny = (self.nx + 7) // 8, (self.ny + 7) // 8
        return np.array([[self.block_max(bi, bj) for bi in range(bnx)] for bj in range(bny)],
                        dtype=np.float32).ravel()

    def max_in_disk(self, cx, cy, r, above=float("-inf")):
        out = (C.c_double * 3)()
        self.L.or_stock_max_in_disk(self.p, float(cx), float(cy), float(r), float(above), out)
        return tuple(out)


def fill_engagement(stock, cx, cy, ux, uy, R):
    """Simulator.fillEngagement (src/sim.js:301-319) restated: (hmax, Hb f32[180])"""
    import numpy as np
    Hb = np.zeros(180, np.float32)
    h = lib().or_fill_engagement(stock.p, float(cx), float(cy), float(ux), float(uy), float(R),
                                 Hb.ctypes.data)
    return h, Hb


def immersion(Hb, tipZ):
    """Simulator.immersion (src/sim.js:343-348) restated: (st, ex) or None"""
    import numpy as np
    Hb = np.ascontiguousarray(Hb, np.float32)
    st, ex = C.c_double(), C.c_double()
    ok = lib().or_immersion(Hb.ctypes.data, float(tipZ), C.byref(st), C.byref(ex))
    return (st.value, ex.value) if ok else None


def tip_lut(shape, D, corner_r=0.0, edge_r=None, edge_dz=0.0):
    """tipLut (src/assembly.js:38-54) restated: f32 table."""
    import numpy as np
    er = np.ascontiguousarray(edge_r if edge_r is not None else [], np.float64)
    L = lib()
    n = L.or_tip_lut(int(shape), float(D), float(corner_r), er.ctypes.data, len(er),
                     float(edge_dz), None)
    out = np.zeros(n, np.float32)
    L.or_tip_lut(int(shape), float(D), float(corner_r), er.ctypes.data, len(er), float(edge_dz),
                 out.ctypes.data)
    return out


def tool_model(D, stickout, kSpindle, mHolder=0.05):
    k, fn = C.c_double(), C.c_double()
    lib().or_tool_model(D, stickout, kSpindle, mHolder, C.byref(k), C.byref(fn))
    return k.value, fn.value
