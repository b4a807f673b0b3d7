"""The zigzag program (512 x 512 stock at 0.1 mm, the flat 10 mm and ball 6 mm layers of the
heightmap workload) run through the simulator's sub-step loop on the device: every sub-step's
engagement, engine run and stamps, layer after layer on one pair of stocks, with the rigid mode
of the reference's target test.  Against the oracle's loop (oracle/simloop.c) on the same
inputs: each sub-step's deflection, axial depth, engagement arc, step count, vibration value
and peak force |(Fx, Fy)| bit for bit, the volume to 1e-12, the per-segment maximum force
(src/cutting.js:280-281) and the final heights bit for bit.  The first 25,000 sub-steps cover
the first layer (a plunge and 22 raster lines with the flat end mill) and the start of the
second (the ball end mill, its engine reconfigured)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_SUB = 25000


def _run(lib_mod, oracle_mod=None):
    from mcsim import workloads as W, path
    from mcsim._abi import make_cell
    layers = W.zigzag_layers()
    cell_of = lambda t: make_cell(t, W.ZIGZAG_MODES, None, W.MAT_A5052, W.ZIGZAG_SPEED, (0.0, 0.0), 0.0, 1, True)
    sub_of = lambda L: path.sim_substeps(L["x"], L["y"], L["z"], L["rapid"], L["line"], L["F"], L["S"],
                                         L["tool"]["N"], 0.1)
    typ = lambda tool: W.MC_TOOL_FLAT if tool["type"] == "flat" else W.MC_TOOL_BALL
    if oracle_mod is None:
        a, i, e = lib_mod.Stock(51.1, 51.1, 20, 0.1), lib_mod.Stock(51.1, 51.1, 20, 0.1), lib_mod.Engine()

        def run(a, i, e, sub, n, tool, share):
            return lib_mod.sim_run([dict(actual=a, ideal=i, engine=e, sub=sub, nsub=n, R=tool["R"],
                                         fluteLen=tool["fluteLen"], rw=0.0, share=share, type=typ(tool))])[0]
    else:
        a, i, e = oracle_mod.OracleStock(51.1, 51.1, 20, 0.1), oracle_mod.OracleStock(51.1, 51.1, 20, 0.1), \
            oracle_mod.OracleEngine()

        def run(a, i, e, sub, n, tool, share):
            return oracle_mod.sim_run(a, i, e, sub, n, tool["R"], tool["fluteLen"], 0.0, share, typ(tool))
    i.set_wall_band(2.0)
    outs, lines = W.zigzag_coupled(layers, (a, i), e, run, cell_of, sub_of, max_substeps=N_SUB)
    return outs, lines, a.H(), i.H(), W.segment_max_force(outs, lines)


def test_zigzag_coupled_matches_oracle():
    import oracle
    from mcsim import lib
    lib.load()
    print("device:", lib.device_info(), flush=True)
    d_out, d_lines, d_H, d_I, d_seg = _run(lib)
    print("device: %d sub-steps, %d segments" % (len(d_out), len(d_seg)), flush=True)
    o_out, o_lines, o_H, o_I, o_seg = _run(None, oracle)
    assert len(d_out) == len(o_out) == N_SUB and d_lines == o_lines
    bad = [k for k, (p, q) in enumerate(zip(d_out, o_out))
           if (p.dx, p.dy, p.ap, p.errVal, p.maxF, p.nsteps, p.cut, p.arc, p.phiSt, p.phiEx) !=
              (q.dx, q.dy, q.ap, q.errVal, q.maxF, q.nsteps, q.cut, q.arc, q.phiSt, q.phiEx)
           or abs(p.vol - q.vol) > 1e-12 * max(1.0, abs(q.vol))]
    assert not bad, (len(bad), bad[:5])
    assert d_seg == o_seg
    assert np.array_equal(d_H.view(np.int32), o_H.view(np.int32))
    assert np.array_equal(d_I.view(np.int32), o_I.view(np.int32))
    # the layers cut: every raster segment of the first layer took load, the ball layer too
    assert sum(1 for v in d_seg.values() if v > 0) >= 0.9 * len(d_seg)
    from mcsim import workloads as W
    layers = W.zigzag_layers()
    assert set(layers[1]["line"][1:]) & set(d_lines)               # the ball layer ran too
    print("per-segment max force: %d segments, max %.1f N, median %.1f N" % (
        len(d_seg), max(d_seg.values()), float(np.median(list(d_seg.values())))), flush=True)
