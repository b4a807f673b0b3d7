"""Path A on the GPU against the oracle and the reference's golden vectors: identical
verdicts, chatter indices and brackets (bit-exact FP64), per-step forces bit-exact."""
import numpy as np
import pytest

import golden_util as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    from mcsim import lib
    lib.load()
    print("device:", lib.device_info(), flush=True)
    return lib


@pytest.fixture(scope="module")
def orc():
    import oracle
    return oracle


def test_chatter_golden_bit_exact(mc, orc):
    g = G.load("chatter")
    cells = [G.cell_from_json(c["cell"]) for c in g["cases"]]
    ap = [c["ap"] for c in g["cases"]]
    out = mc.chatter_batch(cells, ap)
    for i, (r, ref) in enumerate(zip(out, g["results"])):
        o = orc.chatter_at(cells[i], ap[i])
        assert (r.index, r.amp, r.once, r.chatter, r.nz) == (o.index, o.amp, o.once, o.chatter, o.nz), i
        assert r.chatter == ref["chatter"], i
        assert r.nz == ref["nz"]
        # the reference's own index: bit-equal (pow differs from V8's in rare last bits,
        # which the float32 sample stores absorb in every golden case)
        assert r.index == ref["index"], (i, r.index, ref["index"])


def test_time_domain_limit_golden(mc, orc):
    g = G.load("tdl")
    cells = [G.cell_from_json(c["cell"]) for c in g["cases"]]
    out = mc.bisect_batch(cells, 0.05, 10.0, 10)
    for i, (r, ref) in enumerate(zip(out, g["results"])):
        assert r.alim == ref["alim"] and r.lo == ref["lo"] and r.hi == G.num(ref["hi"]), i
        o = orc.time_domain_limit(cells[i])
        assert (r.alim, r.lo, r.hi, r.open, r.trials, r.steps, r.visits) == \
               (o.alim, o.lo, o.hi, o.open, o.trials, o.steps, o.visits), i


def test_engine_trace_bit_exact(mc, orc):
    g = G.load("trace")
    for case, ref in zip(g["cases"], g["results"]):
        c = G.cell_from_json(case["cell"])
        fx, fy, x, y = mc.engine_trace(c, case["ap"], case["steps"])
        ofx, ofy, ox, oy = orc.engine_trace(c, case["ap"], case["steps"])
        for a, b in ((fx, ofx), (fy, ofy), (x, ox), (y, oy)):
            assert np.array_equal(a.view(np.int64), b.view(np.int64))
        if case["cell"]["mat"]["mc"] == 0 or not case["cell"]["nonlinear"]:
            assert np.array_equal(fx, G.f64(ref["Fx"])) and np.array_equal(x, G.f64(ref["x"]))
        else:  # size effect: pow's last bits are not pinned to V8's (~1e-12 relative)
            assert np.allclose(fx, G.f64(ref["Fx"]), rtol=1e-9, atol=1e-9)


def test_simulator_engine_traces(mc, orc):
    """The engine as the simulator drives it (feed direction, stock map, runout, fnMax,
    velocity set from outside) against the reference's traces and the oracle"""
    g = G.load("enginex")
    for case, ref in zip(g["cases"], g["results"]):
        cell, spec = G.enginex_case(case)
        got = mc.engine_trace_ex(cell, spec, case["steps"])
        want = orc.engine_trace_ex(cell, spec, case["steps"])
        for a, b in zip(got, want):          # the oracle: bit for bit
            assert np.array_equal(a, b)
        for a, key in zip(got, ("Fx", "Fy", "x", "y")):
            r = G.f64(ref[key])
            if not case["nonlinear"]:
                assert np.array_equal(a, r), key
            else:   # size effect: pow's last bits are not pinned to V8's
                assert np.allclose(a, r, rtol=1e-9, atol=1e-12 * np.max(np.abs(r))), key


def test_grid_subset_bisection_matches_oracle(mc, orc):
    from mcsim import workloads as W
    cells, opts = W.sld_grid(n_speeds=8, n_tools=6, seed=3)
    out = mc.bisect_batch(cells, **opts)
    ref = orc.bisect_batch(__import__("mcsim._abi", fromlist=["cell_array"]).cell_array(cells),
                           len(cells), threads=8, **opts)
    for i in range(len(cells)):
        a, b = out[i], ref[i]
        assert (a.alim, a.lo, a.hi, a.open, a.trials, a.steps, a.visits) == \
               (b.alim, b.lo, b.hi, b.open, b.trials, b.steps, b.visits), i
    # open cells (their trees stop once the aMax check answers) are part of the case
    assert sum(a.open for a in out) >= 5


def test_bracket_search_matches_oracle(mc, orc):
    from mcsim import workloads as W
    cells, opts = W.sld_40rpm()
    cells = cells[::5]
    lo = [0.3, 0.5, 0.05, 2.0, 0.2, 1.0, 0.9, 0.4]
    hi = [0.6, 5.0, 0.07, 9.0, 9.9, 1.3, 12.0, 0.45]   # includes invalid / failing hints
    out, fb = mc.bracket_batch(cells, lo, hi, 0.05, **opts)
    for i, c in enumerate(cells):
        r, f = orc.bracket_limit(c, lo[i], hi[i], 0.05, **opts)
        assert (out[i].alim, out[i].lo, out[i].hi, out[i].trials, int(fb[i])) == \
               (r.alim, r.lo, r.hi, r.trials, f), i


def test_speculative_search_equals_serial_search(mc, orc, monkeypatch):
    """the speculative wave-per-cell search and the one-trial-at-a-time search agree
    exactly (depths, brackets, serial-path trial/step/visit counts)"""
    from mcsim import workloads as W
    cells, opts = W.sld_grid(n_speeds=12, n_tools=10, seed=7)
    spec = mc.bisect_batch(cells, **opts)
    monkeypatch.setenv("MC_SEARCH", "serial")
    ser = mc.bisect_batch(cells, **opts)
    for i in range(len(cells)):
        a, b = spec[i], ser[i]
        assert (a.alim, a.lo, a.hi, a.open, a.trials, a.steps, a.visits, a.engaged) == \
               (b.alim, b.lo, b.hi, b.open, b.trials, b.steps, b.visits, b.engaged), i


def test_mixed_batch_routes_every_cell_exactly(mc, orc):
    """one batch with cells the search kernel takes and cells it leaves to the
    one-trial-per-group kernel: size-effect exponents 0.5, 1, 2 and 9 (outside the search
    kernel's power function), a 4-flute tool needing 160 visits per step at aMax (above its
    128), and a 12 mm flute whose deep trials change the slice pitch (more slice lags than
    its geometry table holds: the cell is rerun there).  Every result equals the oracle's."""
    from mcsim import workloads as W
    from mcsim._abi import make_cell
    t3 = dict(N=3, R=6.0, helixDeg=30.0, type="flat", fluteLen=24.0)
    t4 = dict(N=4, R=6.0, helixDeg=30.0, type="flat", fluteLen=24.0)
    t2 = dict(N=2, R=5.0, helixDeg=35.0, type="ball", fluteLen=12.0)
    modes = [(2e4, 1000.0, 0.03)]
    arc = (W.JS_PI / 2, W.JS_PI)
    cells = []
    for j, mcx in enumerate((0.1, 0.5, 1.0, 2.0, 9.0, 0.25, 0.0)):
        mat = dict(W.MAT_A5052, mc=mcx)
        cells.append(make_cell(t3, modes, None, mat, 4000 + 900 * j, arc, 0.05, 20, True))
    cells.append(make_cell(t4, modes, None, W.MAT_A5052, 5200, arc, 0.05, 12, True))
    cells.append(make_cell(t2, modes, None, W.MAT_A5052, 6100, arc, 0.05, 12, True))
    opts = dict(aMin=0.05, aMax=20.0, iters=8)
    out = mc.bisect_batch(cells, **opts)
    for i, c in enumerate(cells):
        o = orc.time_domain_limit(c, **opts)
        assert (out[i].alim, out[i].lo, out[i].hi, out[i].open, out[i].trials, out[i].steps,
                out[i].visits) == (o.alim, o.lo, o.hi, o.open, o.trials, o.steps, o.visits), i


def test_graph_replays_match_eager_launches(mc):
    """the captured launch sequence, replayed several times with eager work and
    synchronisations in between, returns exactly what direct launches return"""
    from mcsim import workloads as W
    cells, opts = W.sld_grid(n_speeds=10, n_tools=8, seed=21)
    keep = mc.search_tuning()
    try:
        mc.set_search_tuning(**dict(keep, use_graph=0))
        eager = [(r.alim, r.lo, r.hi, r.trials, r.steps, r.visits) for r in mc.bisect_batch(cells, **opts)]
        mc.set_search_tuning(**dict(keep, use_graph=1))
        for rep in range(3):
            got = [(r.alim, r.lo, r.hi, r.trials, r.steps, r.visits) for r in mc.bisect_batch(cells, **opts)]
            assert got == eager, rep
            mc.chatter_batch(cells[:4], [0.5] * 4)   # eager work of the same stream in between
    finally:
        mc.set_search_tuning(**keep)


def test_autotuner_keeps_results(mc):
    from mcsim import autotune, workloads as W
    cells, opts = W.sld_grid(n_speeds=8, n_tools=6, seed=22)
    keep = mc.search_tuning()
    try:
        ref = [(r.alim, r.lo, r.hi, r.trials) for r in mc.bisect_batch(cells, **opts)]
        best, ms, table = autotune.tune_search(cells, opts, fracs=(1e9, 0.5, 0.3), wide_max=(16, 64))
        print("autotune best", best, "%.2f ms" % ms, flush=True)
        assert mc.search_tuning()["wide_frac"] == best["wide_frac"]
        assert [(r.alim, r.lo, r.hi, r.trials) for r in mc.bisect_batch(cells, **opts)] == ref
    finally:
        mc.set_search_tuning(**keep)
