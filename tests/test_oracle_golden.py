"""The oracle against the reference's own outputs (tests/golden/*.json, produced by
running the reference's JavaScript under Node): chatterAt verdicts and chatter indices,
timeDomainLimit brackets, per-step forces, Stock.stamp maps, lobes, toolModel."""
import numpy as np

import golden_util as G
import oracle


def test_chatter_cases_bit_exact():
    g = G.load("chatter")
    assert len(g["cases"]) >= 50
    for case, ref in zip(g["cases"], g["results"]):
        t = oracle.chatter_at(G.cell_from_json(case["cell"]), case["ap"])
        assert t.chatter == ref["chatter"]
        assert t.index == ref["index"] and t.amp == ref["amp"] and t.nz == ref["nz"]


def test_time_domain_limits_bit_exact():
    g = G.load("tdl")
    for case, ref in zip(g["cases"], g["results"]):
        o = case["opts"]
        r = oracle.time_domain_limit(G.cell_from_json(case["cell"]), o["aMin"], o["aMax"], o["iters"])
        assert (r.alim, r.lo, r.hi) == (ref["alim"], ref["lo"], G.num(ref["hi"]))


def test_force_traces():
    g = G.load("trace")
    for case, ref in zip(g["cases"], g["results"]):
        fx, fy, x, y = oracle.engine_trace(G.cell_from_json(case["cell"]), case["ap"], case["steps"])
        linear = not case["cell"]["nonlinear"] or case["cell"]["mat"]["mc"] == 0
        for a, key in ((fx, "Fx"), (fy, "Fy"), (x, "x"), (y, "y")):
            b = G.f64(ref[key])
            if linear:
                assert np.array_equal(a.view(np.int64), b.view(np.int64)), key
            else:   # size effect: pow's last bits (see test_oracle_math)
                assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * np.max(np.abs(b))), key


def test_stock_maps_bit_exact():
    g = G.load("stock")
    for case, ref in zip(g["cases"], g["results"]):
        s = oracle.OracleStock(case["w"], case["l"], case["h"], case["cell"])
        assert (s.nx, s.ny, s.hx, s.hy) == (ref["nx"], ref["ny"], ref["hx"], ref["hy"])
        if case["wallBand"]:
            s.set_wall_band(case["wallBand"])
        vol = s.stamps(G.stamps_from_json(case["stamps"]), len(case["stamps"]))
        assert np.array_equal(s.H().view(np.int32), G.f32(ref["H"]).view(np.int32))
        assert np.array_equal(s.err(), G.f32(ref["err"]))
        assert np.array_equal(s.wallD(), G.f32(ref["wallD"]))
        assert np.array_equal(s.lineCut(), G.i32(ref["lineCut"]))
        assert np.array_equal(s.lineWall(), G.i32(ref["lineWall"]))
        assert np.array_equal(vol, np.array(ref["vol"]))
        for (x, y), v in zip(case["samples"], ref["sample"]):
            assert s.sample(x, y) == G.num(v)


def test_lobes_and_tool_model():
    import ctypes as C
    from mcsim._abi import Mode
    g = G.load("lobes")
    for case, ref in zip(g["cases"], g["results"]):
        p = case["params"]
        mx = p.get("modesX") or [dict(k=p["k"], fn=p["fn"], zeta=p["zeta"])]
        my = p.get("modesY") or mx
        MX = (Mode * len(mx))(*[Mode(m["k"], m["fn"], m["zeta"]) for m in mx])
        MY = (Mode * len(my))(*[Mode(m["k"], m["fn"], m["zeta"]) for m in my])
        cap = 200000
        rpm, alim = np.zeros(cap), np.zeros(cap)
        ll = np.zeros(p["lobes"], dtype=np.int32)
        L = oracle.lib()
        tot = L.or_stability_lobes(p["N"], p["Ktc"], p["Krc"], p["phiSt"], p["phiEx"], p["lobes"],
                                   p["rpmMin"], p["rpmMax"], MX, len(mx), MY, len(my),
                                   rpm.ctypes.data, alim.ctypes.data, cap, ll.ctypes.data)
        want_r = np.concatenate([G.f64(l["rpm"]) for l in ref["lobes"]]) if ref["lobes"] else np.zeros(0)
        want_a = np.concatenate([G.f64(l["alim"]) for l in ref["lobes"]]) if ref["lobes"] else np.zeros(0)
        assert tot == len(want_r)
        assert np.array_equal(rpm[:tot], want_r) and np.array_equal(alim[:tot], want_a)
        nl = int(np.count_nonzero(ll))
        for S, v in zip(case["speeds"], ref["alimAt"]):
            a = L.or_alim_at(rpm.ctypes.data, alim.ctypes.data, ll.ctypes.data, nl, float(S))
            assert a == G.num(v)
        dc = np.zeros(4)
        L.or_directional_coeffs(p["Krc"] / p["Ktc"], p["phiSt"], p["phiEx"], dc.ctypes.data)
        assert np.array_equal(dc, np.array(ref["dc"]))
    t = G.load("toolmodel")
    for case, ref in zip(t["cases"], t["results"]):
        k, fn = oracle.tool_model(case["D"], case["stickout"], case["kSpindle"], case.get("mHolder", 0.05))
        assert abs(k / ref["k"] - 1) < 1e-14 and abs(fn / ref["fn"] - 1) < 1e-14
