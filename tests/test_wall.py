"""The thin-wall model (wallAt, src/wall.js:11-112) and the simulator's sub-step loop with it on
(workpieceFlex, src/sim.js:407-408,429,443), restated by the oracle (oracle/wall.c,
oracle/simloop.c), against the reference (golden_simw.json):

- wallAt call by call: at recorded calls of three reference Simulator runs, the whole input
  (the stock's heights, the tool position, the engagement map, the load height, the material
  and the tool's static stiffness) and the reference wallAt's answer on it; every field
  bit for bit, null answers included;
- whole runs: the walls found at each cutting sub-step, the deflected stamps, the marks on
  the wall-error map and every final map.  With the linear force law the run is bit for bit.
  With the size effect on, and on the run with 623 walls, the engine's and wallAt's x ** 3
  use the restated fdlibm pow, which differs from Node 12's Math.pow by one ulp on a few per
  cent of inputs: the walls are found at the same sub-steps with the same geometry, their
  stiffness within 1e-14, the positions within 1e-12 mm and every map bit for bit;
- against the mathematics: a straight wall of uniform thickness, where the model's
  compliance is the midpoint sum of the cantilever integral, a^3 (1/3 - 1/(12 NZ^2)) / (E I)
  (NZ = 10 heights), its thickness the wall's to half a cell and its normal the wall's."""
import math

import numpy as np
import pytest

import golden_util as G

WALL_GEOM = ("nx", "ny", "t", "tRoot", "hWall", "a", "bEff", "base", "cx", "cy")
WALL_STIFF = ("k", "m", "fn", "I")


@pytest.fixture(scope="module")
def simw():
    return G.load("simw")


def test_atan2_matches_node():
    """the restated Math.atan2 (fdlibm e_atan2) the wall normal comes from"""
    import oracle
    from test_oracle_math import NODE, bits, node_math
    if NODE is None:
        pytest.skip("node not installed")
    rng = np.random.default_rng(11)
    y, x = rng.normal(0, 3, 60000), rng.normal(0, 3, 60000)
    y[:8] = [0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1e-300, 5.0]
    x[:8] = [1.0, 1.0, -1.0, -1.0, 0.0, -0.0, -1e300, 1.0]
    assert np.array_equal(bits(oracle.math_batch("atan2", y, x)), bits(node_math("atan2", y, x)))


def test_wall_at_matches_reference_calls(simw):
    import oracle
    from mcsim._abi import WALL_FIELDS
    walls = nulls = 0
    for case, res in zip(simw["cases"], simw["results"]):
        st = case["stock"]
        for sn in res["snaps"]:
            s = oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"])
            s.set_H(G.zf32(sn["H"]))
            w = oracle.wall_at(s, sn["x"], sn["y"], sn["z"], sn["R"], G.f32(sn["Hb"]), sn["ux"], sn["uy"],
                               sn["zLoad"], sn["E"], sn["rho"], sn["kTool"])
            if sn["wall"] is None:
                assert w.found == 0, (case["cell"], sn["sub"])
                nulls += 1
            else:
                assert w.found == 1, (case["cell"], sn["sub"])
                got = [getattr(w, f) for f in WALL_FIELDS]
                assert np.array_equal(got, sn["wall"]), (case["cell"], sn["sub"])
                walls += 1
    assert walls >= 10 and nulls >= 8


def _flex_run(oracle, case, res):
    st = case["stock"]
    act = oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"])
    ide = oracle.OracleStock(st["w"], st["l"], st["h"], case["cell"])
    ide.set_wall_band(2.0)
    eng = oracle.OracleEngine()
    cell, fnMax = G.simx_cell(res)
    info = eng.configure(cell, fnMax)
    _, kte, kre = res["wears"][0]
    eng.set_wear(kte, kre)
    sub, n = G.simx_program(res)
    out, walls = oracle.sim_run(act, ide, eng, sub, n, res["toolR"], res["fluteLen"], 0.0, info.stepsPerTooth / 4,
                                G.simx_tool(case, res, oracle.tip_lut)[0],
                                wall=dict(E=res["E"], rho=res["rho"], kTool=res["kTool"]))
    return act, ide, sub, n, out, walls


def check_flex_run(case, res, sub, n, out, walls, act_maps, ide_maps, exact, vol_rtol=0.0):
    """one run with the wall model against the reference's record; exact: every double bit
    for bit, else the tolerances of the module docstring; vol_rtol: the stamps' removed
    volumes (a device sums a stamp's nodes in another order)"""
    from mcsim._abi import MC_WALL_FN_CAP, WALL_FIELDS
    assert len(res["cfgs"]) == 1 and res["cfgs"][0][10] == MC_WALL_FN_CAP
    a = res["actual"]
    cx = np.array([sub[k].x + out[k].dx for k in range(n)])
    cy = np.array([sub[k].y + out[k].dy for k in range(n)])
    ev = np.array([o.errVal for o in out])
    tol = 0.0 if exact else 1e-12
    assert np.max(np.abs(cx - G.f64(a["cx"]))) <= tol and np.max(np.abs(cy - G.f64(a["cy"]))) <= tol
    assert np.allclose(ev, G.f64(a["errVal"]), rtol=0.0 if exact else 1e-11, atol=0.0)
    assert np.allclose([o.vol for o in out], G.f64(a["vol"]), rtol=vol_rtol, atol=0.0)
    runs = dict((k, m) for k, m in res["runs"])
    assert [o.nsteps for o in out] == [runs.get(k, 0) for k in range(n)]
    # setWork's argument at every cutting sub-step: a wall at the same sub-steps, and its fields
    w = res["works"]
    found = [k for k in range(n) if walls[k].found]
    assert [s for s, f in zip(w["sub"], w["wall"]) if f] == found
    # the other setWork calls drop a wall whose deflection decayed out of the cut (:429)
    extra = [(s, f) for s, f in zip(w["sub"], w["wall"]) if s not in runs]
    assert len(w["sub"]) == len(runs) + len(extra) and all(f == 0 for _, f in extra)
    F = G.f64(w["fields"]).reshape(-1, len(WALL_FIELDS))
    got = np.array([[getattr(walls[k], f) for f in WALL_FIELDS] for k in found])
    geo = [WALL_FIELDS.index(f) for f in WALL_GEOM]
    stiff = [WALL_FIELDS.index(f) for f in WALL_STIFF]
    assert np.array_equal(got[:, geo], F[:, geo])
    assert np.allclose(got[:, stiff], F[:, stiff], rtol=0.0 if exact else 1e-14, atol=0.0)
    assert len(found) > 0
    # the marks on the wall-error map after each stamp while a wall is held
    mk = res["marks"]
    marked = [k for k in range(n) if walls[k].marked]
    assert [s - 1 for s in mk["sub"]] == marked
    M = G.f64(mk["args"]).reshape(-1, 6)
    ours = np.array([[cx[k], cy[k], res["toolR"], walls[k].wnx, walls[k].wny, walls[k].markVal] for k in marked])
    assert np.allclose(ours, M, rtol=0.0 if exact else 1e-12, atol=0.0 if exact else 1e-12)
    for name, maps, ref in (("actual", act_maps, res["stock"]), ("commanded", ide_maps, res["commanded"])):
        for key, got_m in maps.items():
            want = G.zi32(ref[key]) if key in ("lineCut", "lineWall") else G.zf32(ref[key])
            assert np.array_equal(got_m, want, equal_nan=True), (case["cell"], name, key)


def _oracle_maps(s, wall_err=True):
    m = dict(H=s.H(), err=s.err(), wallD=s.wallD(), markS=s.markS(), markF=s.markF(), lineCut=s.lineCut(),
             lineWall=s.lineWall())
    if wall_err:
        m["wallErr"] = s.wallErr()
    return m


def test_oracle_flex_loop_matches_reference_runs(simw):
    import oracle
    for ci, (case, res) in enumerate(zip(simw["cases"], simw["results"])):
        act, ide, sub, n, out, walls = _flex_run(oracle, case, res)
        check_flex_run(case, res, sub, n, out, walls, _oracle_maps(act), _oracle_maps(ide, False),
                       exact=not case["nonlinear"] and ci == 0)


def test_wall_model_on_a_uniform_wall():
    """a straight wall 2.5 mm thick and 6 mm tall along y between two cleared floors, the cutter
    against its face: the model's thickness is the wall's to half a cell, its normal the
    wall's, and its stiffness the midpoint sum of the cantilever's compliance integral"""
    import oracle
    cell, h0 = 0.05, 6.0
    s = oracle.OracleStock(30.0, 40.0, 10.0, cell)
    X = np.arange(s.nx) * s.hx
    H = np.where((X >= 15.0) & (X <= 17.5), 0.0, -h0).astype(np.float32)
    s.set_H(np.tile(H, s.ny))
    R, z = 4.0, -h0
    px, py = 17.5 + R - 0.3, 20.0   # 0.3 mm into the wall's +x face, feeding along +y
    ap = h0
    zLoad = z + ap / 2
    # the engagement map of a cutter whose left side is in the wall
    Hb = np.full(180, -np.inf, np.float32)
    for b in range(180):
        phi = (b + 0.5) / 180 * 2 * math.pi
        ex, ey = math.sin(phi) * 0.0 + math.cos(phi) * -1.0, math.sin(phi) * 1.0 + math.cos(phi) * 0.0
        if math.sin(phi) > 0.02 and px + (R - 0.5 * cell) * ex <= 17.5:
            Hb[b] = 0.0
    E, rho, kTool = 70300.0, 2680.0, 1e9
    w = oracle.wall_at(s, px, py, z, R, Hb, 0.0, 1.0, zLoad, E, rho, kTool)
    assert w.found == 1
    # the thinnest of the crossings 5 degrees apart lies within 2.5 degrees of the wall normal;
    # its length is the material between the contact point (on the cutter's edge, up to 0.3 mm
    # into the face here) and the far face, along it, to a cell (the bilinear ramp at the face)
    ang = math.degrees(math.atan2(w.ny, w.nx))
    assert abs(abs(ang) - 180) <= 2.5
    assert 17.2 <= w.cx <= 17.5 and py <= w.cy <= py + R   # the middle of the leading half's contact
    assert abs(w.t - (w.cx - 15.0) / abs(w.nx)) <= cell
    assert w.hWall == pytest.approx(h0, abs=1e-9) and w.base == pytest.approx(-h0, abs=1e-9)
    a = w.a
    assert a == pytest.approx(zLoad + h0)
    I = w.bEff * w.t ** 3 / 12
    midpoint = a ** 3 * (1 / 3 - 1 / 1200) / (E * I)
    assert 1 / w.k == pytest.approx(midpoint, rel=1e-12)
    # the mass is 0.24 of the wall strip's unless the 5 kHz cap applies
    m = 0.24 * rho * 1e-9 * w.t * w.bEff * w.hWall
    fn = math.sqrt(w.k * 1e3 / m) / (2 * math.pi)
    if fn > 5000:
        assert w.fn == 5000 and w.m == pytest.approx(w.k * 1e3 / (2 * math.pi * 5000) ** 2, rel=1e-12)
    else:
        assert w.m == pytest.approx(m, rel=1e-12) and w.fn == pytest.approx(fn, rel=1e-12)
