"""The device lobes (machinig-sim_amd/csrc/lobes.hip: lobe_boundary, lobe_emit, lobe_read)
against the reference's own outputs (tests/golden/golden_lobes.json: stabilityLobes and
alimAt, src/sld.js:23-76; tests/golden/golden_clobes.json: Simulator.lobes' composite curve,
src/sim.js:262-286) and against the oracle's restatement (oracle/sld.c), bit for bit."""
import ctypes as C

import numpy as np
import pytest

import golden_util as G

pytestmark = pytest.mark.gpu


def test_lobes_match_reference_golden():
    from mcsim import lib
    g = G.load("lobes")
    for case, ref in zip(g["cases"], g["results"]):
        p = case["params"]
        mx = p.get("modesX") or [dict(k=p["k"], fn=p["fn"], zeta=p["zeta"])]
        my = p.get("modesY") or mx
        lobes = lib.stability_lobes(p["N"], p["Ktc"], p["Krc"], p["phiSt"], p["phiEx"], p["lobes"],
                                    p["rpmMin"], p["rpmMax"], [(m["k"], m["fn"], m["zeta"]) for m in mx],
                                    [(m["k"], m["fn"], m["zeta"]) for m in my])
        assert len(lobes) == len(ref["lobes"])
        for (r, a), want in zip(lobes, ref["lobes"]):
            assert np.array_equal(r, G.f64(want["rpm"])) and np.array_equal(a, G.f64(want["alim"]))
        for S, v in zip(case["speeds"], ref["alimAt"]):
            assert lib.alim_at(lobes, S) == G.num(v)
        assert np.array_equal(lib.directional_coeffs(p["Krc"] / p["Ktc"], p["phiSt"], p["phiEx"]),
                              np.array(ref["dc"]))


def _oracle_lobes(N, Ktc, Krc, phiSt, phiEx, nl, rpmMin, rpmMax, mx, my):
    import oracle
    from mcsim._abi import Mode
    cap = 400000
    rpm, alim = np.zeros(cap), np.zeros(cap)
    ll = np.zeros(max(1, nl), dtype=np.int32)
    MX = (Mode * len(mx))(*[Mode(*m) for m in mx])
    MY = (Mode * len(my))(*[Mode(*m) for m in my])
    tot = oracle.lib().or_stability_lobes(N, Ktc, Krc, phiSt, phiEx, nl, rpmMin, rpmMax, MX, len(mx), MY, len(my),
                                          rpm.ctypes.data, alim.ctypes.data, cap, ll.ctypes.data)
    return rpm, alim, ll, tot


def test_lobes_match_oracle_multimode():
    """seeded 1-4 modes per axis (every sample count 400..1600), partial arcs; lobe lengths,
    points and alimAt reads equal the oracle's"""
    import oracle
    from mcsim import lib
    rng = np.random.default_rng(3)
    for case in range(24):
        nmx, nmy = (int(v) for v in rng.integers(1, 5, 2))
        mx = [(10 ** rng.uniform(3.7, 4.7), 10 ** rng.uniform(2.7, 3.6), rng.uniform(0.01, 0.05)) for _ in range(nmx)]
        my = [(10 ** rng.uniform(3.7, 4.7), 10 ** rng.uniform(2.7, 3.6), rng.uniform(0.01, 0.05)) for _ in range(nmy)]
        N = int(rng.integers(2, 5))
        phiSt = float(rng.uniform(0, 3.0))
        lobes = lib.stability_lobes(N, 700.0, 210.0, phiSt, np.pi, 12, 2000.0, 30000.0, mx, my)
        rpm, alim, ll, tot = _oracle_lobes(N, 700.0, 210.0, phiSt, np.pi, 12, 2000.0, 30000.0, mx, my)
        assert [len(l[0]) for l in lobes] == [int(v) for v in ll if v > 0]
        got_r = np.concatenate([l[0] for l in lobes]) if lobes else np.zeros(0)
        got_a = np.concatenate([l[1] for l in lobes]) if lobes else np.zeros(0)
        assert len(got_r) == tot
        assert np.array_equal(got_r, rpm[:tot]) and np.array_equal(got_a, alim[:tot])
        for S in np.linspace(2000, 30000, 37):
            want = oracle.lib().or_alim_at(rpm.ctypes.data, alim.ctypes.data, ll.ctypes.data,
                                           int(np.count_nonzero(ll)), float(S))
            assert lib.alim_at(lobes, S) == want


def test_lobes_truncated_output():
    """max_pts cuts the points short but every lobe's full length is still recorded"""
    from mcsim import lib
    a = lib.lobe_args(3, 700.0, 210.0, 0.5, np.pi, 8, 2000.0, 30000.0, [(2e4, 900.0, 0.02)])
    full = lib.stability_lobes(3, 700.0, 210.0, 0.5, np.pi, 8, 2000.0, 30000.0, [(2e4, 900.0, 0.02)])
    cap = 10
    rpm, alim = np.zeros(cap), np.zeros(cap)
    ll = np.zeros(8, dtype=np.int32)
    n = C.c_int(0)
    assert lib.load().mc_stability_lobes(C.byref(a), rpm.ctypes.data, alim.ctypes.data, cap, ll.ctypes.data,
                                         C.byref(n)) == 0
    assert n.value == cap
    assert list(ll[ll > 0]) == [len(l[0]) for l in full]
    assert np.array_equal(rpm, np.concatenate([l[0] for l in full])[:cap])


def test_analytic_limits_of_grid_cells():
    """mc_analytic_limits over the 64 x 256 lobe grid equals the oracle's alimAt on the
    oracle's lobes of each tool over the tool's speed window (+inf where no lobe covers)"""
    import oracle
    from mcsim import lib, workloads as W
    ns, nt = 256, 64
    cells, _ = W.sld_grid(n_speeds=ns, n_tools=nt)
    got = lib.analytic_limits(cells)
    for t in range(nt):
        group = cells[t * ns:(t + 1) * ns]
        c = group[0]
        smin, smax = min(x.S for x in group), max(x.S for x in group)
        fmax = max([c.mx[j].fn for j in range(c.nmx)] + [c.my[j].fn for j in range(c.nmy)])
        nl = int(np.ceil(60 * 1.7 * fmax / (c.tool.N * 0.8 * smin))) + 2
        mx = [(c.mx[j].k, c.mx[j].fn, c.mx[j].zeta) for j in range(c.nmx)]
        my = [(c.my[j].k, c.my[j].fn, c.my[j].zeta) for j in range(c.nmy)]
        rpm, alim, ll, _ = _oracle_lobes(c.tool.N, c.mat.Ktc, c.mat.Krc, c.phiSt, c.phiEx, nl, smin, smax, mx, my)
        for q, x in enumerate(group):
            want = oracle.lib().or_alim_at(rpm.ctypes.data, alim.ctypes.data, ll.ctypes.data,
                                           int(np.count_nonzero(ll)), float(x.S))
            assert got[t * ns + q] == want, (t, q)


def test_analytic_limits_mixed_groups():
    """cells of several tools interleaved, one speed outside every lobe: each cell's limit
    equals its own group's, whatever the order the cells come in"""
    from mcsim import lib, workloads as W
    cells, _ = W.sld_grid(n_speeds=16, n_tools=5, seed=9)
    want = lib.analytic_limits(cells)
    perm = np.random.default_rng(1).permutation(len(cells))
    got = lib.analytic_limits([cells[i] for i in perm])
    assert np.array_equal(got, want[perm])
    one = lib.analytic_limits(cells[:16])
    assert np.array_equal(one, want[:16])


# ---- Simulator.lobes (src/sim.js:262-286): the composite curve and the diagram at rest ----

def test_composite_lobes_match_reference():
    """mc_composite_lobes on the recorded nodes gives Simulator.lobes' composite curve bit for
    bit; the diagram at rest is stabilityLobes of modalNow(0) over 2000-40000 rpm"""
    from mcsim import lib
    from test_lobes import BASE, _clobes_runs, _nodes
    seen = [0, 0]
    for res, run, arc, comp in _clobes_runs():
        ref = run["lobes"]
        if comp:
            assert len(run["calls"]) == 43 and len(ref) == 1
            nodes = [lib.lobe_args(d["N"], d["Ktc"], d["Krc"], d["phiSt"], d["phiEx"], d["lobes"], d["rpmMin"],
                                   d["rpmMax"], d["mx"], d["my"]) for d in _nodes(res, run, arc)]
            rpm, alim = lib.composite_lobes(nodes, 25.0, BASE["rpmMax"] * 1.1)
            assert np.array_equal(rpm, G.zf64(ref[0]["rpm"])) and np.array_equal(alim, G.zf64(ref[0]["alim"]))
            seen[0] += 1
        elif run["calls"]:
            S, mx, my = run["calls"][0]
            assert S == 0
            lob = lib.stability_lobes(res["N"], res["Ktc"], res["Krc"], arc[0], arc[1], BASE["lobes"], BASE["rpmMin"],
                                      BASE["rpmMax"], mx, my)
            assert len(lob) == len(ref)
            for (r, a), want in zip(lob, ref):
                assert np.array_equal(r, G.zf64(want["rpm"])) and np.array_equal(a, G.zf64(want["alim"]))
            seen[1] += 1
    assert seen == [11, 2]


def test_composite_lobes_batched_against_oracle():
    """every recorded curve's nodes in one call, plus a node with no lobes (1e3 throughout)
    and a wide node (1600 speeds): each stretch equals the oracle's composite curve"""
    import oracle
    from mcsim import lib
    from test_lobes import _clobes_runs, _nodes
    nodes = []
    for res, run, arc, comp in _clobes_runs():
        if comp:
            nodes += _nodes(res, run, arc)
    nodes.append(dict(nodes[0], lobes=0))
    nodes.append(dict(nodes[1], rpmMin=2000.0, rpmMax=42000.0))
    args = [lib.lobe_args(d["N"], d["Ktc"], d["Krc"], d["phiSt"], d["phiEx"], d["lobes"], d["rpmMin"], d["rpmMax"],
                          d["mx"], d["my"]) for d in nodes]
    rpm, alim = lib.composite_lobes(args, 25.0, 44000.0)
    want_r, want_a = oracle.composite_lobes(nodes, 25.0, 44000.0)
    assert np.array_equal(rpm, want_r) and np.array_equal(alim, want_a)
    assert np.all(alim[-1640:-1600] == 1e3)
