"""Simulator.lobes' composite curve (src/sim.js:262-286) in the oracle against the
reference's recorded curves (tests/golden/golden_clobes.json), and the argument checks of
mc_composite_lobes, which answer before any device work.  The device lobes themselves are
tested in test_gpu_lobes.py."""
import numpy as np

import golden_util as G


# ---- Simulator.lobes (src/sim.js:262-286): the composite curve and the diagram at rest ----

BASE = dict(lobes=30, rpmMin=2000, rpmMax=40000)


def _node(res, S0, mx, my, phiSt, phiEx):
    """the speed node the reference's loop builds at S0 (its modes: modalNow(S0 + 500))"""
    return dict(N=res["N"], Ktc=res["Ktc"], Krc=res["Krc"], phiSt=phiSt, phiEx=phiEx, lobes=BASE["lobes"],
                rpmMin=S0, rpmMax=S0 + 1000, mx=mx, my=my)


def _clobes_runs():
    """(case result, run, arc, composite?) for every lobes() call the fixtures recorded"""
    g = G.load("clobes")
    for case, res in zip(g["cases"], g["results"]):
        for arc, run in zip(case["arcs"], res["runs"]):
            yield res, run, arc, bool(res["dependent"] and not arc[2])


def _nodes(res, run, arc):
    # modalNow(S0 + LOBE_NODE / 2) per node: S0 = S - 500 exactly (integers)
    return [_node(res, S - 500, mx, my, arc[0], arc[1]) for S, mx, my in run["calls"]]


def test_composite_lobes_oracle_pinned():
    """the oracle's composite restatement against the same recorded curves; then planted
    mistakes (the next node's modes, a window read to its end inclusive, the uncovered
    speeds left at +inf) each break it"""
    import oracle
    for res, run, arc, comp in _clobes_runs():
        if not comp:
            continue
        ref = run["lobes"][0]
        want_r, want_a = G.zf64(ref["rpm"]), G.zf64(ref["alim"])
        nodes = _nodes(res, run, arc)
        rpm, alim = oracle.composite_lobes(nodes, 25.0, BASE["rpmMax"] * 1.1)
        assert np.array_equal(rpm, want_r) and np.array_equal(alim, want_a)
        shifted = [dict(d, mx=e["mx"], my=e["my"]) for d, e in zip(nodes, nodes[1:] + nodes[-1:])]
        assert not np.array_equal(oracle.composite_lobes(shifted, 25.0, 44000.0)[1], want_a)
        wide = [dict(d, rpmMax=d["rpmMax"] + 1e-9 * d["rpmMax"]) for d in nodes]
        assert len(oracle.composite_lobes(wide, 25.0, 44000.0)[0]) != len(want_r)
        if np.any(want_a == 1e3):
            assert np.all(alim[want_a == 1e3] == 1e3) and not np.any(np.isinf(alim))


def test_composite_lobes_checked():
    """the size query, a too-small output and a bad step are reported before any device work"""
    import ctypes as C
    from mcsim import lib
    from mcsim._abi import LobeArgs
    for res, run, arc, comp in _clobes_runs():
        if comp:
            curve = [lib.lobe_args(d["N"], d["Ktc"], d["Krc"], d["phiSt"], d["phiEx"], d["lobes"], d["rpmMin"],
                                   d["rpmMax"], d["mx"], d["my"]) for d in _nodes(res, run, arc)]
            break
    arr = (LobeArgs * 2)(*curve[:2])
    n = C.c_int(0)
    assert lib.load().mc_composite_lobes(arr, 2, 25.0, 44000.0, None, None, 0, C.byref(n)) == 0 and n.value == 80
    out = np.zeros(80)
    assert lib.load().mc_composite_lobes(arr, 2, 25.0, 44000.0, out.ctypes.data, out.ctypes.data, 79, C.byref(n)) != 0
    assert lib.load().mc_composite_lobes(arr, 2, 0.0, 44000.0, None, None, 0, C.byref(n)) != 0
    assert lib.load().mc_composite_lobes(arr, 0, 25.0, 44000.0, out.ctypes.data, out.ctypes.data, 80, C.byref(n)) == 0
    assert n.value == 0
