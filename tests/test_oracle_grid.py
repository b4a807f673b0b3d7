"""The oracle's timeDomainLimit (oracle/engine.c) against the reference's own answers on the
256-speed x 64-tool lobe grid (golden_tdl_grid.json, every cell resolved by the reference's
src/stabtd.js:24-32 under Node): a seeded sample of 384 cells plus the three cells where the
last bit of pow changes the answer (golden_util.POW_ULP_CELLS), identical alim, bracket and
open flag per cell except on those three.  The GPU test covers all 16,384 cells."""
import numpy as np

import golden_util as G


def test_oracle_matches_reference_on_grid_sample():
    import oracle
    from mcsim import workloads as W
    from mcsim._abi import cell_array
    ref = G.grid_limits()
    cells, opts = W.sld_grid()
    assert len(cells) == ref["n"] and opts == ref["opts"]
    idx = np.random.RandomState(2024).choice(len(cells), 384, replace=False)
    idx = np.concatenate([idx[~np.isin(idx, G.POW_ULP_CELLS)], G.POW_ULP_CELLS])
    got = oracle.bisect_batch(cell_array([cells[i] for i in idx]), len(idx), threads=8, **opts)
    ratio = (opts["aMax"] / opts["aMin"]) ** (0.5 ** opts["iters"])   # one bisection step
    for i, r in zip(idx, got):
        same = (r.alim, r.lo, r.hi, bool(r.open)) == (ref["alim"][i], ref["lo"][i], ref["hi"][i], ref["open"][i])
        assert same == (i not in G.POW_ULP_CELLS), i
    # the pow cells: 14077 ends one bisection step (a bracket) from the reference, 13924 two,
    # and 5395 flips at aMax
    r13924, r14077, r5395 = got[-2], got[-1], got[-3]
    for r, i, steps in ((r13924, 13924, 2), (r14077, 14077, 1)):
        assert abs(np.log(r.alim / ref["alim"][i])) <= steps * np.log(ratio) * 1.0001
    assert r5395.open and not ref["open"][5395]
