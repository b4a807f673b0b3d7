"""The search kernel against the reference on the whole 256-speed x 64-tool lobe grid: every
one of the 16,384 cells resolved by the reference's timeDomainLimit (src/stabtd.js:24-32,
golden_tdl_grid.json) must come back with the identical alim, bracket and open flag, except
the three cells where V8's Math.pow and fdlibm's differ in the last bit (POW_ULP_CELLS),
which must carry the oracle's answer bit for bit."""
import numpy as np
import pytest

import golden_util as G

pytestmark = pytest.mark.gpu


def test_search_matches_reference_on_full_grid():
    from mcsim import lib, workloads as W
    from mcsim._abi import cell_array
    print("device:", lib.device_info(), flush=True)
    ref = G.grid_limits()
    cells, opts = W.sld_grid()
    assert len(cells) == ref["n"]
    print("full grid: %d cells" % len(cells), flush=True)
    res = lib.bisect_batch(cell_array(cells), **opts)
    alim = np.array([r.alim for r in res]); lo = np.array([r.lo for r in res])
    hi = np.array([r.hi for r in res]); op = np.array([bool(r.open) for r in res])
    bad = np.nonzero((alim != ref["alim"]) | (lo != ref["lo"]) | (hi != ref["hi"]) | (op != ref["open"]))[0]
    assert tuple(bad) == G.POW_ULP_CELLS, (len(bad), bad[:10].tolist())
    import oracle
    got = oracle.bisect_batch(cell_array([cells[i] for i in bad]), len(bad), threads=len(bad), **opts)
    for i, o in zip(bad, got):
        r = res[i]
        assert (r.alim, r.lo, r.hi, r.open, r.trials, r.steps, r.visits) == \
               (o.alim, o.lo, o.hi, o.open, o.trials, o.steps, o.visits), i
