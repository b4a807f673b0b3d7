"""The learned bracket surrogate around the plain time-domain search.

Data come from the GPU chatter search itself: cells drawn from the lobe-grid
distributions (random tools and random speeds over 2000-30000 rpm, from a seed other
than the grid's) are resolved by mc_bisect_batch, and their log(alim) is the target.  The
MLP (24 -> 256 -> 256 -> 256 -> 1) is trained on the device (BF16 MFMA, Adam); its
prediction p gives the bracket [(1 - w) e^p, (1 + w) e^p], which mc_bisect_bracket_batch
verifies (the low end must be stable, the high end must chatter) before bisecting it down
to the tolerance; a failed bracket falls back to the full plain search.
"""
import math

import numpy as np

from . import lib, workloads as W


def sample_cells(n, seed=2, n_tools=None):
    """n cells: tools from grid_tools (seed), speeds uniform in [2000, 30000] rpm"""
    n_tools = n_tools or max(16, n // 16)
    tools = W.grid_tools(n_tools, seed)
    rng = W.ParkMiller(seed * 7919 + 13)
    cells = []
    for i in range(n):
        t = tools[i % n_tools]
        S = rng.uniform(2000.0, 30000.0)
        arc = W.arc_of(t["imm"] * t["D"], t["D"])
        cells.append(W.make_cell(t["tool"], t["mx"], t["my"], W.MAT_A5052, S, arc, 0.05, t["revs"], True))
    return cells


def make_dataset(cells, opts):
    """features and log(alim) of cells resolved by the plain GPU search"""
    res = lib.bisect_batch(cells, **opts)
    X = lib.surrogate_features(cells)
    y = np.array([math.log(r.alim) for r in res], dtype=np.float32)
    return X, y, res


def train(X, y, seed=1, **kw):
    m = lib.Mlp(seed)
    loss = m.train(X, y, **kw)
    return m, loss


def brackets(model, cells, width=0.15, aMin=0.05, aMax=10.0):
    """[(1 - w) e^p, (1 + w) e^p] clipped to [aMin, aMax]"""
    p = np.exp(model.predict(lib.surrogate_features(cells)).astype(np.float64))
    lo = np.maximum(aMin, (1 - width) * p)
    hi = np.minimum(aMax, (1 + width) * p)
    return lo, hi


def bracketed_search(model, cells, tol=0.05, width=0.15, **opts):
    lo, hi = brackets(model, cells, width, opts.get("aMin", 0.05), opts.get("aMax", 10.0))
    return lib.bracket_batch(cells, lo, hi, tol, **opts)


def residual_sigma(model, X, y):
    """spread of ln(alim) - prediction on held-out rows"""
    r = y.astype(np.float64) - model.predict(X).astype(np.float64)
    return float(np.sqrt(np.mean(r * r)))


def guided_search(model, cells, sigma, **opts):
    """timeDomainLimit with the surrogate as the speculation prior: identical results to
    the plain search, fewer speculative rounds when the prediction is good"""
    mu = model.predict(lib.surrogate_features(cells)).astype(np.float64)
    return lib.prior_batch(cells, mu, np.full(len(cells), sigma), **opts)
