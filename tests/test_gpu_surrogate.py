"""The learned bracket surrogate on the device: BF16 MFMA forward / backward against the
CPU oracle (oracle/mlp.py), hipGraph training (deterministic, loss falls), and the
surrogate-bracketed search against the plain search."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from mcsim import lib
    lib.load()
    print("device:", lib.device_info(), flush=True)
    return lib


def _model(lib, seed=3):
    m = lib.Mlp(seed)
    m.set_scaling(np.zeros(24), np.ones(24), 0.0, 1.0)
    return m


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


BLOCKS = [("W1", 24 * 256), ("b1", 256), ("W2", 65536), ("b2", 256), ("W3", 65536), ("b3", 256),
          ("W4", 256), ("b4", 1)]


def test_gradients_match_oracle(lib):
    import mlp as M
    m = _model(lib)
    rng = np.random.default_rng(1)
    for n in (256, 200):                         # a multiple of the tile and a ragged batch
        X = rng.normal(0, 1, (n, 24)).astype(np.float32)
        t = rng.normal(0, 1, n).astype(np.float32)
        g, loss = m.grads(X, t)
        P = m.params()
        gq, lq = M.grads(P, X.astype(np.float64), t.astype(np.float64), quant=True)
        ge, _ = M.grads(P, X.astype(np.float64), t.astype(np.float64), quant=False)
        o = 0
        for name, k in BLOCKS:
            eq = _rel(g[o:o + k], gq[o:o + k])
            ee = _rel(g[o:o + k], ge[o:o + k])
            print("n=%d %s rel.err vs BF16 oracle %.2e, vs FP64 %.2e" % (n, name, eq, ee), flush=True)
            assert eq < 5e-3, (n, name, eq)          # FP32 accumulation vs FP64
            assert ee < 5e-2, (n, name, ee)          # BF16 quantisation
            o += k
        assert abs(loss - lq) <= 1e-3 * lq


def test_forward_matches_oracle(lib):
    import mlp as M
    m = _model(lib, 4)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (1000, 24)).astype(np.float32)
    y = m.predict(X)
    yo, _ = M.forward(m.params(), X.astype(np.float64), quant=True)
    assert np.max(np.abs(y - yo)) <= 2e-3 * max(1.0, np.max(np.abs(yo)))


def test_training_is_deterministic_and_fits(lib):
    rng = np.random.default_rng(3)
    n = 8192
    X = rng.normal(0, 1, (n, 24)).astype(np.float32)
    t = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3] ** 2).astype(np.float32)
    runs = []
    for _ in range(2):
        m = lib.Mlp(7)
        loss = m.train(X, t, epochs=40, batch=1024, lr=2e-3, lr_final=1e-4, seed=5)
        runs.append((m.params(), loss))
        print("epoch loss first %.4f last %.5f" % (loss[0], loss[-1]), flush=True)
    assert np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][1][-1] < 0.05 * runs[0][1][0]


def test_eager_training_equals_graph_replay(lib, monkeypatch):
    """MC_MLP_EAGER launches the training step's kernels one by one instead of replaying the
    captured graph (for kernel tracers): the same kernels in the same order, so the weights
    and the losses are identical"""
    rng = np.random.default_rng(4)
    n = 4096
    X = rng.normal(0, 1, (n, 24)).astype(np.float32)
    t = (np.cos(X[:, 1]) - 0.2 * X[:, 0] * X[:, 5]).astype(np.float32)
    runs = []
    for eager in (False, True):
        if eager:
            monkeypatch.setenv("MC_MLP_EAGER", "1")
        else:
            monkeypatch.delenv("MC_MLP_EAGER", raising=False)
        m = lib.Mlp(9)
        loss = m.train(X, t, epochs=6, batch=512, lr=2e-3, lr_final=1e-4, seed=2)
        runs.append((m.params(), loss, m.predict(X[:1000])))
    monkeypatch.delenv("MC_MLP_EAGER", raising=False)
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][2], runs[1][2])


def test_surrogate_bracketed_search(lib):
    """train on cells the GPU search resolves, then bracket held-out cells (+-15 %): a
    verified bracket is no wider than the tolerance, a failed one returns the plain
    search's own result, and most verified brackets hold the plain search's limit (the
    verdict is not monotonic in depth for every cell, so a verified bracket can hold another
    stability transition than the one bisection from [aMin, aMax] finds; that share is the
    bracket form's error against plain bisection)"""
    from mcsim import surrogate as S
    opts = dict(aMin=0.05, aMax=10.0, iters=10)
    cells = S.sample_cells(3072, seed=11)
    train, test = cells[:2560], cells[2560:]
    X, y, _ = S.make_dataset(train, opts)
    m, loss = S.train(X, y, epochs=60, batch=512, lr=2e-3, lr_final=1e-4)
    plain = lib.bisect_batch(test, **opts)
    out, fb = S.bracketed_search(m, test, tol=0.05, width=0.15, **opts)
    hit = agree = 0
    tp = ts = 0
    for i in range(len(test)):
        p, s = plain[i], out[i]
        tp += p.trials
        ts += s.trials
        if fb[i]:
            assert (s.alim, s.lo, s.hi, s.open) == (p.alim, p.lo, p.hi, p.open), i
            continue
        hit += 1
        assert s.hi - s.lo <= 0.05 + 1e-12 and s.lo < s.hi
        step = p.hi - p.lo if not p.open else 0.0
        agree += s.lo - step - 1e-9 <= p.alim <= s.hi + step + 1e-9
    print("bracket hit rate %.3f, agreeing with plain %.3f of hits; trials plain %d surrogate %d"
          % (hit / len(test), agree / max(hit, 1), tp, ts), flush=True)
    assert hit > 0.3 * len(test)
    assert agree >= 0.8 * hit


def test_prior_guided_search_is_identical_to_plain(lib):
    """any prior (exact, noisy, wrong, absent) changes only which tree nodes are
    speculated, never the result: limits, brackets and serial-path counts equal the plain
    search's"""
    from mcsim import workloads as W
    cells, opts = W.sld_grid(n_speeds=16, n_tools=12, seed=5)
    plain = lib.bisect_batch(cells, **opts)
    n = len(cells)
    rng = np.random.default_rng(4)
    exact = np.array([math.log(r.alim) for r in plain])
    for mu, sig in ((exact, np.full(n, 0.05)), (exact + rng.normal(0, 0.3, n), np.full(n, 0.3)),
                    (np.full(n, math.log(9.5)), np.full(n, 0.01)), (exact, np.zeros(n))):
        got = lib.prior_batch(cells, mu, sig, **opts)
        for i in range(n):
            a, b = got[i], plain[i]
            assert (a.alim, a.lo, a.hi, a.open, a.trials, a.steps, a.visits, a.engaged) == \
                   (b.alim, b.lo, b.hi, b.open, b.trials, b.steps, b.visits, b.engaged), i


def test_prior_saves_speculative_trials(lib):
    """a good prior (the exact answer within its spread) makes the speculative search run
    fewer trials than the prior-free search over the same cells, while the serial path, and
    so every result, is the same; a confidently wrong prior costs trials instead"""
    from mcsim import workloads as W
    cells, opts = W.sld_grid(n_speeds=32, n_tools=16, seed=7)
    plain = lib.bisect_batch(cells, **opts)
    ex_plain, se_plain = lib.last_search_trials()
    n = len(cells)
    exact = np.array([math.log(r.alim) for r in plain])
    lib.prior_batch(cells, exact, np.full(n, 0.05), **opts)
    ex_good, se_good = lib.last_search_trials()
    lib.prior_batch(cells, np.full(n, math.log(9.5)), np.full(n, 0.01), **opts)
    ex_bad, se_bad = lib.last_search_trials()
    print("executed trials: plain %d, good prior %d, wrong prior %d (serial %d)" % (ex_plain, ex_good, ex_bad, se_plain))
    assert se_plain == se_good == se_bad == sum(r.trials for r in plain)
    assert ex_plain >= se_plain and ex_good >= se_good
    assert ex_good < ex_plain
    assert ex_bad > ex_good


def test_model_file_round_trip(tmp_path):
    """Mlp.save / Mlp.load (the file the JS Surrogate reads): weights, scaling and sigma come
    back, and the predictions are the same bits"""
    from mcsim import lib, workloads as W
    m = lib.Mlp(5)
    xm = np.linspace(-1, 1, lib.MC_SURR_IN).astype(np.float32)
    m.set_scaling(xm, np.full(lib.MC_SURR_IN, 2.0, np.float32), 0.25, 1.5)
    cells, _ = W.sld_grid(n_speeds=8, n_tools=4)
    X = lib.surrogate_features(cells)
    path = str(tmp_path / "m.bin")
    m.save(path, sigma=0.37)
    m2, sigma = lib.Mlp.load(path)
    assert abs(sigma - 0.37) < 1e-7
    assert np.array_equal(m2.params(), m.params())
    assert np.array_equal(m2.predict(X), m.predict(X))
