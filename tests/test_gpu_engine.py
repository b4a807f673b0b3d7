"""The device-resident CuttingEngine (mc_engine_*) against call sequences recorded from the
reference simulator (src/cutting.js:21-326 as src/sim.js:252-257,380-445,645-662 drives it):
state, stats, chip events and rings after every run, bit for bit with the linear force law;
with the size effect within 1e-12 of the reference and bit for bit with the oracle (both use
the restated fdlibm pow, which differs from V8's Math.pow in the last bit on a few % of
inputs).  Also: many engines stepped in one launch give each engine's own answers."""
import sys

import numpy as np
import pytest

import golden_util as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from mcsim import lib as L
    L.load()
    return L


@pytest.fixture(scope="module")
def englog():
    return G.load("englog")


def _linear(g):
    return [(c, r) for c, r in zip(g["cases"], g["results"]) if c.get("nonlinear") is False]


def test_engine_replays_simulator_calls_linear(lib, englog):
    for case, res in _linear(englog):
        print("englog linear", case["tool"], flush=True)
        bad, runs = G.replay_englog(lib.Engine(), res)
        assert not bad, (case["tool"], bad[:3])
        assert runs > 100


def test_engine_replays_simulator_calls_size_effect(lib, englog):
    import oracle
    for case, res in zip(englog["cases"], englog["results"]):
        if case.get("nonlinear") is False:
            continue
        print("englog size effect", case["tool"], flush=True)
        dev, ora = lib.Engine(), oracle.OracleEngine()
        bad, runs = G.replay_englog(dev, res, rtol=1e-12)
        assert not bad, (case["tool"], bad[:3])
        bad_o, runs_o = G.replay_englog(ora, res, rtol=1e-12)
        assert not bad_o and runs_o == runs
        sd, so = dev.state(), ora.state()
        for f, _ in sd._fields_:
            assert G._same(getattr(sd, f), getattr(so, f)), f
        a, b = dev.stats(), ora.stats()
        assert G._same([getattr(a, f) for f, _ in a._fields_], [getattr(b, f) for f, _ in b._fields_])
        for x, y in zip(dev.rings(), ora.rings()):
            assert np.array_equal(x, y)


def test_engine_batch_matches_single_runs(lib, englog):
    """the four linear-law programs stepped together: each batch launch runs every engine
    that has a run pending, and each engine still reproduces its own reference sequence"""
    cases = _linear(englog)
    engines = [lib.Engine() for _ in cases]
    bads = [[] for _ in cases]
    gens = [G.englog_replayer(e, r, bad=b) for e, (_, r), b in zip(engines, cases, bads)]
    pending = {}
    for i, g in enumerate(gens):
        pending[i] = next(g)
    launches = 0
    while pending:
        ids = sorted(pending)
        lib.engine_run_batch([engines[i] for i in ids], [pending[i] for i in ids])
        launches += 1
        for i in ids:
            try:
                pending[i] = gens[i].send(None)
            except StopIteration:
                del pending[i]
    assert launches > 100
    for (case, _), b in zip(cases, bads):
        assert not b, (case["tool"], b[:3])


def test_engine_argument_errors(lib):
    from mcsim._abi import make_cell
    e = lib.Engine()
    with pytest.raises(lib.McError):
        e.run(10)                       # run before configure
    cell = make_cell(dict(N=3, R=5.0, helixDeg=30.0, type="flat", fluteLen=20.0), [(2e4, 1500.0, 0.03)], None,
                     dict(Ktc=700, Krc=210, Kte=20, Kre=15), 8000.0, (0.0, 3.14), 0.05, 10)
    e.configure(cell)
    with pytest.raises(lib.McError):
        e.run(10)                       # configure without setEngagement
    e.set_engagement(1.0, 0.0, 0.05, -1.0, 1.0)
    with pytest.raises(lib.McError):
        e.run(-1)
    with pytest.raises(lib.McError):
        lib.engine_run_batch([e, e], [5, 5])   # one engine twice in a batch
    e.run(0)
    e.run(25)
    assert e.state().stepCount == 25
