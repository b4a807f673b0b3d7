"""The oracle's restated fdlibm against Node's own Math.* (V8 7.8), bit for bit.

sin / cos / tan / atan / hypot / Math.round / the double % are pinned exactly on random
and near-multiple-of-pi/2 inputs.  pow: the restated fdlibm e_pow differs from Node 12's
Math.pow by one ulp on ~4.5 % of random inputs (V8 7.8's pow algorithm is not the one
restated); the test pins that the disagreement never exceeds one ulp and that the
y = 2 special case (the reference's x ** 2) is exactly x * x."""
import os
import shutil
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
NODE = shutil.which("node")
needs_node = pytest.mark.skipif(NODE is None, reason="node not installed")


def node_math(fn, x, y=None):
    d = tempfile.mkdtemp()
    xf, yf, of = (os.path.join(d, n) for n in ("x.bin", "y.bin", "o.bin"))
    np.asarray(x, dtype="<f8").tofile(xf)
    if y is not None:
        np.asarray(y, dtype="<f8").tofile(yf)
    subprocess.run([NODE, os.path.join(HERE, "node_math.mjs"), fn, xf, yf if y is not None else "-", of],
                   check=True)
    return np.fromfile(of, dtype="<f8")


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.int64)


@needs_node
@pytest.mark.parametrize("fn,lo,hi", [("sin", -10, 10), ("cos", -10, 10), ("sin", 0, 2 * np.pi),
                                      ("cos", 0, 2 * np.pi), ("tan", -1.5, 1.5), ("atan", -20, 20)])
def test_trig_bit_exact(fn, lo, hi):
    rng = np.random.default_rng(hash(fn) % 1000)
    x = rng.uniform(lo, hi, 100000)
    k = rng.integers(1, 6, 20000)
    x = np.concatenate([x, k * (np.pi / 2) * (1 + rng.normal(0, 1e-10, 20000)), [0.0, -0.0, np.pi / 4]])
    assert np.array_equal(bits(oracle.math_batch(fn, x)), bits(node_math(fn, x)))


@needs_node
def test_hypot_bit_exact():
    rng = np.random.default_rng(4)
    x, y = rng.normal(0, 500, 100000), rng.normal(0, 500, 100000)
    y[:100] = 0.0
    assert np.array_equal(bits(oracle.math_batch("hypot", x, y)), bits(node_math("hypot", x, y)))


@needs_node
def test_pow_within_one_ulp_and_square_exact():
    rng = np.random.default_rng(5)
    x = rng.uniform(0.001, 30, 100000)
    y = rng.uniform(0.05, 0.3, 100000)
    a, b = bits(oracle.math_batch("pow", x, y)), bits(node_math("pow", x, y))
    assert np.max(np.abs(a - b)) <= 1
    assert np.mean(a != b) < 0.08
    z = rng.normal(0, 1e4, 10000)
    assert np.array_equal(bits(node_math("sq", z)), bits(z * z))
    assert np.array_equal(bits(oracle.math_batch("pow", z, np.full_like(z, 2.0))), bits(z * z))


@needs_node
def test_js_remainder_is_exact_fmod():
    rng = np.random.default_rng(6)
    x = rng.uniform(-30, 30, 50000)
    y = np.full_like(x, 2 * 3.141592653589793)
    assert np.array_equal(bits(np.fmod(x, y)), bits(node_math("fmod", x, y)))


@needs_node
def test_math_round():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(0, 200, 20000), np.arange(0, 100) + 0.5, [0.49999999999999994]])
    lib = oracle.lib()
    ours = np.array([lib.or_js_round(float(v)) for v in x])
    assert np.array_equal(ours, node_math("round", x))


def test_constants_match_their_bit_patterns():
    # the hex words fdlibm documents for its constants (independent of the decimal literals)
    want = {0.0: 0, 1.66666666666666019037e-01: 0x3FC555555555553E,
            6.93147182464599609375e-01: 0x3FE62E4300000000, 9.61796700954437255859e-01: 0x3FEEC709E0000000}
    for v, h in want.items():
        assert struct.unpack("<Q", struct.pack("<d", v))[0] == h
