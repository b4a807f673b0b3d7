"""The surrogate's CPU oracle (oracle/mlp.py) pinned without the device: BF16 rounding on
known values, and its exact (unquantised) gradient against central finite differences of
its own loss on parameters of every block."""
import numpy as np

import mlp as M  # oracle/mlp.py


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7                      # BF16 spacing at 1
    x = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 2 + 2 ** -20, -2.5, 0.0, 3.0e38])
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                   # tie -> even
    assert r[2] == one + 2 * ulp         # tie -> even (up)
    assert r[3] == one + ulp             # above the tie
    assert r[4] == -2.5 and r[5] == 0.0
    assert np.isfinite(r[6])


def _params(rng):
    P = rng.normal(0, 1, M.NPARAM)
    W1, b1, W2, b2, W3, b3, W4, b4 = M.unpack(P)   # views into P
    W1 *= np.sqrt(1.0 / M.IN)
    for W in (W2, W3):
        W *= np.sqrt(1.0 / M.HID)
    W4 *= np.sqrt(1.0 / M.HID)
    for b in (b1, b2, b3):
        b *= 0.1
    return P


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    P = _params(rng)
    X = rng.normal(0, 1, (16, M.IN))
    t = rng.normal(0, 1, 16)
    g, _ = M.grads(P, X, t, quant=False)
    sizes = [M.IN * M.HID, M.HID, M.HID * M.HID, M.HID, M.HID * M.HID, M.HID, M.HID, 1]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    h = 1e-6
    for blk in range(8):
        for i in rng.choice(sizes[blk], size=min(6, sizes[blk]), replace=False):
            j = offs[blk] + i
            Pp, Pm = P.copy(), P.copy()
            Pp[j] += h
            Pm[j] -= h
            lp = M.grads(Pp, X, t, quant=False)[1]
            lm = M.grads(Pm, X, t, quant=False)[1]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (blk, i, fd, g[j])


def test_quantised_gradient_close_to_exact():
    rng = np.random.default_rng(9)
    P = _params(rng)
    X = rng.normal(0, 1, (64, M.IN))
    t = rng.normal(0, 1, 64)
    gq, lq = M.grads(P, X, t, quant=True)
    ge, le = M.grads(P, X, t, quant=False)
    assert np.linalg.norm(gq - ge) <= 0.05 * np.linalg.norm(ge)
    assert abs(lq - le) <= 0.05 * le
