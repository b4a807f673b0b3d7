"""The seeded full-size coupled program (config 4: a zigzag pocket, then chorded contour
passes): its move count, its bounds on the stock, its depth order and its determinism."""
import numpy as np

from mcsim import workloads as W


def test_full_program_shape():
    width = 102.35
    x, y, z, r, ln, F, S = W.coupled_full_program(3, width, 20000)
    assert len(x) == 20001 and len(ln) == len(F) == len(S) == len(r) == len(x)
    assert ln == list(range(1, 20002))
    cut = np.array(r) == 0
    xs, ys, zs = np.array(x), np.array(y), np.array(z)
    # every cutting move stays on the stock; a rapid never goes below the floor cut before it
    # (each contour loop is entered by a rapid down to the depth of the loop before it)
    assert np.all((xs[cut] >= 0) & (xs[cut] <= width) & (ys[cut] >= 0) & (ys[cut] <= width))
    for i in np.nonzero(~cut)[0]:
        if i > 4:
            assert zs[i] >= zs[:i][cut[:i]].min() - 1e-12, i
    # contour chords are short (the discretizer's chords), the pocket's passes long
    d = np.hypot(np.diff(xs), np.diff(ys))[cut[1:]]
    assert np.median(d) <= 0.0901 and d.max() > 50
    # each contour loop is deeper than the one before: the cutting depths never rise past the
    # pocket floor once the loops start
    floor = zs[4]
    assert np.all(zs[cut] <= floor + 1e-12)
    assert W.coupled_full_program(3, width, 20000) == (x, y, z, r, ln, F, S)
    assert W.coupled_full_program(4, width, 20000) != (x, y, z, r, ln, F, S)
