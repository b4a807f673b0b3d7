"""The no-allocation stamp batch (mc_stamp_batch_noalloc) against mc_stamp_batch, through the
C ABI: two calls in a row on the same stocks, the second smaller than the first, each
compared map for map (bit-exact) and volume for volume (1e-12 relative: both sum the
per-stamp volume with FP64 atomics, in an order that varies).  A batch over the limit is
refused before anything is applied."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.int32), np.asarray(b).view(np.int32))


def test_noalloc_stamps_equal_the_allocating_entry_point():
    from mcsim import lib, path, workloads as W
    from mcsim._abi import MC_OK, Stamp
    L = lib.load()
    print("device:", lib.device_info(), flush=True)
    rec, _ = path.schedule(W.zigzag_segments(120), 0.1)
    arr = path.as_stamp_array(rec)
    n1, n2 = 40000, 9000
    assert len(arr) >= n1 + n2
    first = (Stamp * n1).from_buffer_copy(bytes(memoryview(arr))[:C.sizeof(Stamp) * n1])
    second = (Stamp * n2).from_buffer_copy(bytes(memoryview(arr))[C.sizeof(Stamp) * n1:C.sizeof(Stamp) * (n1 + n2)])
    for s in second:          # a second pass 1 mm deeper, so that it removes material again
        s.tipZ -= 1.0
    a, b = lib.Stock(51.1, 51.1, 20, 0.1), lib.Stock(51.1, 51.1, 20, 0.1)
    for call, batch in enumerate((first, second)):
        n = len(batch)
        va, vb = np.zeros(n), np.zeros(n)
        print("call %d: %d stamps" % (call, n), flush=True)
        assert L.mc_stamp_batch_noalloc(a.p, batch, n, va.ctypes.data) == MC_OK, L.mc_last_error()
        assert L.mc_stamp_batch(b.p, batch, n, vb.ctypes.data) == MC_OK, L.mc_last_error()
        assert same_bits(a.H(), b.H())
        for x, y in zip(a.side(), b.side()):
            assert same_bits(x, y)
        assert a.dirty() == b.dirty()
        assert (va > 0).sum() > n // 10
        np.testing.assert_allclose(va, vb, rtol=1e-12, atol=1e-12)
    # over the limit: refused, nothing applied
    H = a.H()
    big = (Stamp * 65537)()
    for i in range(len(big)):
        big[i] = first[i % n1]
    assert L.mc_stamp_batch_noalloc(a.p, big, len(big), None) == 2
    assert b"MC_NOALLOC_MAX_STAMPS" in L.mc_last_error()
    assert same_bits(a.H(), H)
    a.close()
    b.close()
