"""The allocation-free reduction (gm_reduce_scalar_noalloc) against
gm_reduce_scalar through the C-ABI: every monoid and type, dense, iso and
empty arrays, two calls in a row with the second on a smaller input (the
fixed scratch keeps the first call's partials), and no growth of the device
arena while it runs."""
import ctypes

import numpy as np
import pytest

from graphblas_amd import capi

pytestmark = pytest.mark.gpu

# (add_op, type code, numpy dtype): PLUS/MIN/MAX/ANY on int32, int64, fp64;
# LOR (MAX), LAND (MIN), ANY on bool
CASES = [(op, tc, dt) for tc, dt in ((6, np.int32), (8, np.int64), (11, np.float64))
         for op in (1, 2, 3, 4)] + [(3, 1, np.uint8), (2, 1, np.uint8), (4, 1, np.uint8)]


def _values(rng, n, dt):
    if dt == np.float64:
        return rng.standard_normal(n)
    if dt == np.uint8:
        return (rng.random(n) < 0.9).astype(np.uint8)
    info = np.iinfo(dt)
    return rng.integers(info.min // 4, info.max // 4, n, dtype=dt)


def _reduce_both(L, x, op, tc, dt, iso, out_dev):
    n = len(x) if not iso else 5000
    d = L.gm_malloc(max(x.nbytes, 8))
    assert d
    try:
        if x.nbytes:
            assert L.gm_memcpy_h2d(d, x.ctypes.data, x.nbytes) == 0
        ref = np.zeros(1, dtype=dt)
        rc = L.gm_reduce_scalar(ref.ctypes.data, op, tc, d if len(x) else None, n, iso)
        assert rc == 0, capi.last_error()
        in_use = L.gm_arena_bytes_in_use()
        rc = L.gm_reduce_scalar_noalloc(out_dev, op, tc, d if len(x) else None, n, iso, None)
        assert rc == 0, capi.last_error()
        assert L.gm_synchronize() == 0
        assert L.gm_arena_bytes_in_use() == in_use, "the allocation-free call allocated"
        got = np.zeros(1, dtype=dt)
        assert L.gm_memcpy_d2h(got.ctypes.data, out_dev, got.nbytes) == 0
        return got[0], ref[0]
    finally:
        L.gm_free(d)


@pytest.mark.parametrize("op,tc,dt", CASES)
def test_noalloc_matches_reduce_scalar(grb, op, tc, dt):
    L = capi.lib()
    rng = np.random.default_rng(op * 100 + tc)
    out_dev = L.gm_malloc(8)
    assert out_dev
    try:
        # two calls in a row, the second on a smaller input
        for n in (3_000_017, 1_234):
            print(f"[noalloc] op={op} type={tc} n={n}", flush=True)
            x = _values(rng, n, dt)
            got, ref = _reduce_both(L, x, op, tc, dt, 0, out_dev)
            assert got.tobytes() == ref.tobytes(), (n, got, ref)
        # iso (x[0] n times) and empty
        got, ref = _reduce_both(L, _values(rng, 1, dt), op, tc, dt, 1, out_dev)
        assert got.tobytes() == ref.tobytes(), ("iso", got, ref)
        got, ref = _reduce_both(L, np.zeros(0, dtype=dt), op, tc, dt, 0, out_dev)
        assert got.tobytes() == ref.tobytes(), ("empty", got, ref)
    finally:
        L.gm_free(out_dev)


def test_noalloc_sum_equals_numpy(grb):
    # an independent check of the value itself: int64 sum with wrap-around
    L = capi.lib()
    rng = np.random.default_rng(7)
    out_dev = L.gm_malloc(8)
    try:
        for n in (1 << 22, 999):
            x = rng.integers(-(1 << 40), 1 << 40, n, dtype=np.int64)
            got, ref = _reduce_both(L, x, 1, 8, np.int64, 0, out_dev)
            assert int(got) == int(x.sum()) == int(ref)
    finally:
        L.gm_free(out_dev)


def test_noalloc_argument_errors(grb):
    L = capi.lib()
    assert L.gm_reduce_scalar_noalloc(None, 1, 8, None, 0, 0, None) == capi.GrB_INVALID_VALUE
    out_dev = L.gm_malloc(8)
    try:
        assert L.gm_reduce_scalar_noalloc(out_dev, 1, 8, None, 10, 0, None) == capi.GrB_INVALID_VALUE
        assert L.gm_reduce_scalar_noalloc(out_dev, 1, 2, out_dev, 1, 0, None) == -8  # int8: not supported
    finally:
        L.gm_free(out_dev)
