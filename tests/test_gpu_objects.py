"""Object-model calls around the product on the device: extractTuples of
bool and int32 matrices, GrB_Matrix_reduce_BOOL, GrB_Vector_reduce_<T>
and GrB_Vector_dup, each against numpy on the same seeded inputs (the
reference's GrB_Matrix_extractTuples / GrB_reduce / GrB_Vector_dup,
Source/extractTuples/, Source/reduce/GB_reduce_to_scalar.c,
Source/vector/GrB_Vector_dup.c).  Integer and bool results are exact; the
fp64 sum is within 1e-12 of a compensated sum, relative to the sum of the
magnitudes (the bound of any summation order).
"""
import ctypes
import math

import numpy as np
import pytest

from graphblas_amd import capi
from helpers import rand_csr

pytestmark = pytest.mark.gpu

GrB_INSUFFICIENT_SPACE = -103


def _extract(A: capi.Matrix, fn: str, dt):
    L = capi.lib()
    nv = A.nvals
    I = np.empty(max(nv, 1), dtype=np.uint64)
    J = np.empty(max(nv, 1), dtype=np.uint64)
    X = np.empty(max(nv, 1), dtype=dt)
    n = ctypes.c_uint64(nv)
    rc = getattr(L, fn)(I.ctypes.data, J.ctypes.data, X.ctypes.data, ctypes.byref(n), A.h)
    assert rc == 0, capi.last_error()
    return I[:n.value], J[:n.value], X[:n.value]


@pytest.mark.parametrize("tname,fn,dt", [("bool", "GrB_Matrix_extractTuples_BOOL", np.bool_),
                                         ("int32", "GrB_Matrix_extractTuples_INT32", np.int32)])
def test_extract_tuples(grb, tname, fn, dt):
    rng = np.random.default_rng(5)
    M = rand_csr(200, 150, 0.05, rng, "int32")
    if tname == "bool":
        M.x = (M.x % 2).astype(np.uint8)
    A = capi.Matrix.from_csr(M, tname)
    I, J, X = _extract(A, fn, dt)
    rows = np.repeat(np.arange(M.nrows), np.diff(M.p))
    assert np.array_equal(I, rows) and np.array_equal(J, M.j)
    want = M.x.astype(bool) if tname == "bool" else M.x.astype(np.int32)
    assert np.array_equal(X.astype(want.dtype), want)
    # a count smaller than nvals is an error, not a short copy
    small = ctypes.c_uint64(max(A.nvals - 1, 0))
    buf = np.empty(max(A.nvals, 1), dtype=np.uint64)
    xb = np.empty(max(A.nvals, 1), dtype=dt)
    rc = getattr(capi.lib(), fn)(buf.ctypes.data, buf.ctypes.data, xb.ctypes.data,
                                 ctypes.byref(small), A.h)
    assert rc == GrB_INSUFFICIENT_SPACE


def test_matrix_reduce_bool(grb):
    L = capi.lib()
    rng = np.random.default_rng(6)
    for density, ones in ((0.05, 0.5), (0.05, 1.0), (0.05, 0.0), (0.0, 0.5)):
        M = rand_csr(300, 300, density, rng, "int32")
        M.x = (rng.random(M.nnz) < ones).astype(np.uint8)
        A = capi.Matrix.from_csr(M, "bool")
        for mon, want in (("GrB_LOR_MONOID_BOOL", bool(M.x.any())),
                          ("GrB_LAND_MONOID_BOOL", bool(M.x.all()))):
            c = ctypes.c_bool(not want)
            rc = L.GrB_Matrix_reduce_BOOL(ctypes.byref(c), None, ctypes.c_void_p(capi.handle(mon)),
                                          A.h, None)
            assert rc == 0, capi.last_error()
            assert c.value == want, (mon, density, ones)


@pytest.mark.parametrize("tname,fn,dt", [("int32", "GrB_Vector_reduce_INT32", ctypes.c_int32),
                                         ("fp64", "GrB_Vector_reduce_FP64", ctypes.c_double),
                                         ("bool", "GrB_Vector_reduce_BOOL", ctypes.c_bool)])
def test_vector_reduce(grb, tname, fn, dt):
    L = capi.lib()
    rng = np.random.default_rng(7)
    n = 100000
    idx = np.sort(rng.choice(n, 30000, replace=False))
    if tname == "int32":
        vals = rng.integers(-1000, 1000, len(idx)).astype(np.int32)
        cases = (("GrB_PLUS_MONOID_INT32", int(vals.astype(np.int64).sum())),
                 ("GrB_MIN_MONOID_INT32", int(vals.min())), ("GrB_MAX_MONOID_INT32", int(vals.max())))
    elif tname == "fp64":
        vals = rng.standard_normal(len(idx))
        cases = (("GrB_PLUS_MONOID_FP64", math.fsum(vals)),
                 ("GrB_MIN_MONOID_FP64", float(vals.min())), ("GrB_MAX_MONOID_FP64", float(vals.max())))
    else:
        vals = rng.random(len(idx)) < 0.5
        cases = (("GrB_LOR_MONOID_BOOL", bool(vals.any())), ("GrB_LAND_MONOID_BOOL", bool(vals.all())))
    u = capi.Vector.from_arrays(idx, vals, tname, n)
    for mon, want in cases:
        c = dt()
        rc = getattr(L, fn)(ctypes.byref(c), None, ctypes.c_void_p(capi.handle(mon)), u.h, None)
        assert rc == 0, capi.last_error()
        if tname == "fp64" and "PLUS" in mon:
            assert abs(c.value - want) <= 1e-12 * float(np.abs(vals).sum()), (c.value, want)
        else:
            assert c.value == want, (mon, c.value, want)


def test_vector_dup(grb):
    L = capi.lib()
    rng = np.random.default_rng(8)
    n = 5000
    idx = np.sort(rng.choice(n, 700, replace=False))
    vals = rng.integers(0, 100, len(idx)).astype(np.int64)
    u = capi.Vector.from_arrays(idx, vals, "int64", n)
    h = ctypes.c_void_p()
    assert L.GrB_Vector_dup(ctypes.byref(h), u.h) == 0, capi.last_error()
    w = capi.Vector.__new__(capi.Vector)
    w.h, w.n, w.type_name = h, n, "int64"
    wi, wx = w.to_arrays()
    assert np.array_equal(wi, idx) and np.array_equal(wx, vals)
    # the copy is its own object: clearing the source leaves it intact
    assert L.GrB_Vector_clear(u.h) == 0
    assert u.nvals == 0
    wi, wx = w.to_arrays()
    assert np.array_equal(wi, idx) and np.array_equal(wx, vals)
