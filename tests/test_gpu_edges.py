"""Edge shapes on the GPU against the oracle: empty dimensions, 1x1, a star
(one row with every column, exercising the longest-row paths), aliased
output, and results that are empty."""
import ctypes

import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from graphblas_amd import workloads as W
from helpers import DevCsr, assert_csr_equal, download, gm_dot3, gm_saxpy3, rand_csr

pytestmark = pytest.mark.gpu


def _empty(m, n):
    return W.Csr(m, n, np.zeros(m + 1, dtype=np.int64), np.zeros(0, dtype=np.int64),
                 np.zeros(0, dtype=np.float64))


@pytest.mark.parametrize("shape", [(0, 30, 50), (40, 0, 50), (40, 30, 0), (1, 1, 1)])
def test_saxpy3_degenerate_dimensions(grb, shape):
    m, k, n = shape
    rng = np.random.default_rng(m * 100 + k * 10 + n)
    A = rand_csr(m, k, 0.5, rng, "fp64") if m and k else _empty(m, k)
    B = rand_csr(k, n, 0.5, rng, "fp64") if k and n else _empty(k, n)
    ref = O.saxpy3("plus_times_fp64", A, B)
    got = gm_saxpy3("GrB_PLUS_TIMES_SEMIRING_FP64", A, B)
    assert got.nrows == m
    assert_csr_equal(got, ref, fp_tol=1e-12, what=str(shape))


def test_dot3_zero_rows(grb):
    rng = np.random.default_rng(3)
    M = _empty(0, 20)
    A = _empty(0, 15)
    B = rand_csr(20, 15, 0.3, rng, "int64")
    got = gm_dot3("GxB_PLUS_PAIR_INT64", M, A, B)
    assert got.nrows == 0 and got.nnz == 0


def test_star_graph_hub_rows(grb):
    """Vertex 0 adjacent to all others: one row of length n-1 (bitmap, workgroup
    bin, global hash table) next to rows of length 1."""
    n = 5000
    p = np.zeros(n + 1, dtype=np.int64)
    rows = [np.arange(1, n, dtype=np.int64)] + [np.array([0], dtype=np.int64)] * (n - 1)
    p[1:] = np.cumsum([len(r) for r in rows])
    S = W.Csr(n, n, p, np.concatenate(rows))
    # C<S> = S*S' (dot3): every entry's intersection is empty (no triangles)
    got = gm_dot3("GxB_PLUS_PAIR_INT64", S, S, S)
    assert got.nnz == 0
    # C = S*S (saxpy3): row 0 gets n-1 entries of value... its row is S(0,:)*S
    ref = O.saxpy3("plus_pair_int64", S, S)
    got = gm_saxpy3("GxB_PLUS_PAIR_INT64", S, S)
    assert_csr_equal(got, ref, what="star S*S")
    # C<!S> = S*S (the complemented form, as in BFS)
    ref = O.saxpy3("plus_pair_int64", S, S, M=S, mask_comp=True)
    got = gm_saxpy3("GxB_PLUS_PAIR_INT64", S, S, M=S, mask_comp=True)
    assert_csr_equal(got, ref, what="star C<!S>")


def test_grb_mxm_output_aliases_input(grb):
    """C = C*C with C aliased to both inputs (the reference allows aliasing)."""
    rng = np.random.default_rng(5)
    A = rand_csr(60, 60, 0.05, rng, "int64")
    C = capi.Matrix.from_csr(A, "int64")
    rc = capi.lib().GrB_mxm(C.h, None, None, capi.handle("GrB_PLUS_TIMES_SEMIRING_INT64"),
                            C.h, C.h, None)
    assert rc == 0, capi.last_error()
    ref = O.saxpy3("plus_times_int64", A, A)
    assert_csr_equal(C.to_csr(), ref, what="C = C*C")


def test_transpose_zero_rows_and_cols(grb):
    for (m, n) in [(0, 7), (7, 0), (1, 1)]:
        A = _empty(m, n) if m * n == 0 else W.Csr(1, 1, np.array([0, 1], dtype=np.int64),
                                                   np.array([0], dtype=np.int64), np.array([2.5]))
        dA = DevCsr(A, 11)
        out = capi.GmCsrOut()
        rc = capi.lib().gm_transpose(ctypes.byref(out), ctypes.byref(dA.view))
        dA.free()
        assert rc == 0, capi.last_error()
        got = download(out, np.float64)
        assert_csr_equal(got, O.transpose(A), what=f"transpose {m}x{n}")


def test_vector_empty_frontier(grb):
    G = W.rmat_sym(8, 8, seed=1)
    Gd = capi.Matrix.from_csr(G, "bool")
    q = capi.Vector("int64", G.nrows)
    w = capi.Vector("int64", G.nrows)
    assert capi.vxm(w, None, "GxB_ANY_SECONDI_INT64", q, Gd) == 0, capi.last_error()
    assert w.nvals == 0


@pytest.mark.parametrize("tname,iso", [("int64", False), ("fp64", False), ("bool", True),
                                       ("int32", False)])
def test_serialize_roundtrip(grb, tname, iso):
    L = capi.lib()
    rng = np.random.default_rng(4)
    A = rand_csr(50, 70, 0.1, rng)
    if tname == "bool":
        A.x = None
    else:
        A.x = {"int64": lambda: rng.integers(-9, 9, A.nnz), "fp64": lambda: rng.random(A.nnz),
               "int32": lambda: rng.integers(-9, 9, A.nnz).astype(np.int32)}[tname]()
    M = capi.Matrix.from_csr(A, tname)
    blob, size = ctypes.c_void_p(), ctypes.c_uint64()
    assert L.GxB_Matrix_serialize(ctypes.byref(blob), ctypes.byref(size), M.h, None) == 0
    bound = ctypes.c_uint64()
    assert L.GrB_Matrix_serializeSize(ctypes.byref(bound), M.h) == 0 and bound.value >= size.value
    out = ctypes.c_void_p()
    assert L.GxB_Matrix_deserialize(ctypes.byref(out), None, blob, size.value, None) == 0
    capi._libc.free(blob)
    B = capi.Matrix.__new__(capi.Matrix)
    B.h, B.type_name, B.nrows, B.ncols = out, tname, 50, 70
    got = B.to_csr()
    ref = M.to_csr()
    assert np.array_equal(got.p, ref.p) and np.array_equal(got.j, ref.j)
    assert np.array_equal(np.asarray(got.x), np.asarray(ref.x))
