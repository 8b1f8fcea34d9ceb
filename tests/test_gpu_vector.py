"""GrB_vxm / GrB_mxv (vectors as 1 x n rows on GrB_mxm's kernels) on the GPU
against the oracle."""
import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from graphblas_amd import workloads as W
from helpers import assert_csr_equal, rand_csr

pytestmark = pytest.mark.gpu


def _row(idx, vals, n):
    """1 x n CSR of a sparse vector."""
    order = np.argsort(idx)
    return W.Csr(1, n, np.array([0, len(idx)], dtype=np.int64), np.asarray(idx, dtype=np.int64)[order],
                 None if vals is None else np.asarray(vals)[order])


def _got(w, n):
    I, X = w.to_arrays()
    return W.Csr(1, n, np.array([0, len(I)], dtype=np.int64), I, X)


def test_vxm_bfs_step(grb):
    """q'<!v'> = q'*A with ANY_SECONDI (the BFS step; GrB_DESC_RSC)."""
    G = W.rmat_sym(12, 16, seed=3)
    n = G.nrows
    Gd = capi.Matrix.from_csr(G, "bool")
    rng = np.random.default_rng(1)
    src = np.sort(rng.choice(n, 20, replace=False))
    q = capi.Vector.from_arrays(src, np.ones(20, dtype=np.int64), "int64", n)
    vis = np.union1d(src, rng.choice(n, 200, replace=False))
    v = capi.Vector.from_arrays(vis, np.ones(len(vis), dtype=bool), "bool", n)
    w = capi.Vector("int64", n)
    assert capi.vxm(w, v, "GxB_ANY_SECONDI_INT64", q, Gd, "RSC") == 0, capi.last_error()
    F = _row(src, np.ones(20, dtype=np.int64), n)
    V = _row(vis, None, n)
    ref = O.saxpy3("any_secondi_int64", F, G, M=V, mask_comp=True)
    assert_csr_equal(_got(w, n), ref, what="vxm bfs")


@pytest.mark.parametrize("desc", [None, "T0"])
def test_mxv_plus_times(grb, desc):
    rng = np.random.default_rng(7)
    A = rand_csr(120, 90, 0.05, rng, "fp64")
    Ad = capi.Matrix.from_csr(A, "fp64")
    ulen = 90 if desc is None else 120
    idx = np.sort(rng.choice(ulen, 30, replace=False))
    vals = rng.random(30)
    u = capi.Vector.from_arrays(idx, vals, "fp64", ulen)
    wlen = 120 if desc is None else 90
    w = capi.Vector("fp64", wlen)
    assert capi.mxv(w, None, "GrB_PLUS_TIMES_SEMIRING_FP64", Ad, u, desc) == 0, capi.last_error()
    U = _row(idx, vals, ulen)
    # w' = u' * A'  (A*u), or u' * A (A'*u)
    B = O.transpose(A) if desc is None else A
    ref = O.saxpy3("plus_times_fp64", U, B)
    assert_csr_equal(_got(w, wlen), ref, fp_tol=1e-12, what=f"mxv {desc}")
    # the dense product agrees too
    dense = A.to_scipy().toarray()
    uu = np.zeros(ulen)
    uu[idx] = vals
    dw = dense @ uu if desc is None else dense.T @ uu
    I, X = w.to_arrays()
    assert np.allclose(dw[I], X, rtol=1e-12, atol=0)


def test_vxm_transposed_with_accum_and_mask(grb):
    rng = np.random.default_rng(9)
    A = rand_csr(80, 60, 0.08, rng, "int64")
    Ad = capi.Matrix.from_csr(A, "int64")
    idx = np.sort(rng.choice(60, 25, replace=False))
    vals = rng.integers(-5, 6, 25)
    u = capi.Vector.from_arrays(idx, vals, "int64", 60)
    w0i = np.sort(rng.choice(80, 30, replace=False))
    w0x = rng.integers(-5, 6, 30)
    w = capi.Vector.from_arrays(w0i, w0x, "int64", 80)
    mi = np.sort(rng.choice(80, 40, replace=False))
    m = capi.Vector.from_arrays(mi, np.ones(40, dtype=bool), "bool", 80)
    # w'<m'> += u' * A'   (GrB_INP1 = TRAN)
    assert capi.vxm(w, m, "GrB_PLUS_TIMES_SEMIRING_INT64", u, Ad, "T1", accum="GrB_PLUS_INT64") == 0, \
        capi.last_error()
    U = _row(idx, vals.astype(np.int64), 60)
    M = _row(mi, None, 80)
    T = O.saxpy3("plus_times_int64", U, O.transpose(A), M=M)
    C0 = _row(w0i, w0x.astype(np.int64), 80)
    ref = O.accum_mask(C0, M, T, op=1, dtype=np.int64)
    assert_csr_equal(_got(w, 80), ref, what="w<m> += u'A'")
