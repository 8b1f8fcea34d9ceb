"""Bitmap and full matrices and the products with a bitmap or full result
(saxbit, saxpy4, saxpy5) on the GPU against the oracle.

The oracle computes every product in CSR (its saxpy3 restatement, which sums
each entry's terms for k ascending and keeps the last writer for ANY, like
the dense-result kernel), and C += A*B as accum(C, A*B) (its accum_mask
restatement).  Bar: bit-exact for bool/integer semirings and ANY; plus_times
fp64 within 1e-12 relative per entry (both sum in k order, so it is
bit-exact in practice).
"""
import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from graphblas_amd.workloads import Csr
from helpers import SR, assert_csr_equal, rand_csr

pytestmark = pytest.mark.gpu

FP_TOL = 1e-12
VT = {"GrB_PLUS_TIMES_SEMIRING_FP64": "fp64", "GrB_MIN_PLUS_SEMIRING_INT32": "int32",
      "GrB_PLUS_TIMES_SEMIRING_INT64": "int64", "GrB_MIN_PLUS_SEMIRING_INT64": "int64",
      "GrB_MAX_PLUS_SEMIRING_INT32": "int32", "GrB_MIN_PLUS_SEMIRING_FP64": "fp64"}
TN = {"fp64": "fp64", "int32": "int32", "int64": "int64", None: "bool"}
ZT = {"plus_pair_int64": "int64", "plus_pair_int32": "int32", "any_pair_bool": "bool",
      "plus_times_fp64": "fp64", "min_plus_int32": "int32", "any_secondi_int64": "int64",
      "plus_times_int64": "int64", "min_plus_int64": "int64", "max_plus_int32": "int32",
      "min_plus_fp64": "fp64"}


def _dense_of(A: Csr, dtype):
    """(values, present) of a host CSR."""
    x = np.zeros((A.nrows, A.ncols), dtype=dtype)
    b = np.zeros((A.nrows, A.ncols), dtype=bool)
    rows = np.repeat(np.arange(A.nrows), np.diff(A.p))
    b[rows, A.j] = True
    if A.x is not None:
        x[rows, A.j] = A.x
    else:
        x[rows, A.j] = 1
    return x, b


def _csr_of(x, b):
    rows, cols = np.nonzero(b)
    p = np.zeros(x.shape[0] + 1, dtype=np.int64)
    np.add.at(p, rows + 1, 1)
    p = np.cumsum(p)
    return Csr(x.shape[0], x.shape[1], p, cols.astype(np.int64), x[rows, cols].copy())


def _matrix(A: Csr, fmt: str, tname: str):
    """A GrB_Matrix holding A as 'sparse', 'bitmap' or 'full' (full: A is dense)."""
    if fmt == "sparse":
        return capi.Matrix.from_csr(A, tname)
    x, b = _dense_of(A, capi.NP_TYPES[tname])
    if fmt == "full":
        assert b.all()
        return capi.Matrix.from_dense(x, None, tname)
    return capi.Matrix.from_dense(x, b, tname)


def test_pack_unpack_and_conform(grb):
    rng = np.random.default_rng(5)
    A = rand_csr(70, 90, 0.2, rng, "fp64")
    x, b = _dense_of(A, np.float64)
    M = capi.Matrix.from_dense(x, b, "fp64")
    assert M.sparsity == capi.Matrix.BITMAP and M.nvals == A.nnz
    assert_csr_equal(M.to_csr(), A, what="bitmap -> CSR")
    gx, gb = M.to_dense()
    assert np.array_equal(gb, b) and np.array_equal(gx[b], x[b])
    # sparse only, then bitmap only, then full only (entries missing: stays bitmap)
    M.set_sparsity_control(capi.Matrix.SPARSE)
    assert M.sparsity == capi.Matrix.SPARSE
    assert_csr_equal(M.to_csr(), A, what="conform to sparse")
    M.set_sparsity_control(capi.Matrix.BITMAP)
    assert M.sparsity == capi.Matrix.BITMAP
    assert_csr_equal(M.to_csr(), A, what="conform to bitmap")
    M.set_sparsity_control(capi.Matrix.FULL)
    assert M.sparsity == capi.Matrix.BITMAP
    # a full matrix
    F = capi.Matrix.from_dense(x, None, "fp64")
    assert F.sparsity == capi.Matrix.FULL and F.nvals == 70 * 90
    fx, fb = F.to_dense()
    assert fb.all() and np.array_equal(fx, x)
    F.set_sparsity_control(capi.Matrix.SPARSE)
    assert F.sparsity == capi.Matrix.SPARSE and F.nvals == 70 * 90
    F.set_sparsity_control(capi.Matrix.FULL)
    assert F.sparsity == capi.Matrix.FULL
    assert np.array_equal(F.to_dense()[0], x)
    # reduce and dup of a bitmap matrix
    assert abs(capi.reduce_sum_fp64(M) - float(A.x.sum())) <= 1e-12 * max(1.0, abs(float(A.x.sum())))


CASES = [("sparse", "bitmap"), ("bitmap", "sparse"), ("full", "full"), ("bitmap", "bitmap"),
         ("sparse", "full")]


@pytest.mark.parametrize("handle", list(SR))
@pytest.mark.parametrize("fa,fb", CASES)
def test_saxbit_semirings(grb, handle, fa, fb):
    """C = A*B with a bitmap or full operand: C bitmap (saxbit)."""
    print(f"[dense] saxbit {handle} A {fa} B {fb}", flush=True)
    rng = np.random.default_rng(hash((handle, fa, fb)) % (2 ** 32))
    vt = VT.get(handle)
    name, _ = SR[handle]
    tn = TN[vt] if vt else ("bool" if "pair" in name or "secondi" in name else "int64")
    m, k, n = 120, 100, 2500          # two column tiles of C
    A = rand_csr(m, k, 1.0 if fa == "full" else 0.3, rng, vt)
    B = rand_csr(k, n, 1.0 if fb == "full" else (0.02 if fb == "sparse" else 0.3), rng, vt)
    Ad, Bd = _matrix(A, fa, tn), _matrix(B, fb, tn)
    C = capi.Matrix(ZT[name], m, n)
    capi.mxm_checked(C, None, handle, Ad, Bd)
    assert C.sparsity == capi.Matrix.BITMAP, "saxbit gives a bitmap C"
    ref = O.saxpy3(name, A, B)
    assert_csr_equal(C.to_csr(), ref, fp_tol=FP_TOL if "fp64" in name else None, what=handle)


@pytest.mark.parametrize("mask", ["bitmap", "bitmap_comp", "sparse_comp", "full_valued"])
@pytest.mark.parametrize("handle", ["GrB_PLUS_TIMES_SEMIRING_FP64", "GxB_PLUS_PAIR_INT64",
                                    "GxB_ANY_SECONDI_INT64", "GrB_MIN_PLUS_SEMIRING_INT32"])
def test_saxbit_masks(grb, handle, mask):
    print(f"[dense] saxbit mask {mask} {handle}", flush=True)
    rng = np.random.default_rng(hash((handle, mask)) % (2 ** 32))
    vt = VT.get(handle)
    name, _ = SR[handle]
    tn = TN[vt] if vt else "bool"
    m, k, n = 90, 60, 300
    A = rand_csr(m, k, 0.3, rng, vt)
    B = rand_csr(k, n, 0.3, rng, vt)
    Mk = rand_csr(m, n, 0.4, rng, "int32")
    comp = mask.endswith("comp")
    struct = mask in ("bitmap", "bitmap_comp", "sparse_comp")
    if mask == "full_valued":
        # a full mask whose zero values exclude their entries
        xm = rng.integers(0, 2, (m, n)).astype(np.int32)
        Md = capi.Matrix.from_dense(xm, None, "int32")
        Mref = _csr_of(xm, np.ones((m, n), dtype=bool))
    else:
        Mref = Mk
        Md = _matrix(Mk, "sparse" if mask.startswith("sparse") else "bitmap", "int32")
    Ad, Bd = capi.Matrix.from_csr(A, tn), _matrix(B, "bitmap", tn)
    C = capi.Matrix(ZT[name], m, n)
    desc = {(True, True): "RSC", (True, False): "RC", (False, True): "S", (False, False): None}[(comp, struct)]
    capi.mxm_checked(C, Md, handle, Ad, Bd, desc)
    ref = O.saxpy3(name, A, B, M=Mref, mask_comp=comp, mask_struct=struct)
    assert_csr_equal(C.to_csr(), ref, fp_tol=FP_TOL if "fp64" in name else None, what=mask)


@pytest.mark.parametrize("which", ["saxpy4", "saxpy5"])
@pytest.mark.parametrize("handle,accum", [("GrB_PLUS_TIMES_SEMIRING_FP64", "GrB_PLUS_FP64"),
                                          ("GrB_MIN_PLUS_SEMIRING_INT32", "GrB_MIN_INT32"),
                                          ("GrB_PLUS_TIMES_SEMIRING_INT64", "GrB_PLUS_INT64")])
def test_saxpy45_full_accumulate(grb, which, handle, accum):
    """C += A*B with C full and the monoid as accum: saxpy4 (A sparse, B full)
    and saxpy5 (A full, B sparse); C stays full and equals accum(C, A*B)."""
    print(f"[dense] {which} {handle}", flush=True)
    rng = np.random.default_rng(hash((which, handle)) % (2 ** 32))
    vt = VT[handle]
    name, _ = SR[handle]
    dt = capi.NP_TYPES[vt]
    m, k, n = 80, 70, 2200
    if which == "saxpy4":
        A = rand_csr(m, k, 0.1, rng, vt)
        B = rand_csr(k, n, 1.0, rng, vt)
        Ad, Bd = capi.Matrix.from_csr(A, vt), _matrix(B, "full", vt)
    else:
        A = rand_csr(m, k, 1.0, rng, vt)
        B = rand_csr(k, n, 0.05, rng, vt)
        Ad, Bd = _matrix(A, "full", vt), capi.Matrix.from_csr(B, vt)
    cx = (rng.random((m, n)) if vt == "fp64" else rng.integers(-50, 50, (m, n))).astype(dt)
    C = capi.Matrix.from_dense(cx, None, vt)
    capi.mxm_checked(C, None, handle, Ad, Bd, None, accum)
    assert C.sparsity == capi.Matrix.FULL, "C += A*B keeps C full"
    T = O.saxpy3(name, A, B)
    Cin = _csr_of(cx, np.ones((m, n), dtype=bool))
    op = {"GrB_PLUS_FP64": 1, "GrB_PLUS_INT64": 1, "GrB_MIN_INT32": 4}[accum]
    ref = O.accum_mask(Cin, None, T, op=op, dtype=dt)
    assert_csr_equal(C.to_csr(), ref, fp_tol=FP_TOL if vt == "fp64" else None, what=which)


def test_dense_operands_through_sparse_paths(grb):
    """A transposed bitmap operand and a sparse-enough case take the CSR form
    of the dense matrix (dot3 / saxpy3) and give the same result."""
    rng = np.random.default_rng(9)
    A = rand_csr(150, 120, 0.1, rng, "fp64")
    B = rand_csr(150, 200, 0.1, rng, "fp64")
    Ad = _matrix(A, "bitmap", "fp64")
    Bd = _matrix(B, "bitmap", "fp64")
    C = capi.Matrix("fp64", 120, 200)
    capi.mxm_checked(C, None, "GrB_PLUS_TIMES_SEMIRING_FP64", Ad, Bd, "T0")   # C = A'*B
    ref = O.saxpy3("plus_times_fp64", O.transpose(A), B)
    assert_csr_equal(C.to_csr(), ref, fp_tol=FP_TOL, what="A' * B bitmap")
    # {S} * {B} with few nonempty rows of A: C sparse (saxpy3 on B's CSR form)
    A2 = rand_csr(400, 150, 0.1, rng, "fp64", empty_rows=0.9)
    C2 = capi.Matrix("fp64", 400, 200)
    capi.mxm_checked(C2, None, "GrB_PLUS_TIMES_SEMIRING_FP64", capi.Matrix.from_csr(A2, "fp64"), Bd)
    assert C2.sparsity == capi.Matrix.SPARSE
    assert_csr_equal(C2.to_csr(), O.saxpy3("plus_times_fp64", A2, B), fp_tol=FP_TOL, what="S*B")
    # C's sparsity control applies to the result: bitmap result held sparse
    C3 = capi.Matrix("fp64", 150, 200)
    C3.set_sparsity_control(capi.Matrix.SPARSE)
    Ad2 = _matrix(rand_csr(150, 150, 0.3, rng, "fp64"), "bitmap", "fp64")
    capi.mxm_checked(C3, None, "GrB_PLUS_TIMES_SEMIRING_FP64", Ad2, Bd)
    assert C3.sparsity == capi.Matrix.SPARSE


def test_saxbit_cabi_errors(grb):
    L = capi.lib()
    out = capi.GmDenseOut()
    sr = capi.GmSemiring(1, 1, 11)
    # both forms of A given: invalid
    a = capi.GmCsr()
    d = capi.GmDense()
    import ctypes
    rc = L.gm_saxbit(ctypes.byref(out), None, None, False, True, ctypes.byref(a), ctypes.byref(d),
                     ctypes.byref(a), None, ctypes.byref(sr), None)
    assert rc == capi.GrB_INVALID_VALUE
