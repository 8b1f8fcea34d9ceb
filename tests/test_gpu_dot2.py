"""dot2 (every dot product of a small, dense C = A*B' or C<!M> = A*B') on the
GPU against the oracle: C-ABI and GrB_mxm (bit-exact for integer/bool,
plus_times fp64 within 1e-12 relative per entry)."""
import ctypes

import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from graphblas_amd import workloads as W
from helpers import SR, DevCsr, assert_csr_equal, download, rand_csr

pytestmark = pytest.mark.gpu

FP_TOL = 1e-12
ZT = {8: np.int64, 6: np.int32, 1: np.uint8, 11: np.float64}


def _vals(handle):
    return {"GrB_PLUS_TIMES_SEMIRING_FP64": "fp64", "GrB_MIN_PLUS_SEMIRING_INT32": "int32",
            "GrB_PLUS_TIMES_SEMIRING_INT64": "int64", "GrB_MIN_PLUS_SEMIRING_INT64": "int64",
            "GrB_MAX_PLUS_SEMIRING_INT32": "int32", "GrB_MIN_PLUS_SEMIRING_FP64": "fp64"}.get(handle)


def _gm_dot2(handle, A, B, M=None, mask_struct=True):
    name, trip = SR[handle]
    ab = trip[2]
    dA, dB = DevCsr(A, ab), DevCsr(B, ab)
    dM = DevCsr(M, 8) if M is not None else None
    out = capi.GmCsrOut()
    rc = capi.lib().gm_dot2(ctypes.byref(out), ctypes.byref(dM.view) if dM else None, M is not None,
                            mask_struct, ctypes.byref(dA.view), ctypes.byref(dB.view),
                            ctypes.byref(capi.GmSemiring(*trip)), None)
    for d in (dA, dB, dM):
        if d is not None:
            d.free()
    assert rc == 0, capi.last_error()
    return download(out, ZT[trip[2]])


def _pairs(A, B, M, struct):
    """Structural mask of every (i, j) with A(i,:) and B(j,:) nonempty, less
    the entries where M(i,j) is true."""
    ra = np.diff(A.p) > 0
    rb = np.flatnonzero(np.diff(B.p) > 0)
    p = [0]
    js = []
    for i in range(A.nrows):
        row = rb if ra[i] else rb[:0]
        if M is not None and ra[i]:
            mj = M.j[M.p[i]:M.p[i + 1]]
            if not struct:
                mj = mj[M.x[M.p[i]:M.p[i + 1]] != 0]
            row = np.setdiff1d(row, mj)
        js.append(row.astype(np.int64))
        p.append(p[-1] + len(row))
    return W.Csr(A.nrows, B.nrows, np.array(p, dtype=np.int64),
                 np.concatenate(js) if js else np.zeros(0, np.int64), None)


@pytest.mark.parametrize("handle", list(SR))
@pytest.mark.parametrize("masked", ["none", "comp_struct", "comp_valued"])
def test_dot2_cabi(grb, handle, masked):
    """C = A*B' and C<!M> = A*B' over every pair of nonempty rows, with empty
    rows on both sides."""
    print(f"[dot2] {handle} {masked}", flush=True)
    rng = np.random.default_rng(hash((handle, masked)) % 2 ** 32)
    vt = _vals(handle)
    A = rand_csr(40, 300, 0.15, rng, vt, empty_rows=0.2)
    B = rand_csr(35, 300, 0.15, rng, vt, empty_rows=0.2)
    M = None
    if masked != "none":
        M = rand_csr(40, 35, 0.3, rng, "int64")
        M = W.Csr(M.nrows, M.ncols, M.p, M.j, rng.integers(0, 2, M.nnz).astype(np.int64))
    struct = masked != "comp_valued"
    name, _ = SR[handle]
    # the oracle's dot3 over every allowed pair (dot semantics: an ANY monoid
    # keeps the first match, the smallest k)
    ref = O.dot3(name, _pairs(A, B, M, struct), A, B)
    got = _gm_dot2(handle, A, B, M=M, mask_struct=struct)
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what=f"dot2 {masked}")


def test_nonempty_rows(grb):
    rng = np.random.default_rng(5)
    for m, dens, er in [(1, 0.5, 0.0), (300, 0.02, 0.4), (50, 0.0, 0.0), (1000, 0.01, 0.9)]:
        A = rand_csr(m, 200, dens, rng, empty_rows=er)
        dA = DevCsr(A, 1)
        c = ctypes.c_int64(-1)
        rc = capi.lib().gm_nonempty_rows(ctypes.byref(c), ctypes.byref(dA.view))
        dA.free()
        assert rc == 0, capi.last_error()
        assert c.value == int((np.diff(A.p) > 0).sum())


@pytest.mark.parametrize("handle,tname", [("GxB_PLUS_PAIR_INT64", "int64"),
                                          ("GrB_PLUS_TIMES_SEMIRING_FP64", "fp64"),
                                          ("GrB_MIN_PLUS_SEMIRING_INT32", "int32")])
def test_grb_mxm_dot2_tiny_c(grb, handle, tname):
    """GrB_mxm C = A*B' (GrB_DESC_T1) with 8 x 10 nonempty rows: the dot2 rule
    (C tiny) takes it; C<!M> with GrB_DESC_SCT1 too; and C = A*B with the
    GxB_AxB_DOT hint transposes B first."""
    print(f"[dot2] GrB_mxm tiny C {handle}", flush=True)
    rng = np.random.default_rng(hash(handle) % 2 ** 32)
    vt = _vals(handle)
    A = rand_csr(8, 5000, 0.05, rng, vt)
    B = rand_csr(10, 5000, 0.05, rng, vt)
    name, trip = SR[handle]
    ctl = capi.lib().GB_AxB_dot2_control
    assert ctl(float(A.nnz), float(B.nnz), 8.0, 10.0, 5000.0)
    tol = FP_TOL if tname == "fp64" else None
    Ad, Bd = capi.Matrix.from_csr(A, tname), capi.Matrix.from_csr(B, tname)
    # C = A*B'
    C = capi.Matrix(tname, 8, 10)
    capi.mxm_checked(C, None, handle, Ad, Bd, "T1")
    ref = O.saxpy3(name, A, O.transpose(B))
    assert_csr_equal(C.to_csr(), ref, fp_tol=tol, what="C=A*B' dot2")
    # C<!M,struct> = A*B'
    M = rand_csr(8, 10, 0.4, rng)
    Md = capi.Matrix.from_csr(M, "bool")
    C2 = capi.Matrix(tname, 8, 10)
    capi.mxm_checked(C2, Md, handle, Ad, Bd, "SCT1")
    ref2 = O.saxpy3(name, A, O.transpose(B), M=M, mask_comp=True)
    assert_csr_equal(C2.to_csr(), ref2, fp_tol=tol, what="C<!M>=A*B' dot2")
    # C = A*Bt with GxB_AxB_DOT: Bt is transposed to rows of B', then dot2
    Bt = O.transpose(B)
    Btd = capi.Matrix.from_csr(Bt, tname)
    d = capi.Descriptor(axb=capi.GxB_AxB_DOT)
    C3 = capi.Matrix(tname, 8, 10)
    capi.mxm_checked(C3, None, handle, Ad, Btd, d)
    d.free()
    assert_csr_equal(C3.to_csr(), ref, fp_tol=tol, what="C=A*B GxB_AxB_DOT")
