"""Transpose, accumulate/mask, and GrB_mxm shapes that use them, on the GPU
against the oracle (bit-exact for integer/bool, fp64 within 1e-12)."""
import ctypes

import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from graphblas_amd import workloads as W
from helpers import DevCsr, assert_csr_equal, download, power_law_csr, rand_csr

pytestmark = pytest.mark.gpu

TYPES = {"bool": (1, np.uint8), "int32": (6, np.int32), "int64": (8, np.int64),
         "fp64": (11, np.float64)}
OPS = {"none": 0, "plus": 1, "minus": 2, "times": 3, "min": 4, "max": 5, "first": 6,
       "second": 7, "any": 8, "pair": 9, "lor": 10, "land": 11}


def _vals(rng, n, tname):
    if tname == "bool":
        return rng.integers(0, 2, n).astype(np.uint8)
    if tname == "int32":
        return rng.integers(-2 ** 31, 2 ** 31 - 1, n, dtype=np.int64).astype(np.int32)
    if tname == "int64":
        return rng.integers(-1000, 1000, n).astype(np.int64)
    return rng.standard_normal(n)


def _with(m, x):
    return W.Csr(m.nrows, m.ncols, m.p, m.j, x)


@pytest.mark.parametrize("shape", ["rect", "powerlaw", "empty", "iso"])
@pytest.mark.parametrize("tname", ["int64", "fp64", "bool"])
def test_transpose(grb, shape, tname):
    print(f"[post] transpose {shape} {tname}", flush=True)
    code, dt = TYPES[tname]
    rng = np.random.default_rng(hash((shape, tname)) % 2 ** 32)
    if shape == "rect":
        A = rand_csr(130, 70, 0.08, rng)
    elif shape == "powerlaw":
        A = power_law_csr(500, 20, rng)
    elif shape == "empty":
        A = rand_csr(40, 90, 0.0, rng)
    else:
        A = rand_csr(60, 80, 0.1, rng)
    iso = shape == "iso"
    A = _with(A, _vals(rng, A.nnz, tname))
    if iso:
        A.x = np.full(A.nnz, A.x[0] if A.nnz else dt(1), dtype=dt)
    dA = DevCsr(A, code, iso=iso)
    out = capi.GmCsrOut()
    rc = capi.lib().gm_transpose(ctypes.byref(out), ctypes.byref(dA.view))
    dA.free()
    assert rc == 0, capi.last_error()
    assert out.iso == (1 if iso else 0)
    got = download(out, dt)
    ref = O.transpose(A)
    assert_csr_equal(got, ref, what="transpose")


def test_transpose_large(grb):
    """About 2.2M entries over 2^19 columns, two of them held by a third of
    the rows: the sort runs several radix passes over the column bits, and
    every column's rows must still come out ascending."""
    print("[post] transpose large", flush=True)
    rng = np.random.default_rng(77)
    m, n = 300000, 1 << 19
    keys = rng.integers(0, m * n, 2_000_000, dtype=np.int64)
    hub_rows = np.arange(0, m, 3, dtype=np.int64)
    keys = np.unique(np.concatenate([keys, hub_rows * n + 7, hub_rows * n + (n - 1)]))
    rows, cols = keys // n, keys % n
    p = np.zeros(m + 1, dtype=np.int64)
    np.add.at(p, rows + 1, 1)
    A = W.Csr(m, n, np.cumsum(p), cols, rng.standard_normal(len(keys)))
    dA = DevCsr(A, TYPES["fp64"][0])
    out = capi.GmCsrOut()
    rc = capi.lib().gm_transpose(ctypes.byref(out), ctypes.byref(dA.view))
    dA.free()
    assert rc == 0, capi.last_error()
    got = download(out, np.float64)
    assert_csr_equal(got, O.transpose(A), what="transpose large")


@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("tname", ["int32", "int64", "fp64", "bool"])
def test_accum_mask_ops(grb, op, tname):
    code, dt = TYPES[tname]
    rng = np.random.default_rng(hash((op, tname)) % 2 ** 32)
    C = rand_csr(80, 60, 0.1, rng)
    C = _with(C, _vals(rng, C.nnz, tname))
    T = rand_csr(80, 60, 0.1, rng)
    T = _with(T, _vals(rng, T.nnz, tname))
    M = rand_csr(80, 60, 0.2, rng)
    for comp, struct, replace in [(False, True, False), (True, True, False), (False, False, True),
                                  (True, False, True)]:
        Mv = _with(M, rng.integers(0, 2, M.nnz).astype(np.int64))
        ref = O.accum_mask(C, Mv, T, mask_comp=comp, mask_struct=struct, replace=replace,
                           op=OPS[op], dtype=dt)
        dC, dM, dT = DevCsr(C, code), DevCsr(Mv, 8), DevCsr(T, code)
        out = capi.GmCsrOut()
        rc = capi.lib().gm_accum_mask(ctypes.byref(out), ctypes.byref(dC.view), ctypes.byref(dM.view),
                                      comp, struct, replace, ctypes.byref(dT.view), OPS[op])
        for d in (dC, dM, dT):
            d.free()
        assert rc == 0, capi.last_error()
        got = download(out, dt)
        assert_csr_equal(got, ref, what=f"{op} comp={comp} struct={struct} replace={replace}")


def test_accum_mask_no_mask_no_c(grb):
    rng = np.random.default_rng(3)
    T = rand_csr(50, 40, 0.1, rng)
    T = _with(T, _vals(rng, T.nnz, "int64"))
    dT = DevCsr(T, 8)
    out = capi.GmCsrOut()
    rc = capi.lib().gm_accum_mask(ctypes.byref(out), None, None, False, True, False,
                                  ctypes.byref(dT.view), 1)
    dT.free()
    assert rc == 0, capi.last_error()
    assert_csr_equal(download(out, np.int64), T, what="C = T")


# --- GrB_mxm shapes that need a transpose or the accum/mask step ------------

def _mat(m, tname):
    return capi.Matrix.from_csr(m, tname)


def test_grb_mxm_transposed_operands(grb):
    rng = np.random.default_rng(11)
    A = _with(rand_csr(90, 70, 0.06, rng), None)
    A.x = rng.random(A.nnz)
    B = _with(rand_csr(90, 50, 0.06, rng), None)
    B.x = rng.random(B.nnz)
    Ad, Bd = _mat(A, "fp64"), _mat(B, "fp64")
    # C = A'*B (T0): 70 x 50
    C = capi.Matrix("fp64", 70, 50)
    capi.mxm_checked(C, None, "GrB_PLUS_TIMES_SEMIRING_FP64", Ad, Bd, "T0")
    ref = O.saxpy3("plus_times_fp64", O.transpose(A), B)
    assert_csr_equal(C.to_csr(), ref, fp_tol=1e-12, what="A'*B")
    # C = A*B' without a mask (T1): 90 x 90 with B2 90 x 70
    B2 = _with(rand_csr(90, 70, 0.06, rng), None)
    B2.x = rng.random(B2.nnz)
    C2 = capi.Matrix("fp64", 90, 90)
    capi.mxm_checked(C2, None, "GrB_PLUS_TIMES_SEMIRING_FP64", Ad, _mat(B2, "fp64"), "T1")
    ref = O.saxpy3("plus_times_fp64", A, O.transpose(B2))
    assert_csr_equal(C2.to_csr(), ref, fp_tol=1e-12, what="A*B'")
    # C<!M> = A*B' (RSCT1)
    M = rand_csr(90, 90, 0.1, rng)
    C3 = capi.Matrix("fp64", 90, 90)
    capi.mxm_checked(C3, _mat(M, "bool"), "GrB_PLUS_TIMES_SEMIRING_FP64", Ad, _mat(B2, "fp64"),
                     "RSCT1")
    ref = O.saxpy3("plus_times_fp64", A, O.transpose(B2), M=M, mask_comp=True)
    assert_csr_equal(C3.to_csr(), ref, fp_tol=1e-12, what="C<!M>=A*B'")
    # C<M> = A'*B' (ST0T1): transpose A, then dot3
    Ab = rand_csr(70, 90, 0.06, rng)
    Ab.x = rng.random(Ab.nnz)
    Bb = rand_csr(80, 70, 0.06, rng)
    Bb.x = rng.random(Bb.nnz)
    Mb = rand_csr(90, 80, 0.2, rng)
    C4 = capi.Matrix("fp64", 90, 80)
    capi.mxm_checked(C4, _mat(Mb, "bool"), "GrB_PLUS_TIMES_SEMIRING_FP64", _mat(Ab, "fp64"),
                     _mat(Bb, "fp64"), "ST0T1")
    ref = O.dot3("plus_times_fp64", Mb, O.transpose(Ab), Bb)
    assert_csr_equal(C4.to_csr(), ref, fp_tol=1e-12, what="C<M>=A'*B'")


@pytest.mark.parametrize("accum,desc", [("GrB_PLUS_INT64", None), ("GrB_MIN_INT64", "S"),
                                        (None, None), ("GrB_SECOND_INT64", "R"),
                                        ("GrB_PLUS_INT64", "SC")])
def test_grb_mxm_accum_and_mask_on_nonempty_c(grb, accum, desc):
    print(f"[post] GrB_mxm accum={accum} desc={desc}", flush=True)
    rng = np.random.default_rng(17)
    A = rand_csr(60, 50, 0.08, rng, "int64")
    B = rand_csr(50, 40, 0.08, rng, "int64")
    M = rand_csr(60, 40, 0.3, rng, "int64")
    C0 = rand_csr(60, 40, 0.2, rng, "int64")
    C = _mat(C0, "int64")
    Md = _mat(M, "int64")
    comp = desc is not None and "C" in desc
    struct = desc is not None and "S" in desc
    replace = desc is not None and "R" in desc
    capi.mxm_checked(C, Md, "GrB_PLUS_TIMES_SEMIRING_INT64", _mat(A, "int64"), _mat(B, "int64"),
                     desc, accum=accum)
    T = O.saxpy3("plus_times_int64", A, B, M=M, mask_comp=comp, mask_struct=struct)
    op = {None: 0, "GrB_PLUS_INT64": 1, "GrB_MIN_INT64": 4, "GrB_SECOND_INT64": 7}[accum]
    ref = O.accum_mask(C0, M, T, mask_comp=comp, mask_struct=struct, replace=replace, op=op,
                       dtype=np.int64)
    assert_csr_equal(C.to_csr(), ref, what=f"accum {accum} desc {desc}")


def test_grb_mxm_accum_without_mask(grb):
    rng = np.random.default_rng(23)
    A = rand_csr(40, 30, 0.1, rng, "fp64")
    B = rand_csr(30, 20, 0.1, rng, "fp64")
    C0 = rand_csr(40, 20, 0.2, rng, "fp64")
    C = _mat(C0, "fp64")
    capi.mxm_checked(C, None, "GrB_PLUS_TIMES_SEMIRING_FP64", _mat(A, "fp64"), _mat(B, "fp64"),
                     None, accum="GrB_PLUS_FP64")
    T = O.saxpy3("plus_times_fp64", A, B)
    ref = O.accum_mask(C0, None, T, op=1, dtype=np.float64)
    assert_csr_equal(C.to_csr(), ref, fp_tol=1e-12, what="C += A*B")


def test_grb_mxm_axb_dot_hint(grb):
    """C<M> = A*B with GxB_AxB_METHOD = GxB_AxB_DOT: B is transposed and dot3
    runs; the result equals the saxpy3 one."""
    L = capi.lib()
    rng = np.random.default_rng(29)
    A = rand_csr(70, 60, 0.08, rng, "int64")
    B = rand_csr(60, 50, 0.08, rng, "int64")
    M = rand_csr(70, 50, 0.2, rng, "int64")
    d = ctypes.c_void_p()
    assert L.GrB_Descriptor_new(ctypes.byref(d)) == 0
    assert L.GrB_Descriptor_set(d, 7090, 7083) == 0          # GxB_AxB_METHOD = GxB_AxB_DOT
    assert L.GrB_Descriptor_set(d, 1, 4) == 0                # structural mask
    C = capi.Matrix("int64", 70, 50)
    Md, Ad, Bd = _mat(M, "int64"), _mat(A, "int64"), _mat(B, "int64")   # keep the handles alive
    rc = L.GrB_mxm(C.h, Md.h, None, capi.handle("GrB_PLUS_TIMES_SEMIRING_INT64"), Ad.h, Bd.h, d)
    L.GrB_Descriptor_free(ctypes.byref(d))
    assert rc == 0, capi.last_error()
    ref = O.saxpy3("plus_times_int64", A, B, M=M)
    assert_csr_equal(C.to_csr(), ref, what="axb dot")
