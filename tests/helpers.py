"""Shared helpers for the tests: random CSR inputs, C-ABI calls, comparisons."""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from graphblas_amd import capi
from graphblas_amd.workloads import Csr

# semiring handle -> (oracle name, kernel triple)
SR = {
    "GxB_PLUS_PAIR_INT64": ("plus_pair_int64", (1, 3, 8)),
    "GxB_PLUS_PAIR_INT32": ("plus_pair_int32", (1, 3, 6)),
    "GxB_ANY_PAIR_BOOL": ("any_pair_bool", (4, 3, 1)),
    "GrB_PLUS_TIMES_SEMIRING_FP64": ("plus_times_fp64", (1, 1, 11)),
    "GrB_MIN_PLUS_SEMIRING_INT32": ("min_plus_int32", (2, 2, 6)),
    "GxB_ANY_SECONDI_INT64": ("any_secondi_int64", (4, 4, 8)),
    "GrB_PLUS_TIMES_SEMIRING_INT64": ("plus_times_int64", (1, 1, 8)),
    "GrB_MIN_PLUS_SEMIRING_INT64": ("min_plus_int64", (2, 2, 8)),
    "GrB_MAX_PLUS_SEMIRING_INT32": ("max_plus_int32", (3, 2, 6)),
    "GrB_MIN_PLUS_SEMIRING_FP64": ("min_plus_fp64", (2, 2, 11)),
}

ZT = {"int64": np.int64, "int32": np.int32, "bool": np.uint8, "fp64": np.float64}


def rand_csr(nrows, ncols, density, rng, values=None, empty_rows=0.0) -> Csr:
    """Random CSR with sorted rows.  values: None, 'fp64', 'int32' (1..100), 'int64'."""
    m = sp.random(nrows, ncols, density=density, format="csr", random_state=rng,
                  data_rvs=lambda k: np.ones(k))
    if empty_rows > 0 and nrows > 0:
        keep = rng.random(nrows) >= empty_rows
        m = sp.diags(keep.astype(np.float64)) @ m
        m = m.tocsr()
        m.eliminate_zeros()
    m.sort_indices()
    p = m.indptr.astype(np.int64)
    j = m.indices.astype(np.int64)
    nnz = len(j)
    x = None
    if values == "fp64":
        x = rng.random(nnz)
    elif values == "int32":
        x = rng.integers(1, 101, nnz).astype(np.int32)
    elif values == "int64":
        x = rng.integers(-50, 51, nnz).astype(np.int64)
    return Csr(nrows, ncols, p, j, x)


def power_law_csr(n, avg, rng, alpha=1.3, values=None) -> Csr:
    """Rows with heavy-tailed lengths (some rows dense)."""
    lens = np.minimum((rng.pareto(alpha, n) + 1) * avg / 3, n).astype(np.int64)
    p = np.zeros(n + 1, dtype=np.int64)
    p[1:] = np.cumsum(lens)
    j = np.empty(p[-1], dtype=np.int64)
    for i in range(n):
        j[p[i]:p[i + 1]] = np.sort(rng.choice(n, size=lens[i], replace=False))
    x = None
    if values == "fp64":
        x = rng.random(len(j))
    elif values == "int32":
        x = rng.integers(1, 101, len(j)).astype(np.int32)
    elif values == "int64":
        x = rng.integers(-50, 51, len(j)).astype(np.int64)
    return Csr(n, n, p, j, x)


def assert_csr_equal(got: Csr, ref: Csr, fp_tol=None, what=""):
    assert got.nrows == ref.nrows, what
    assert np.array_equal(got.p, ref.p), f"{what}: row pointers differ"
    assert np.array_equal(got.j, ref.j), f"{what}: column indices differ"
    if ref.x is None:
        return
    gx = np.asarray(got.x).astype(ref.x.dtype) if got.x is not None else None
    if fp_tol is None:
        assert np.array_equal(gx, ref.x), f"{what}: values differ"
    else:
        den = np.maximum(np.abs(ref.x), np.finfo(np.float64).tiny)
        err = np.max(np.abs(gx - ref.x) / den) if len(ref.x) else 0.0
        assert err <= fp_tol, f"{what}: max rel err {err}"


# ---------------------------------------------------------------------------
# direct C-ABI (gm_*) calls on device copies of host CSR arrays
# ---------------------------------------------------------------------------

DevCsr = capi.DeviceCsr


def download(out: capi.GmCsrOut, zdtype) -> Csr:
    L = capi.lib()
    n, nz = out.nrows, out.nvals
    p = np.empty(n + 1, dtype=np.int64)
    j = np.empty(max(nz, 1), dtype=np.int64)
    assert L.gm_memcpy_d2h(p.ctypes.data, out.p, p.nbytes) == 0
    if nz:
        assert L.gm_memcpy_d2h(j.ctypes.data, out.j, 8 * nz) == 0
    if out.iso:
        x1 = np.empty(1, dtype=zdtype)
        assert L.gm_memcpy_d2h(x1.ctypes.data, out.x, x1.nbytes) == 0
        x = np.full(nz, x1[0], dtype=zdtype)
    else:
        x = np.empty(max(nz, 1), dtype=zdtype)
        if nz:
            assert L.gm_memcpy_d2h(x.ctypes.data, out.x, x.itemsize * nz) == 0
        x = x[:nz]
    for q in (out.p, out.j, out.x):
        L.gm_free(q)
    return Csr(n, out.ncols, p, j[:nz], x)


def gm_dot3(handle: str, M: Csr, A: Csr, B: Csr, mtype=8, abtype=None, mask_struct=True,
            params=None) -> Csr:
    _, trip = SR[handle]
    ztype = trip[2]
    abtype = abtype or ztype
    dM, dA, dB = DevCsr(M, mtype), DevCsr(A, abtype), DevCsr(B, abtype)
    out = capi.GmCsrOut()
    s = capi.GmSemiring(*trip)
    if params and len(params) == 3:
        params = tuple(params) + (0,)
    prm = capi.GmDot3Params(*params) if params else None
    rc = capi.lib().gm_dot3(ctypes.byref(out), ctypes.byref(dM.view), mask_struct,
                            ctypes.byref(dA.view), ctypes.byref(dB.view), ctypes.byref(s),
                            ctypes.byref(prm) if prm else None)
    for d in (dM, dA, dB):
        d.free()
    assert rc == 0, f"gm_dot3 rc={rc}: {capi.last_error()}"
    zname = {8: "int64", 6: "int32", 1: "bool", 11: "fp64"}[ztype]
    return download(out, ZT[zname])


def gm_saxpy3(handle: str, A: Csr, B: Csr, M: Csr = None, mask_comp=False, mask_struct=True,
              mtype=8, abtype=None, params=None) -> Csr:
    _, trip = SR[handle]
    ztype = trip[2]
    abtype = abtype or ztype
    dA, dB = DevCsr(A, abtype), DevCsr(B, abtype)
    dM = DevCsr(M, mtype) if M is not None else None
    out = capi.GmCsrOut()
    s = capi.GmSemiring(*trip)
    if params and len(params) == 3:
        params = tuple(params) + (0,)
    prm = capi.GmSaxpy3Params(*params) if params else None
    rc = capi.lib().gm_saxpy3(ctypes.byref(out), ctypes.byref(dM.view) if dM else None,
                              mask_comp, mask_struct, ctypes.byref(dA.view),
                              ctypes.byref(dB.view), ctypes.byref(s),
                              ctypes.byref(prm) if prm else None)
    for d in (dA, dB) + ((dM,) if dM else ()):
        d.free()
    assert rc == 0, f"gm_saxpy3 rc={rc}: {capi.last_error()}"
    zname = {8: "int64", 6: "int32", 1: "bool", 11: "fp64"}[ztype]
    return download(out, ZT[zname])


# ---------------------------------------------------------------------------
# batched calls
# ---------------------------------------------------------------------------

GmDot3Request = capi.GmDot3Request
GmSaxpy3Request = capi.GmSaxpy3Request


def _bind_batch():
    return capi.lib()


def slice_stacked(C: Csr, r0, r1, col_sub=0) -> Csr:
    p = C.p[r0:r1 + 1] - C.p[r0]
    j = C.j[C.p[r0]:C.p[r1]] - col_sub
    x = None if C.x is None else C.x[C.p[r0]:C.p[r1]]
    return Csr(r1 - r0, 0, p.copy(), j.copy(), None if x is None else x.copy())


def gm_batch_dot3(handle, reqs, abtype=None, mtype=8, mask_struct=True, shared=False, params=None):
    """reqs: list of (M, A, B) host Csr; shared: every request has M = A = B
    (uploaded once).  Returns (stacked Csr, row_base, col_base)."""
    L = _bind_batch()
    _, trip = SR[handle]
    abtype = abtype or trip[2]
    keep = []
    arr = (GmDot3Request * max(len(reqs), 1))()
    for r, (M, A, B) in enumerate(reqs):
        if shared:
            dA = DevCsr(A, abtype)
            keep.append(dA)
            v = dA.view
            arr[r] = GmDot3Request(v, v, v)
        else:
            dM, dA, dB = DevCsr(M, mtype), DevCsr(A, abtype), DevCsr(B, abtype)
            keep += [dM, dA, dB]
            arr[r] = GmDot3Request(dM.view, dA.view, dB.view)
    rb = np.zeros(len(reqs) + 1, dtype=np.int64)
    cb = np.zeros(len(reqs) + 1, dtype=np.int64)
    out = capi.GmCsrOut()
    s = capi.GmSemiring(*trip)
    prm = capi.GmDot3Params(*params) if params else None
    rc = L.gm_batch_dot3(ctypes.byref(out), rb.ctypes.data, cb.ctypes.data, arr, len(reqs),
                         mask_struct, ctypes.byref(s), ctypes.byref(prm) if prm else None)
    assert rc == 0, f"gm_batch_dot3 rc={rc}: {capi.last_error()}"
    zname = {8: "int64", 6: "int32", 1: "bool", 11: "fp64"}[trip[2]]
    # split on the device as well, and compare with the host slices
    parts = (capi.GmCsrOut * max(len(reqs), 1))()
    rc = L.gm_batch_split(parts, ctypes.byref(out), rb.ctypes.data, cb.ctypes.data, len(reqs))
    assert rc == 0, f"gm_batch_split rc={rc}: {capi.last_error()}"
    split = [download(parts[r], ZT[zname]) for r in range(len(reqs))]
    C = download(out, ZT[zname])
    for d in keep:
        d.free()
    return C, rb, cb, split


def gm_batch_saxpy3(handle, reqs, B: Csr, mask_comp=False, mask_struct=True, abtype=None, mtype=8,
                    params=None):
    """reqs: list of (M or None, A) host Csr against the shared B."""
    L = _bind_batch()
    _, trip = SR[handle]
    abtype = abtype or trip[2]
    keep = []
    dB = DevCsr(B, abtype)
    keep.append(dB)
    arr = (GmSaxpy3Request * max(len(reqs), 1))()
    for r, (M, A) in enumerate(reqs):
        dA = DevCsr(A, abtype)
        keep.append(dA)
        if M is not None:
            dM = DevCsr(M, mtype)
            keep.append(dM)
            arr[r] = GmSaxpy3Request(dM.view, dA.view)
        else:
            arr[r] = GmSaxpy3Request(capi.GmCsr(), dA.view)
    rb = np.zeros(len(reqs) + 1, dtype=np.int64)
    out = capi.GmCsrOut()
    s = capi.GmSemiring(*trip)
    if params and len(params) == 3:
        params = tuple(params) + (0,)
    prm = capi.GmSaxpy3Params(*params) if params else None
    rc = L.gm_batch_saxpy3(ctypes.byref(out), rb.ctypes.data, arr, len(reqs), mask_comp, mask_struct,
                           ctypes.byref(dB.view), ctypes.byref(s), ctypes.byref(prm) if prm else None)
    assert rc == 0, f"gm_batch_saxpy3 rc={rc}: {capi.last_error()}"
    zname = {8: "int64", 6: "int32", 1: "bool", 11: "fp64"}[trip[2]]
    C = download(out, ZT[zname])
    for d in keep:
        d.free()
    return C, rb
