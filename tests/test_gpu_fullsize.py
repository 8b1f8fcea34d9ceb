"""GPU parity at the full sizes of the benchmark workloads.

* Erdos-Renyi n = 2^20, 8 entries per row, C = A*B plus_times fp64 through
  GrB_mxm: pattern exact and values within 1e-12 relative of the oracle, and
  the max relative error of both paths (the oracle's ascending-k sum and the
  GPU) stated against the compensated double-double oracle.
* RMAT scale 22 triangle count C<L> = L*U' plus_pair int64 through GrB_mxm:
  nnz(C), sum(C) and 64 per-band digests of the whole C against the fixture
  that tests/golden/make_fullsize.py wrote with the oracle, and two bands of
  rows (one holding the hub rows) bit-exact against the live oracle.
* The 4096-graph batch (n log-uniform in [512, 65536], edge factor in
  [4, 32], plus_pair int32 / min_plus int32): every request of the batched
  path equal to the one-call-per-request path, and 256 requests spread over
  the sizes equal to the oracle.
* Served BFS steps on RMAT scale 20 with k log-uniform in [1, 256] sources:
  every request equal to the oracle.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import pyoracle as O
from digest import band_digest
from graphblas_amd import capi
from graphblas_amd import workloads as W
from helpers import DevCsr, assert_csr_equal

pytestmark = pytest.mark.gpu

FP_TOL = 1e-12          # plus_times fp64, relative, per entry
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_saxpy3_er20_plus_times_fp64(grb):
    n = 1 << 20
    print("[fullsize] saxpy3 ER 2^20 x 8 plus_times fp64", flush=True)
    A, B = W.er(n, n, 8, seed=1), W.er(n, n, 8, seed=2)
    Ad, Bd = capi.Matrix.from_csr(A, "fp64"), capi.Matrix.from_csr(B, "fp64")
    C = capi.Matrix("fp64", n, n)
    capi.mxm_checked(C, None, "GrB_PLUS_TIMES_SEMIRING_FP64", Ad, Bd)
    got = C.to_csr()
    del C, Ad, Bd
    ref = O.saxpy3("plus_times_fp64", A, B)
    assert np.array_equal(got.p, ref.p), "row pointers"
    assert np.array_equal(got.j, ref.j), "column indices"
    dd = O.saxpy3_plus_times_fp64_dd(A, B, ref)
    den = np.abs(dd)
    assert np.all(den > 0)            # values in [0,1): no cancellation to zero
    err_gpu = float(np.max(np.abs(got.x - dd) / den))
    err_seq = float(np.max(np.abs(ref.x - dd) / den))
    err_vs = float(np.max(np.abs(got.x - ref.x) / np.abs(ref.x)))
    print(f"[fullsize] nnz(C)={len(got.j)} max rel err vs compensated: GPU {err_gpu:.3g}, "
          f"sequential oracle {err_seq:.3g}; GPU vs oracle {err_vs:.3g}", flush=True)
    assert err_vs <= FP_TOL and err_gpu <= FP_TOL and err_seq <= FP_TOL
    # linearity: sum(A*B) = (A' 1) . (B 1)
    colsum_a = np.bincount(A.j, weights=A.x, minlength=n)
    rowsum_b = np.add.reduceat(B.x, B.p[:-1])
    assert abs(O.neumaier_sum(got.x) - float(colsum_a @ rowsum_b)) <= 1e-9 * float(colsum_a @ rowsum_b)


def test_tricount_rmat22_bands(grb):
    with open(os.path.join(GOLDEN, "fullsize_rmat22.json")) as f:
        gold = json.load(f)
    print("[fullsize] dot3 tricount RMAT 22", flush=True)
    _, L, U = W.tricount_inputs(22, 16, seed=1)
    assert L.nnz == gold["nnz_L"]
    Ld, Ud = capi.Matrix.from_csr(L, "bool"), capi.Matrix.from_csr(U, "bool")
    C = capi.Matrix("int64", L.nrows, L.nrows)
    capi.mxm_checked(C, Ld, "GxB_PLUS_PAIR_INT64", Ld, Ud, "ST1")
    assert capi.reduce_sum_int64(C) == gold["sum_C"]
    got = C.to_csr()
    del C, Ld, Ud
    assert int(got.p[-1]) == gold["nnz_C"]
    bands = band_digest(got.p, got.j, got.x, gold["nbands"])
    for b, ((nz, d), (gnz, gd)) in enumerate(zip(bands, gold["bands"])):
        assert nz == gnz and d == int(gd), f"band {b} of {gold['nbands']} differs"
    # two bands bit-exact against the live oracle: the first (hub rows, low
    # vertex ids under RMAT's skew) and one from the middle
    n = L.nrows
    for r0 in (0, n // 2):
        r1 = r0 + n // 64
        Ls = W.Csr(r1 - r0, L.ncols, L.p[r0:r1 + 1] - L.p[r0], L.j[L.p[r0]:L.p[r1]].copy())
        # mask and A operand: rows r0..r1 of L; B: all of U
        ref = O.dot3("plus_pair_int64", Ls, Ls, U)
        sub = W.Csr(r1 - r0, n, got.p[r0:r1 + 1] - got.p[r0], got.j[got.p[r0]:got.p[r1]],
                    got.x[got.p[r0]:got.p[r1]])
        assert_csr_equal(sub, ref, what=f"rows {r0}..{r1}")


BATCH_SR = {"plus_pair_int32": (1, 3, 6), "min_plus_int32": (2, 2, 6)}
ORACLE_SR = {"plus_pair_int32": "plus_pair_int32", "min_plus_int32": "min_plus_int32"}


def _download_rows(out, r0, r1, col_base, zdtype):
    """Rows [r0, r1) of a CSR result in HBM as host arrays (p from 0)."""
    L = capi.lib()
    p = np.empty(r1 - r0 + 1, dtype=np.int64)
    L.gm_memcpy_d2h(p.ctypes.data, out.p + 8 * r0, p.nbytes)
    lo, hi = int(p[0]), int(p[-1])
    j = np.empty(max(hi - lo, 1), dtype=np.int64)
    x = np.empty(max(hi - lo, 1), dtype=zdtype)
    if hi > lo:
        L.gm_memcpy_d2h(j.ctypes.data, out.j + 8 * lo, 8 * (hi - lo))
        if out.iso:
            L.gm_memcpy_d2h(x.ctypes.data, out.x, x.itemsize)
            x[:] = x[0]
        else:
            L.gm_memcpy_d2h(x.ctypes.data, out.x + x.itemsize * lo, x.itemsize * (hi - lo))
    return p - lo, j[:hi - lo] - col_base, x[:hi - lo]


def test_batch_4096_full(grb):
    """The whole 4096-request batch: batched == one call per request for every
    request, and 256 requests (every 16th by size) == the oracle."""
    L = capi.lib()
    specs = W.batch_specs(4096, seed=1)
    print("[fullsize] batch of 4096 graphs: generating", flush=True)
    db = W.DeviceBatch(specs)
    try:
        groups = {name: [g for g, s in enumerate(specs) if s.semiring == name] for name in BATCH_SR}
        stacked = {}
        for name, idx in groups.items():
            reqs = (capi.GmDot3Request * len(idx))()
            for r, g in enumerate(idx):
                v = db.views[g]
                reqs[r] = capi.GmDot3Request(v, v, v)
            rb = np.zeros(len(idx) + 1, dtype=np.int64)
            cb = np.zeros(len(idx) + 1, dtype=np.int64)
            out = capi.GmCsrOut()
            sr = capi.GmSemiring(*BATCH_SR[name])
            rc = L.gm_batch_dot3(ctypes.byref(out), rb.ctypes.data, cb.ctypes.data, reqs, len(idx),
                                 False, ctypes.byref(sr), None)
            assert rc == 0, capi.last_error()
            stacked[name] = (out, rb, cb, {g: r for r, g in enumerate(idx)})
        print("[fullsize] batch: batched path done; comparing with one call per request", flush=True)
        order = np.argsort([s.n for s in specs], kind="stable")
        check = set(int(g) for g in order[::16])
        nchecked = 0
        for g, s in enumerate(specs):
            out, rb, cb, pos = stacked[s.semiring]
            r = pos[g]
            bp, bj, bx = _download_rows(out, int(rb[r]), int(rb[r + 1]), int(cb[r]), np.int32)
            one = capi.GmCsrOut()
            sr = capi.GmSemiring(*BATCH_SR[s.semiring])
            v = db.views[g]
            rc = L.gm_dot3(ctypes.byref(one), ctypes.byref(v), False, ctypes.byref(v),
                           ctypes.byref(v), ctypes.byref(sr), None)
            assert rc == 0, capi.last_error()
            pp, pj, px = _download_rows(one, 0, s.n, 0, np.int32)
            for q in (one.p, one.j, one.x):
                L.gm_free(q)
            assert np.array_equal(bp, pp) and np.array_equal(bj, pj) and np.array_equal(bx, px), \
                f"request {g} (n={s.n}, ef={s.ef}, {s.semiring}): batched != one call"
            if g in check:
                A = W.batch_graph_host(s)
                ref = O.dot3(ORACLE_SR[s.semiring], A, A, A, mask_struct=False)
                assert_csr_equal(W.Csr(s.n, s.n, pp, pj, px), ref,
                                 what=f"request {g} (n={s.n}) vs oracle")
                nchecked += 1
        assert nchecked == 256
        for out, _, _, _ in stacked.values():
            for q in (out.p, out.j, out.x):
                L.gm_free(q)
    finally:
        db.free()


def test_served_rmat20_k_loguniform(grb):
    """BFS steps C<!V> = F*A (any_secondi int64) on RMAT scale 20 with k
    log-uniform in [1, 256] sources, served in mixed batches: every request
    equals the oracle."""
    G = W.rmat_sym(20, 16, seed=1)
    dG = DevCsr(G, 1)
    rng = np.random.default_rng(11)
    reqs = []
    for r in range(512):
        k = W.loguniform_int(11, r, 1, 256)
        reqs.append(np.sort(rng.choice(G.nrows, size=k, replace=False)).astype(np.int64))
    print(f"[fullsize] served RMAT 20: {len(reqs)} requests, k up to {max(len(s) for s in reqs)}",
          flush=True)
    srv = capi.Server(dG.view, (4, 4, 8), workers=2, max_batch_reqs=256)
    try:
        ids = [srv.submit(s) for s in reqs]
        mixed = 0
        for src, rid in zip(reqs, ids):
            res = srv.wait(rid)
            assert res.status == 0, capi.last_error()
            mixed += res.batch_size > 1
            p, j, x = srv.fetch(rid, res)
            k = len(src)
            F = W.Csr(k, G.nrows, np.arange(k + 1, dtype=np.int64), src)
            ref = O.saxpy3("any_secondi_int64", F, G, M=F, mask_comp=True)
            assert_csr_equal(W.Csr(k, G.ncols, p, j, x), ref, what=f"request {rid} (k={k})")
            srv.release(rid)
        assert mixed > 0
    finally:
        srv.close()
        dG.free()
