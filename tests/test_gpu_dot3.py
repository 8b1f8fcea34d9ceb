"""dot3 (C<M> = A*B') on the GPU against the oracle.

Bar: bit-exact pattern and values for bool/integer semirings; plus_times fp64
within 1e-12 relative error per entry (summation order may differ).  Calls go
through the C-ABI (gm_dot3) and through the drop-in GrB_mxm.
"""
import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from graphblas_amd import workloads as W
from helpers import SR, DevCsr, assert_csr_equal, gm_dot3, power_law_csr, rand_csr

pytestmark = pytest.mark.gpu

FP_TOL = 1e-12   # plus_times fp64, relative, per entry


def _vals(handle):
    return {"GrB_PLUS_TIMES_SEMIRING_FP64": "fp64", "GrB_MIN_PLUS_SEMIRING_INT32": "int32",
            "GrB_PLUS_TIMES_SEMIRING_INT64": "int64", "GrB_MIN_PLUS_SEMIRING_INT64": "int64",
            "GrB_MAX_PLUS_SEMIRING_INT32": "int32", "GrB_MIN_PLUS_SEMIRING_FP64": "fp64"}.get(handle)


@pytest.mark.parametrize("scale", [8, 10, 12, 14])
def test_tricount_rmat_grb_mxm(grb, scale):
    """C<L> = L*U' with plus_pair_int64 (GrB_DESC_ST1), the triangle count."""
    print(f"[dot3] tricount rmat scale={scale}", flush=True)
    A, L, U = W.tricount_inputs(scale, 16, seed=1)
    Ld, Ud = capi.Matrix.from_csr(L, "bool"), capi.Matrix.from_csr(U, "bool")
    C = capi.Matrix("int64", L.nrows, L.nrows)
    capi.mxm_checked(C, Ld, "GxB_PLUS_PAIR_INT64", Ld, Ud, "ST1")
    ref = O.dot3("plus_pair_int64", L, L, U)
    assert_csr_equal(C.to_csr(), ref, what="tricount")
    assert capi.reduce_sum_int64(C) == int(ref.x.sum())


@pytest.mark.parametrize("handle", list(SR))
@pytest.mark.parametrize("shape", ["uniform", "powerlaw", "empty_rows"])
def test_dot3_semirings_cabi(grb, handle, shape):
    print(f"[dot3] {handle} {shape}", flush=True)
    rng = np.random.default_rng(hash((handle, shape)) % (2 ** 32))
    vt = _vals(handle)
    if shape == "uniform":
        A = rand_csr(300, 200, 0.05, rng, vt)
        B = rand_csr(250, 200, 0.05, rng, vt)
        M = rand_csr(300, 250, 0.05, rng, "int64")
    elif shape == "powerlaw":
        A = power_law_csr(400, 12, rng, values=vt)
        B = power_law_csr(400, 12, rng, values=vt)
        M = power_law_csr(400, 10, rng, values="int64")
    else:
        A = rand_csr(300, 300, 0.03, rng, vt, empty_rows=0.3)
        B = rand_csr(300, 300, 0.03, rng, vt, empty_rows=0.3)
        M = rand_csr(300, 300, 0.05, rng, "int64", empty_rows=0.2)
    name, _ = SR[handle]
    ref = O.dot3(name, M, A, B, mask_struct=True)
    got = gm_dot3(handle, M, A, B)
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what=handle)


@pytest.mark.parametrize("params", [(1, 1, 1, -1), (1, 2, 100000, -1), (1, 10 ** 9, 10 ** 9, -1),
                                    (10 ** 9, 10 ** 9, 10 ** 9, -1), (0, 0, 0, 0),
                                    (1, 1, 1, 1), (1, 2, 100000, 1), (1, 10 ** 9, 10 ** 9, 1),
                                    (10 ** 9, 10 ** 9, 10 ** 9, 1), (0, 0, 0, 1), (0, 0, 0, 20)])
@pytest.mark.parametrize("handle", ["GxB_PLUS_PAIR_INT64", "GrB_PLUS_TIMES_SEMIRING_FP64",
                                    "GrB_MIN_PLUS_SEMIRING_INT32", "GxB_ANY_SECONDI_INT64"])
def test_dot3_forced_bins(grb, params, handle):
    """Every bin (thread / team16 / wave / workgroup), with and without row
    bitmaps (bitmap_min_len 1: every row probed through a bitmap), gives the
    oracle's result."""
    print(f"[dot3] forced bins {params} {handle}", flush=True)
    rng = np.random.default_rng(7)
    vt = _vals(handle)
    A = power_law_csr(500, 20, rng, values=vt)
    B = power_law_csr(500, 20, rng, values=vt)
    M = power_law_csr(500, 15, rng, values="int64")
    name, _ = SR[handle]
    ref = O.dot3(name, M, A, B)
    # row mode off, so the chosen bins see every entry
    got = gm_dot3(handle, M, A, B, params=tuple(params) + (0, -1))
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what=str(params))


def test_dot3_valued_mask(grb):
    """A valued mask: entries whose value is 0 do not produce C(i,j)."""
    rng = np.random.default_rng(3)
    A = rand_csr(200, 150, 0.08, rng, "int32")
    B = rand_csr(180, 150, 0.08, rng, "int32")
    M = rand_csr(200, 180, 0.1, rng, "int64")
    M.x[rng.random(M.nnz) < 0.4] = 0
    ref = O.dot3("min_plus_int32", M, A, B, mask_struct=False)
    got = gm_dot3("GrB_MIN_PLUS_SEMIRING_INT32", M, A, B, mask_struct=False)
    assert_csr_equal(got, ref, what="valued mask")


def test_dot3_empty(grb):
    rng = np.random.default_rng(4)
    A = rand_csr(50, 40, 0.0, rng)
    B = rand_csr(60, 40, 0.1, rng)
    M = rand_csr(50, 60, 0.0, rng, "int64")
    got = gm_dot3("GxB_PLUS_PAIR_INT64", M, A, B)
    assert got.nnz == 0 and np.all(got.p == 0)
    M2 = rand_csr(50, 60, 0.2, rng, "int64")
    got = gm_dot3("GxB_PLUS_PAIR_INT64", M2, A, B)
    assert got.nnz == 0 and np.all(got.p == 0)


def test_dot3_grb_mxm_min_plus_batch_shape(grb):
    """C<A> = A*A' with min_plus int32 on a weighted RMAT (the batched shape)."""
    G = W.rmat_sym(11, 8, seed=5)
    G.x = W.sym_weights_i32(G, seed=9)
    Gd = capi.Matrix.from_csr(G, "int32")
    C = capi.Matrix("int32", G.nrows, G.nrows)
    capi.mxm_checked(C, Gd, "GrB_MIN_PLUS_SEMIRING_INT32", Gd, Gd, "ST1")
    ref = O.dot3("min_plus_int32", G, G, G)
    assert_csr_equal(C.to_csr(), ref, what="min_plus C<A>=A*A'")


def test_dot3_deterministic_fp64(grb):
    rng = np.random.default_rng(12)
    A = power_law_csr(400, 30, rng, values="fp64")
    M = power_law_csr(400, 30, rng, values="int64")
    r1 = gm_dot3("GrB_PLUS_TIMES_SEMIRING_FP64", M, A, A)
    r2 = gm_dot3("GrB_PLUS_TIMES_SEMIRING_FP64", M, A, A)
    assert np.array_equal(r1.x, r2.x)


@pytest.mark.parametrize("handle", ["GxB_PLUS_PAIR_INT64", "GrB_PLUS_TIMES_SEMIRING_FP64",
                                    "GrB_MIN_PLUS_SEMIRING_INT32", "GxB_ANY_SECONDI_INT64",
                                    "GxB_ANY_PAIR_BOOL"])
def test_dot3_bitmaps_distinct_operands(grb, handle):
    """A and B distinct with long rows on both sides (separate bitmap sets)."""
    rng = np.random.default_rng(99)
    vt = _vals(handle)
    A = power_law_csr(600, 40, rng, values=vt)
    B = power_law_csr(600, 40, rng, values=vt)
    M = power_law_csr(600, 20, rng, values="int64")
    name, _ = SR[handle]
    ref = O.dot3(name, M, A, B)
    got = gm_dot3(handle, M, A, B, params=(0, 0, 0, 8))
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what="bitmaps")


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("params", [(1, 2, 100000, -1), (1, 64, 100000, 8), (1, 16, 100000, 1),
                                    (0, 0, 0, 0), (1, 2, 3, -1), (1, 2, 3, 8)])
@pytest.mark.parametrize("handle", ["GxB_PLUS_PAIR_INT64", "GrB_PLUS_TIMES_SEMIRING_FP64",
                                    "GrB_MIN_PLUS_SEMIRING_INT32", "GxB_ANY_SECONDI_INT64"])
def test_dot3_entry_order(grb, order, params, handle):
    """Team entries in mask order (1) and grouped by probed row over the XCD
    regions (2) give the oracle's result, with and without probe indexes."""
    print(f"[dot3] order={order} {params} {handle}", flush=True)
    rng = np.random.default_rng(21)
    vt = _vals(handle)
    A = power_law_csr(700, 25, rng, values=vt)
    B = power_law_csr(700, 25, rng, values=vt)
    M = power_law_csr(700, 20, rng, values="int64")
    name, _ = SR[handle]
    ref = O.dot3(name, M, A, B)
    got = gm_dot3(handle, M, A, B, params=tuple(params) + (order, -1))
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what=f"order {order}")


def test_dot3_grouped_tricount_default(grb):
    """A mask of more than 2^20 entries takes the grouped order by default."""
    A, L, U = W.tricount_inputs(17, 16, seed=2)
    assert L.nnz >= (1 << 20)
    print(f"[dot3] grouped default tricount scale 17 nnz(L)={L.nnz}", flush=True)
    ref = O.dot3("plus_pair_int64", L, L, U)
    got = gm_dot3("GxB_PLUS_PAIR_INT64", L, L, U, abtype=1, mtype=1)
    assert_csr_equal(got, ref, what="tricount scale 17")


def _wide_rows(m, n, long_rows, long_len, short_len, rng):
    """m rows over n columns: rows in long_rows have about long_len sorted
    distinct columns spread over [0, n), the others about short_len; half of
    each row comes from 4000 shared columns."""
    hot = np.random.default_rng(7).choice(n, 4000, replace=False)   # shared by all rows
    p = [0]
    js = []
    for r in range(m):
        k = long_len if r in long_rows else short_len
        row = np.unique(np.concatenate([rng.choice(hot, k // 2, replace=False),
                                        rng.choice(n, k - k // 2, replace=False)]))
        js.append(row.astype(np.int64))
        p.append(p[-1] + len(row))
    return W.Csr(m, n, np.array(p, dtype=np.int64), np.concatenate(js), None)


@pytest.mark.parametrize("bins", [(2, 100000), (2, 3), (100000, 100000)])
@pytest.mark.parametrize("handle", ["GxB_PLUS_PAIR_INT64", "GxB_PLUS_PAIR_INT32", "GxB_ANY_PAIR_BOOL",
                                    "GxB_ANY_SECONDI_INT64"])
def test_dot3_lds_chunks(grb, bins, handle):
    """Grouped bins with long rows spanning several LDS chunks of 2^19
    columns: the wavefront and workgroup bins (chunked LDS bitmaps) and the
    team16 bin (global probes), against the oracle."""
    team16_max, wave_max = bins
    print(f"[dot3] lds chunks {bins} {handle}", flush=True)
    rng = np.random.default_rng(hash((bins, handle)) % 2 ** 32)
    n = 1_500_000
    A = _wide_rows(240, n, set(range(0, 240, 12)), 3000, 150, rng)
    B = _wide_rows(200, n, set(range(3, 200, 10)), 2500, 120, rng)
    M = rand_csr(240, 200, 0.35, rng)
    name, _ = SR[handle]
    ref = O.dot3(name, M, A, B)
    got = gm_dot3(handle, M, A, B, params=(1, team16_max, wave_max, 8, 2, -1))
    assert_csr_equal(got, ref, what=f"lds chunks {bins}")


@pytest.mark.parametrize("params", [(1, 64, 100000, 8), (8, 512, 4096, 128), (1, 2, 100000, 16)])
@pytest.mark.parametrize("handle", list(SR))
def test_dot3_lds_whole_rows(grb, params, handle):
    """Every row spans at most 2^16 columns: the grouped team16 and wavefront
    bins probe whole long-row bitmaps (with word ranks for valued semirings)
    from LDS.  Against the oracle, structural and valued masks."""
    print(f"[dot3] lds whole rows {params} {handle}", flush=True)
    rng = np.random.default_rng(hash((params, handle)) % 2 ** 32)
    vt = _vals(handle)
    A = power_law_csr(3000, 40, rng, values=vt)
    B = power_law_csr(3000, 40, rng, values=vt)
    M = power_law_csr(3000, 30, rng, values="int64")
    name, _ = SR[handle]
    for struct in (True, False):
        ref = O.dot3(name, M, A, B, mask_struct=struct)
        got = gm_dot3(handle, M, A, B, mask_struct=struct, params=tuple(params) + (2, -1))
        assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None,
                         what=f"lds whole rows struct={struct}")


@pytest.mark.parametrize("row_min", [1, 4])
@pytest.mark.parametrize("bitmaps", [-1, 8])
@pytest.mark.parametrize("handle", list(SR))
def test_dot3_row_mode(grb, row_min, bitmaps, handle):
    """Whole mask rows against an LDS bitmap of A(i,:) (with the probe of
    short A rows into long indexed B rows mixed in) give the oracle's result."""
    print(f"[dot3] row mode {row_min} {bitmaps} {handle}", flush=True)
    rng = np.random.default_rng(hash((row_min, bitmaps, handle)) % 2 ** 32)
    vt = _vals(handle)
    A = power_law_csr(600, 30, rng, values=vt)
    B = power_law_csr(600, 30, rng, values=vt)
    M = power_law_csr(600, 25, rng, values="int64")
    name, _ = SR[handle]
    for struct in (True, False):
        ref = O.dot3(name, M, A, B, mask_struct=struct)
        got = gm_dot3(handle, M, A, B, mask_struct=struct, params=(0, 0, 0, bitmaps, 0, row_min))
        assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None,
                         what=f"row mode {row_min} struct={struct}")


def test_dot3_row_mode_wide_rows(grb):
    """A rows spanning more than 2^17 columns stay with the binned kernels."""
    rng = np.random.default_rng(8)
    n = 300000
    A = rand_csr(200, n, 0.0005, rng, "int32")
    B = rand_csr(150, n, 0.0005, rng, "int32")
    M = rand_csr(200, 150, 0.3, rng, "int64")
    ref = O.dot3("min_plus_int32", M, A, B)
    got = gm_dot3("GrB_MIN_PLUS_SEMIRING_INT32", M, A, B, params=(0, 0, 0, 0, 0, 1))
    assert_csr_equal(got, ref, what="wide rows")


def test_dot3_work_model(grb):
    """gm_dot3_work (the list lengths behind the roofline's algorithmic bytes)
    against numpy on a power-law mask."""
    import ctypes
    rng = np.random.default_rng(21)
    A = power_law_csr(600, 12, rng)
    B = power_law_csr(500, 12, rng)
    M = rand_csr(600, 500, 0.05, rng)
    la, lb = np.diff(A.p), np.diff(B.p)
    rows = np.repeat(np.arange(M.nrows), np.diff(M.p))
    a, b = la[rows], lb[M.j]
    dM, dA, dB = DevCsr(M, 1), DevCsr(A, 1), DevCsr(B, 1)
    mn, ln = ctypes.c_int64(), ctypes.c_int64()
    rc = capi.lib().gm_dot3_work(ctypes.byref(mn), ctypes.byref(ln), ctypes.byref(dM.view),
                                 ctypes.byref(dA.view), ctypes.byref(dB.view))
    assert rc == 0, capi.last_error()
    assert mn.value == int(np.minimum(a, b).sum()) and ln.value == int((a + b).sum())


@pytest.mark.parametrize("slice_max", [1, 700, 5000])
def test_dot3_row_slices(grb, slice_max):
    """Masks beyond the 2^32 entries one run packs are cut into row slices and
    the slices' results joined; the slice size is lowered here so that small
    masks take that path: a valued mask, an iso result and the triangle
    count through GrB_mxm, all against the oracle."""
    print(f"[dot3] row slices of {slice_max} entries", flush=True)
    L = capi.lib()
    rng = np.random.default_rng(9)
    A = rand_csr(300, 150, 0.06, rng, "int32")
    B = rand_csr(250, 150, 0.06, rng, "int32")
    M = rand_csr(300, 250, 0.08, rng, "int64")
    M.x[rng.random(M.nnz) < 0.3] = 0
    _, Lr, Ur = W.tricount_inputs(9, 16, seed=2)
    L.gm_debug_dot3_slice_max(slice_max)
    try:
        ref = O.dot3("min_plus_int32", M, A, B, mask_struct=False)
        got = gm_dot3("GrB_MIN_PLUS_SEMIRING_INT32", M, A, B, mask_struct=False)
        assert_csr_equal(got, ref, what="sliced valued mask")
        ref = O.dot3("any_pair_bool", M, A, B)
        got = gm_dot3("GxB_ANY_PAIR_BOOL", M, A, B)
        assert_csr_equal(got, ref, what="sliced iso")
        Ld, Ud = capi.Matrix.from_csr(Lr, "bool"), capi.Matrix.from_csr(Ur, "bool")
        C = capi.Matrix("int64", Lr.nrows, Lr.nrows)
        capi.mxm_checked(C, Ld, "GxB_PLUS_PAIR_INT64", Ld, Ud, "ST1")
        ref = O.dot3("plus_pair_int64", Lr, Lr, Ur)
        assert_csr_equal(C.to_csr(), ref, what="sliced tricount")
    finally:
        L.gm_debug_dot3_slice_max(0)
