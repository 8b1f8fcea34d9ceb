"""saxpy3 (C = A*B, C<M> = A*B, C<!M> = A*B) on the GPU against the oracle.

Bar: bit-exact for bool/integer semirings and ANY (the GPU keeps the last
writer in the reference's k order, like the oracle); plus_times fp64 within
1e-12 relative error per entry (the numeric phase applies products in the
reference's order, so it is usually bit-exact too).
"""
import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from graphblas_amd import workloads as W
from helpers import SR, assert_csr_equal, gm_saxpy3, power_law_csr, rand_csr

pytestmark = pytest.mark.gpu

FP_TOL = 1e-12


def _vals(handle):
    return {"GrB_PLUS_TIMES_SEMIRING_FP64": "fp64", "GrB_MIN_PLUS_SEMIRING_INT32": "int32",
            "GrB_PLUS_TIMES_SEMIRING_INT64": "int64", "GrB_MIN_PLUS_SEMIRING_INT64": "int64",
            "GrB_MAX_PLUS_SEMIRING_INT32": "int32", "GrB_MIN_PLUS_SEMIRING_FP64": "fp64"}.get(handle)


@pytest.mark.parametrize("n", [1024, 1 << 14])
def test_er_plus_times_grb_mxm(grb, n):
    """C = A*B, plus_times fp64, ER with exactly 8 entries per row (GrB_mxm)."""
    print(f"[saxpy3] ER n={n}", flush=True)
    A, B = W.er(n, n, 8, seed=1), W.er(n, n, 8, seed=2)
    Ad, Bd = capi.Matrix.from_csr(A, "fp64"), capi.Matrix.from_csr(B, "fp64")
    C = capi.Matrix("fp64", n, n)
    capi.mxm_checked(C, None, "GrB_PLUS_TIMES_SEMIRING_FP64", Ad, Bd)
    ref = O.saxpy3("plus_times_fp64", A, B)
    got = C.to_csr()
    assert_csr_equal(got, ref, fp_tol=FP_TOL, what="ER")
    # ascending-k accumulation order: identical bits to the oracle
    assert np.array_equal(got.x, ref.x)


@pytest.mark.parametrize("handle", list(SR))
@pytest.mark.parametrize("shape", ["uniform", "powerlaw", "empty_rows"])
def test_saxpy3_semirings_cabi(grb, handle, shape):
    print(f"[saxpy3] {handle} {shape}", flush=True)
    rng = np.random.default_rng(hash((handle, shape, 1)) % (2 ** 32))
    vt = _vals(handle)
    if shape == "uniform":
        A = rand_csr(300, 200, 0.04, rng, vt)
        B = rand_csr(200, 250, 0.04, rng, vt)
    elif shape == "powerlaw":
        A = power_law_csr(500, 10, rng, values=vt)
        B = power_law_csr(500, 10, rng, values=vt)
    else:
        A = rand_csr(300, 300, 0.03, rng, vt, empty_rows=0.3)
        B = rand_csr(300, 300, 0.03, rng, vt, empty_rows=0.3)
    name, _ = SR[handle]
    ref = O.saxpy3(name, A, B)
    got = gm_saxpy3(handle, A, B)
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what=handle)


@pytest.mark.parametrize("comp", [False, True])
@pytest.mark.parametrize("struct", [True, False])
@pytest.mark.parametrize("handle", ["GxB_PLUS_PAIR_INT64", "GrB_PLUS_TIMES_SEMIRING_FP64",
                                    "GxB_ANY_SECONDI_INT64", "GrB_MIN_PLUS_SEMIRING_INT32"])
def test_saxpy3_masked(grb, comp, struct, handle):
    print(f"[saxpy3] masked comp={comp} struct={struct} {handle}", flush=True)
    rng = np.random.default_rng(21 + comp + 2 * struct)
    vt = _vals(handle)
    A = power_law_csr(400, 8, rng, values=vt)
    B = power_law_csr(400, 8, rng, values=vt)
    M = rand_csr(400, 400, 0.05, rng, "int64")
    if not struct:
        M.x[rng.random(M.nnz) < 0.3] = 0
    name, _ = SR[handle]
    ref = O.saxpy3(name, A, B, M=M, mask_comp=comp, mask_struct=struct)
    got = gm_saxpy3(handle, A, B, M=M, mask_comp=comp, mask_struct=struct)
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what="masked")


@pytest.mark.parametrize("params", [(1, 1, 1, 0), (64, 512, 2048, 0), (1, 512, 2048, 0),
                                    (64, 64, 64, 0), (64, 512, 2048, 2), (16, 512, 2048, 1),
                                    (64, 64, 64, 2), (8, 16, 32, 1)])
@pytest.mark.parametrize("handle", ["GxB_PLUS_PAIR_INT64", "GrB_PLUS_TIMES_SEMIRING_FP64",
                                    "GxB_ANY_SECONDI_INT64"])
def test_saxpy3_forced_bins(grb, params, handle):
    """wave128 / wave1024 / block4096 / global tables agree."""
    print(f"[saxpy3] forced bins {params} {handle}", flush=True)
    rng = np.random.default_rng(5)
    vt = _vals(handle)
    A = power_law_csr(600, 12, rng, values=vt)
    B = power_law_csr(600, 12, rng, values=vt)
    name, _ = SR[handle]
    ref = O.saxpy3(name, A, B)
    got = gm_saxpy3(handle, A, B, params=params)
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what=str(params))


def test_bfs_step_grb_mxm(grb):
    """C<!V> = F*A with any_secondi int64 and GrB_DESC_RSC (frontier expansion)."""
    G = W.rmat_sym(12, 16, seed=3)
    n = G.nrows
    rng = np.random.default_rng(8)
    k = 37
    src = rng.choice(n, size=k, replace=False)
    F = W.Csr(k, n, np.arange(k + 1, dtype=np.int64), np.sort(src)[np.argsort(np.argsort(src))].astype(np.int64))
    F = W.Csr(k, n, np.arange(k + 1, dtype=np.int64), src.astype(np.int64))
    V = F  # visited = the sources
    Fd = capi.Matrix.from_csr(F, "bool")
    Gd = capi.Matrix.from_csr(G, "bool")
    Vd = capi.Matrix.from_csr(V, "bool")
    C = capi.Matrix("int64", k, n)
    capi.mxm_checked(C, Vd, "GxB_ANY_SECONDI_INT64", Fd, Gd, "RSC")
    got = C.to_csr()
    ref = O.saxpy3("any_secondi_int64", F, G, M=V, mask_comp=True, mask_struct=True)
    assert_csr_equal(got, ref, what="bfs")
    # witness validity: C(r, i) = k requires F(r, k) and A(k, i)
    for r in range(k):
        for q in range(got.p[r], got.p[r + 1]):
            kk = got.x[q]
            assert kk == src[r]
            assert got.j[q] in G.j[G.p[kk]:G.p[kk + 1]]


def test_saxpy3_rmat_square_heavy_rows(grb):
    """A*A on a power-law graph: heavy rows exercise block and global tables."""
    G = W.rmat_sym(10, 16, seed=2)
    Gd = capi.Matrix.from_csr(G, "bool")
    C = capi.Matrix("int64", G.nrows, G.nrows)
    capi.mxm_checked(C, None, "GxB_PLUS_PAIR_INT64", Gd, Gd)
    ref = O.saxpy3("plus_pair_int64", G, G)
    assert_csr_equal(C.to_csr(), ref, what="A*A")


def test_saxpy3_empty(grb):
    rng = np.random.default_rng(1)
    A = rand_csr(40, 30, 0.0, rng, "fp64")
    B = rand_csr(30, 50, 0.2, rng, "fp64")
    got = gm_saxpy3("GrB_PLUS_TIMES_SEMIRING_FP64", A, B)
    assert got.nnz == 0 and np.all(got.p == 0)


def test_saxpy3_deterministic(grb):
    rng = np.random.default_rng(2)
    A = power_law_csr(800, 20, rng, values="fp64")
    r1 = gm_saxpy3("GrB_PLUS_TIMES_SEMIRING_FP64", A, A)
    r2 = gm_saxpy3("GrB_PLUS_TIMES_SEMIRING_FP64", A, A)
    assert np.array_equal(r1.x, r2.x) and np.array_equal(r1.j, r2.j)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("handle", ["GrB_PLUS_TIMES_SEMIRING_FP64", "GxB_ANY_PAIR_BOOL",
                                    "GrB_MIN_PLUS_SEMIRING_INT32", "GxB_PLUS_PAIR_INT32"])
def test_saxpy3_fused_vs_two_phase(grb, mode, handle):
    """The single-pass fused path and the symbolic+numeric path agree with the oracle
    on a mix of small and big rows (ER rows plus power-law rows)."""
    print(f"[saxpy3] mode={mode} {handle}", flush=True)
    rng = np.random.default_rng(40 + mode)
    vt = _vals(handle)
    A = power_law_csr(700, 6, rng, values=vt)
    B = power_law_csr(700, 6, rng, values=vt)
    name, _ = SR[handle]
    ref = O.saxpy3(name, A, B)
    got = gm_saxpy3(handle, A, B, params=(64, 512, 2048, mode))
    assert_csr_equal(got, ref, fp_tol=FP_TOL if "fp64" in name else None, what=f"mode {mode}")


@pytest.mark.parametrize("mode", [1, 2])
def test_saxpy3_fused_masked_bfs_shape(grb, mode):
    """C<!M> = F*A with many single-entry rows (small and hub sources)."""
    G = W.rmat_sym(11, 16, seed=6)
    n = G.nrows
    rng = np.random.default_rng(9)
    k = 300
    src = rng.choice(n, size=k, replace=False).astype(np.int64)
    src[:5] = np.argsort(-np.diff(G.p))[:5]       # include the hubs
    F = W.Csr(k, n, np.arange(k + 1, dtype=np.int64), src)
    V = rand_csr(k, n, 0.01, rng, "int64")
    ref = O.saxpy3("any_secondi_int64", F, G, M=V, mask_comp=True)
    got = gm_saxpy3("GxB_ANY_SECONDI_INT64", F, G, M=V, mask_comp=True,
                    params=(64, 512, 2048, mode))
    assert_csr_equal(got, ref, what=f"bfs mode {mode}")


def _short_rows_csr(n, ncols, maxlen, rng, vt, full=0.0):
    """CSR with distinct sorted columns; a fraction `full` of the rows have
    maxlen entries, the others a length uniform in [0, maxlen]."""
    lens = np.where(rng.random(n) < full, maxlen, rng.integers(0, maxlen + 1, size=n))
    p = np.zeros(n + 1, dtype=np.int64)
    p[1:] = np.cumsum(lens)
    j = np.concatenate([np.sort(rng.choice(ncols, size=int(l), replace=False)) for l in lens]
                       ).astype(np.int64) if p[-1] else np.zeros(0, dtype=np.int64)
    if vt == "fp64":
        x = rng.random(int(p[-1]))
    else:
        x = rng.integers(-50, 50, size=int(p[-1])).astype(np.int64)
    return W.Csr(n, ncols, p, j, x)


@pytest.mark.parametrize("mask", ["none", "M", "!M"])
@pytest.mark.parametrize("handle", ["GrB_PLUS_TIMES_SEMIRING_FP64", "GrB_PLUS_TIMES_SEMIRING_INT64",
                                    "GrB_MIN_PLUS_SEMIRING_INT64"])
def test_saxpy3_row_records_grb_mxm(grb, mask, handle):
    """GrB_mxm with B's row records (rows of at most 16 entries at stride 16,
    half of them full, so that the padding stays under the records' limit and
    the fused kernel reads long rows in two groups of eight) and rows of A
    whose flops cross the single-pass limit: against the oracle."""
    print(f"[saxpy3] row records {handle} mask={mask}", flush=True)
    rng = np.random.default_rng(77 + len(mask))
    vt = "fp64" if "FP64" in handle else "int64"
    n = 1500
    A = _short_rows_csr(n, n, 9, rng, vt)
    B = _short_rows_csr(n, n, 16, rng, vt, full=0.5)
    pad = 16 * n - B.nnz
    assert pad <= B.nnz // 2 + 1024, "B must qualify for row records"
    Ad, Bd = capi.Matrix.from_csr(A, vt), capi.Matrix.from_csr(B, vt)
    C = capi.Matrix(vt, n, n)
    name, _ = SR[handle]
    if mask == "none":
        capi.mxm_checked(C, None, handle, Ad, Bd)
        ref = O.saxpy3(name, A, B)
    else:
        M = rand_csr(n, n, 0.02, rng, "int64")
        Md = capi.Matrix.from_csr(M, "int64")
        d = capi.Descriptor(mask_struct=True, mask_comp=(mask == "!M"), axb=capi.GxB_AxB_SAXPY)
        try:
            capi.mxm_checked(C, Md, handle, Ad, Bd, d)
        finally:
            d.free()
        ref = O.saxpy3(name, A, B, M=M, mask_comp=(mask == "!M"), mask_struct=True)
    assert_csr_equal(C.to_csr(), ref, fp_tol=FP_TOL if vt == "fp64" else None, what=f"records {mask}")


@pytest.mark.parametrize("table", [0, 1, 2])
@pytest.mark.parametrize("ncols,bins", [(300, (8, 16, 64, 2)), (3000, (8, 16, 64, 2)),
                                        (3000, (8, 512, 2048, 2)), (40000, (8, 16, 32, 2))])
@pytest.mark.parametrize("handle", ["GxB_PLUS_PAIR_INT64", "GrB_PLUS_TIMES_SEMIRING_FP64",
                                    "GrB_MIN_PLUS_SEMIRING_INT32"])
def test_saxpy3_hash_and_gustavson_tables(grb, table, ncols, bins, handle):
    """Row tables indexed by hashing (table 1), by the column itself where a
    table holds every column of C (table 2, Gustavson's dense workspace, also
    for global-bin rows), or by the reference's rule (0); with the bins forced
    small so that rows land in the LDS tables of 128 / 1024 / 4096 slots and
    in HBM tables, and with a mask; all against the oracle."""
    print(f"[saxpy3] table={table} ncols={ncols} bins={bins} {handle}", flush=True)
    rng = np.random.default_rng(ncols + table)
    vt = _vals(handle)
    A = power_law_csr(400, 10, rng, values=vt)
    B = rand_csr(400, ncols, min(0.5, 40.0 / ncols), rng, vt)
    M = rand_csr(400, ncols, min(0.5, 60.0 / ncols), rng, "int64")
    name, _ = SR[handle]
    tol = FP_TOL if "fp64" in name else None
    params = tuple(bins) + (table,)
    ref = O.saxpy3(name, A, B)
    got = gm_saxpy3(handle, A, B, params=params)
    assert_csr_equal(got, ref, fp_tol=tol, what=f"table {table}")
    ref = O.saxpy3(name, A, B, M=M, mask_comp=True, mask_struct=True)
    got = gm_saxpy3(handle, A, B, M=M, mask_comp=True, params=params)
    assert_csr_equal(got, ref, fp_tol=tol, what=f"table {table} <!M>")


@pytest.mark.parametrize("axb", ["hash", "gustavson"])
def test_saxpy3_method_hints_grb_mxm(grb, axb):
    """GrB_mxm with the GxB_AxB_HASH / GxB_AxB_GUSTAVSON descriptor hints."""
    print(f"[saxpy3] GrB_mxm hint {axb}", flush=True)
    rng = np.random.default_rng(3)
    A = power_law_csr(500, 12, rng, values="fp64")
    B = rand_csr(500, 2000, 0.02, rng, "fp64")
    Ad, Bd = capi.Matrix.from_csr(A, "fp64"), capi.Matrix.from_csr(B, "fp64")
    C = capi.Matrix("fp64", 500, 2000)
    d = capi.Descriptor(axb=capi.GxB_AxB_HASH if axb == "hash" else capi.GxB_AxB_GUSTAVSON)
    try:
        capi.mxm_checked(C, None, "GrB_PLUS_TIMES_SEMIRING_FP64", Ad, Bd, d)
    finally:
        d.free()
    assert_csr_equal(C.to_csr(), O.saxpy3("plus_times_fp64", A, B), fp_tol=FP_TOL, what=axb)
