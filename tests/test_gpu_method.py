"""GrB_mxm's choice between saxpy3 and dot3 for C<M>=A*B: the flop count on the
GPU against numpy, and both methods against the oracle's masked saxpy3
(bit-exact for integers, fp64 within 1e-12)."""
import ctypes

import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from graphblas_amd import workloads as W
from helpers import DevCsr, assert_csr_equal, power_law_csr, rand_csr

pytestmark = pytest.mark.gpu


def _flops_host(A, B):
    rl = np.diff(B.p)
    return int(rl[A.j].sum()) if A.nnz else 0


@pytest.mark.parametrize("case", ["random", "powerlaw", "empty_a", "rect"])
def test_flopcount(grb, case):
    print(f"[method] flopcount {case}", flush=True)
    rng = np.random.default_rng(hash(case) % 2 ** 32)
    if case == "random":
        A, B = rand_csr(300, 200, 0.05, rng), rand_csr(200, 250, 0.05, rng)
    elif case == "powerlaw":
        A = power_law_csr(3000, 12, rng)
        B = A
    elif case == "empty_a":
        A, B = rand_csr(50, 40, 0.0, rng), rand_csr(40, 30, 0.2, rng)
    else:
        A, B = rand_csr(7, 5000, 0.01, rng), rand_csr(5000, 9, 0.3, rng)
    dA, dB = DevCsr(A, 1), DevCsr(B, 1)
    f = ctypes.c_int64(-1)
    rc = capi.lib().gm_flopcount(ctypes.byref(f), ctypes.byref(dA.view), ctypes.byref(dB.view))
    dA.free()
    dB.free()
    assert rc == 0, capi.last_error()
    assert f.value == _flops_host(A, B)


def _hub_graph(n, hubs, avg, rng):
    """Symmetric pattern: `hubs` vertices joined to every other vertex, plus
    random edges, no self loops.  Its square has about hubs*n^2 flops."""
    r = list(rng.integers(0, n, avg * n // 2))
    c = list(rng.integers(0, n, avg * n // 2))
    for h in range(hubs):
        v = n // 3 + 7 * h
        r += [v] * n
        c += list(range(n))
    r, c = np.array(r), np.array(c)
    keep = r != c
    r, c = np.concatenate([r[keep], c[keep]]), np.concatenate([c[keep], r[keep]])
    key = np.unique(r * n + c)
    rows, cols = key // n, key % n
    p = np.zeros(n + 1, dtype=np.int64)
    np.add.at(p, rows + 1, 1)
    return W.Csr(n, n, np.cumsum(p), cols.astype(np.int64), None)


def _grb_masked(A, M, sem_handle, tname, hint=None):
    L = capi.lib()
    d = None
    if hint is not None:
        d = ctypes.c_void_p()
        assert L.GrB_Descriptor_new(ctypes.byref(d)) == 0
        assert L.GrB_Descriptor_set(d, 7090, hint) == 0      # GxB_AxB_METHOD
    C = capi.Matrix(tname, A.nrows, A.ncols)
    Ad, Md = capi.Matrix.from_csr(A, tname), capi.Matrix.from_csr(M, tname)
    rc = L.GrB_mxm(C.h, Md.h, None, capi.handle(sem_handle), Ad.h, Ad.h, d)
    if d is not None:
        L.GrB_Descriptor_free(ctypes.byref(d))
    assert rc == 0, capi.last_error()
    return C.to_csr()


@pytest.mark.parametrize("sem", [("plus_pair_int64", "GxB_PLUS_PAIR_INT64", "int64"),
                                 ("plus_times_fp64", "GrB_PLUS_TIMES_SEMIRING_FP64", "fp64"),
                                 ("min_plus_int64", "GrB_MIN_PLUS_SEMIRING_INT64", "int64")])
def test_masked_product_by_flop_rule(grb, sem):
    """C<A>=A*A on a graph with hubs: the flop rule picks dot3 on A'; the
    result equals the oracle's masked saxpy3 and the GxB_AxB_SAXPY-forced run."""
    oname, handle, tname = sem
    print(f"[method] C<A>=A*A {oname}", flush=True)
    rng = np.random.default_rng(41)
    n = 2000
    A = _hub_graph(n, 2, 8, rng)
    if tname == "fp64":
        A = W.Csr(A.nrows, A.ncols, A.p, A.j, rng.standard_normal(A.nnz))
    elif oname != "plus_pair_int64":
        A = W.Csr(A.nrows, A.ncols, A.p, A.j, rng.integers(-50, 50, A.nnz).astype(np.int64))
    fl = _flops_host(A, A)
    assert capi.lib().GB_mxm_prefer_dot3(fl, A.nnz, A.nnz, n, A.nnz, n), "rule should pick dot3"
    # no descriptor: A's values are the mask's (an int64 zero drops out)
    ref = O.saxpy3(oname, A, A, M=A, mask_struct=A.x is None)
    tol = 1e-12 if tname == "fp64" else None
    got = _grb_masked(A, A, handle, tname)
    assert_csr_equal(got, ref, fp_tol=tol, what=f"{oname} (dot3 by flop rule)")
    forced = _grb_masked(A, A, handle, tname, hint=7085)          # GxB_AxB_SAXPY
    assert_csr_equal(forced, ref, fp_tol=tol, what=f"{oname} (saxpy3 forced)")


def test_masked_product_rule_keeps_saxpy(grb):
    """A sparse random product stays on saxpy3 and is still right."""
    print("[method] C<M>=A*B random", flush=True)
    rng = np.random.default_rng(43)
    A = rand_csr(400, 400, 0.02, rng, "int64")
    fl = _flops_host(A, A)
    assert not capi.lib().GB_mxm_prefer_dot3(fl, A.nnz, A.nnz, 400, A.nnz, 400)
    ref = O.saxpy3("plus_times_int64", A, A, M=A, mask_struct=False)
    got = _grb_masked(A, A, "GrB_PLUS_TIMES_SEMIRING_INT64", "int64")
    assert_csr_equal(got, ref, what="saxpy3 kept")
