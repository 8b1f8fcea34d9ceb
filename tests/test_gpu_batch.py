"""Batched GrB_mxm: many independent requests in one launch sequence.

Every request's slice of the stacked result must equal the oracle run on that
request alone (bit-exact for the integer semirings used here), both sliced on
the host and materialised on the device by gm_batch_split.
"""
import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import workloads as W
from helpers import (SR, assert_csr_equal, gm_batch_dot3, gm_batch_saxpy3, power_law_csr,
                     rand_csr, slice_stacked)

pytestmark = pytest.mark.gpu


def _graphs(nreq, seed, weights):
    rng = np.random.default_rng(seed)
    gs = []
    for r in range(nreq):
        scale = int(rng.integers(5, 11))
        ef = int(rng.integers(4, 17))
        G = W.rmat_sym(scale, ef, seed=seed * 1000 + r)
        if weights:
            G.x = W.sym_weights_i32(G, seed=r + 1)
        gs.append(G)
    return gs


@pytest.mark.parametrize("handle", ["GxB_PLUS_PAIR_INT32", "GrB_MIN_PLUS_SEMIRING_INT32"])
def test_batch_dot3_shared(grb, handle):
    """C_r<A_r> = A_r*A_r' for uneven graphs (the batched triangle/min-plus shape)."""
    print(f"[batch] dot3 shared {handle}", flush=True)
    weighted = handle == "GrB_MIN_PLUS_SEMIRING_INT32"
    gs = _graphs(40, 3, weighted)
    C, rb, cb, split = gm_batch_dot3(handle, [(G, G, G) for G in gs], abtype=6 if weighted else 1,
                                     shared=True)
    name, _ = SR[handle]
    for r, G in enumerate(gs):
        ref = O.dot3(name, G, G, G)
        got = slice_stacked(C, rb[r], rb[r + 1], cb[r])
        assert_csr_equal(got, ref, what=f"request {r}")
        assert_csr_equal(split[r], ref, what=f"split {r}")


@pytest.mark.parametrize("handle", ["GxB_PLUS_PAIR_INT64", "GrB_PLUS_TIMES_SEMIRING_FP64",
                                    "GxB_ANY_SECONDI_INT64"])
def test_batch_dot3_distinct(grb, handle):
    print(f"[batch] dot3 distinct {handle}", flush=True)
    rng = np.random.default_rng(11)
    vt = {"GrB_PLUS_TIMES_SEMIRING_FP64": "fp64"}.get(handle)
    reqs = []
    for r in range(25):
        nm, nb, k = int(rng.integers(1, 200)), int(rng.integers(1, 200)), int(rng.integers(1, 150))
        A = rand_csr(nm, k, 0.1, rng, vt)
        B = rand_csr(nb, k, 0.1, rng, vt)
        M = rand_csr(nm, nb, 0.1, rng, "int64")
        reqs.append((M, A, B))
    C, rb, cb, split = gm_batch_dot3(handle, reqs)
    name, _ = SR[handle]
    for r, (M, A, B) in enumerate(reqs):
        ref = O.dot3(name, M, A, B)
        tol = 1e-12 if "fp64" in name else None
        assert_csr_equal(slice_stacked(C, rb[r], rb[r + 1], cb[r]), ref, fp_tol=tol, what=f"req {r}")
        assert_csr_equal(split[r], ref, fp_tol=tol, what=f"split {r}")


@pytest.mark.parametrize("mode", [0, 2])
def test_batch_saxpy3_bfs(grb, mode):
    """C_r<!V_r> = F_r*A for many frontiers against one graph (served shape)."""
    print(f"[batch] saxpy3 bfs mode={mode}", flush=True)
    G = W.rmat_sym(12, 16, seed=4)
    n = G.nrows
    rng = np.random.default_rng(5)
    reqs = []
    for r in range(30):
        k = int(rng.integers(1, 40))
        src = rng.choice(n, size=k, replace=False).astype(np.int64)
        F = W.Csr(k, n, np.arange(k + 1, dtype=np.int64), src)
        V = rand_csr(k, n, 0.02, rng, "int64")
        reqs.append((V, F))
    C, rb = gm_batch_saxpy3("GxB_ANY_SECONDI_INT64", reqs, G, mask_comp=True, abtype=1,
                            params=(64, 512, 2048, mode))
    for r, (V, F) in enumerate(reqs):
        ref = O.saxpy3("any_secondi_int64", F, G, M=V, mask_comp=True)
        assert_csr_equal(slice_stacked(C, rb[r], rb[r + 1]), ref, what=f"req {r}")


def test_batch_saxpy3_unmasked_fp64(grb):
    rng = np.random.default_rng(6)
    B = power_law_csr(300, 10, rng, values="fp64")
    reqs = [(None, rand_csr(int(rng.integers(1, 50)), 300, 0.03, rng, "fp64")) for _ in range(20)]
    C, rb = gm_batch_saxpy3("GrB_PLUS_TIMES_SEMIRING_FP64", reqs, B)
    for r, (_, A) in enumerate(reqs):
        ref = O.saxpy3("plus_times_fp64", A, B)
        assert_csr_equal(slice_stacked(C, rb[r], rb[r + 1]), ref, fp_tol=1e-12, what=f"req {r}")


def test_batch_dot3_empty_and_tiny_parts(grb):
    """Stacking over many tiny requests, requests with no entries at all and
    one large request: a 4096-element stacking tile then spans dozens of
    parts, empty parts share their base with the next one, and the large
    request fills whole tiles by itself."""
    print("[batch] dot3 empty and tiny parts", flush=True)
    rng = np.random.default_rng(19)
    reqs = []
    for r in range(300):
        nm, nb, k = int(rng.integers(1, 12)), int(rng.integers(1, 12)), int(rng.integers(1, 20))
        dens = 0.0 if r % 3 == 1 else 0.4
        A = rand_csr(nm, k, dens, rng)
        B = rand_csr(nb, k, 0.0 if r % 5 == 2 else 0.4, rng)
        M = rand_csr(nm, nb, 0.0 if r % 7 == 3 else 0.5, rng, "int64")
        reqs.append((M, A, B))
    A = rand_csr(3000, 2000, 0.01, rng)
    B = rand_csr(2500, 2000, 0.01, rng)
    M = rand_csr(3000, 2500, 0.005, rng, "int64")
    reqs.insert(150, (M, A, B))
    C, rb, cb, split = gm_batch_dot3("GxB_PLUS_PAIR_INT64", reqs)
    name, _ = SR["GxB_PLUS_PAIR_INT64"]
    for r, (M, A, B) in enumerate(reqs):
        ref = O.dot3(name, M, A, B)
        assert_csr_equal(slice_stacked(C, rb[r], rb[r + 1], cb[r]), ref, what=f"req {r}")
        assert_csr_equal(split[r], ref, what=f"split {r}")
