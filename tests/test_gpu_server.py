"""Served frontier expansion (gm_server_*) on the GPU against the oracle.

Each request C<!V> = F*A (ANY_SECONDI_INT64, C(r,i) = src[r]) must give the
oracle's result bit-exactly whether it ran alone or stacked with others in a
batch, with one or several worker threads.
"""
import numpy as np
import pytest

import pyoracle as O
from graphblas_amd import capi
from graphblas_amd import workloads as W
from helpers import DevCsr, assert_csr_equal, rand_csr

pytestmark = pytest.mark.gpu

ANY_SECONDI = (4, 4, 8)


def _requests(n, nreq, seed, kmax=64, vis_frac=0.5):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nreq):
        k = int(rng.integers(1, kmax + 1))
        src = rng.choice(n, size=k, replace=False).astype(np.int64)
        V = None
        if rng.random() < vis_frac:
            V = rand_csr(k, n, 0.01, rng)
        out.append((src, V))
    return out


def _check(srv, G, reqs, ids):
    results = []
    for (src, V), rid in zip(reqs, ids):
        res = srv.wait(rid)
        assert res.status == 0, capi.last_error()
        p, j, x = srv.fetch(rid, res)
        k = len(src)
        F = W.Csr(k, G.nrows, np.arange(k + 1, dtype=np.int64), src)
        Vm = V if V is not None else F
        ref = O.saxpy3("any_secondi_int64", F, G, M=Vm, mask_comp=True)
        assert_csr_equal(W.Csr(k, G.ncols, p, j, x), ref, what=f"request {rid}")
        results.append(res)
        srv.release(rid)
    return results


@pytest.mark.parametrize("workers,budget", [(1, 1 << 24), (2, 1 << 24), (2, 20000), (3, 1)])
def test_server_matches_oracle(grb, workers, budget):
    print(f"[server] workers={workers} budget={budget}", flush=True)
    G = W.rmat_sym(12, 16, seed=4)
    dG = DevCsr(G, 1)
    srv = capi.Server(dG.view, ANY_SECONDI, workers=workers, flop_budget=budget)
    try:
        reqs = _requests(G.nrows, 80, seed=workers * 7 + budget % 5)
        ids = [srv.submit(src, None if V is None else (V.p, V.j)) for src, V in reqs]
        res = _check(srv, G, reqs, ids)
        st = srv.stats()
        assert st["completed"] == len(reqs)
        assert st["pending"] == 0
        assert sum(1 for r in res if r.latency_us >= r.queue_us >= 0) == len(reqs)
        if budget == 1:
            assert all(r.batch_size == 1 for r in res)   # every request over budget runs alone
    finally:
        srv.close()
        dG.free()


def test_server_alone_equals_mixed(grb):
    """A request's rows are the same whether it ran alone or in a batch."""
    G = W.rmat_sym(11, 16, seed=9)
    dG = DevCsr(G, 1)
    reqs = _requests(G.nrows, 40, seed=3)
    try:
        alone = capi.Server(dG.view, ANY_SECONDI, workers=1, max_batch_reqs=1)
        mixed = capi.Server(dG.view, ANY_SECONDI, workers=2, max_batch_reqs=64)
        ia = [alone.submit(s, None if V is None else (V.p, V.j)) for s, V in reqs]
        im = [mixed.submit(s, None if V is None else (V.p, V.j)) for s, V in reqs]
        batched = 0
        for a, m in zip(ia, im):
            ra, rm = alone.wait(a), mixed.wait(m)
            assert ra.batch_size == 1
            batched += rm.batch_size > 1
            pa, ja, xa = alone.fetch(a, ra)
            pm, jm, xm = mixed.fetch(m, rm)
            assert np.array_equal(pa, pm) and np.array_equal(ja, jm) and np.array_equal(xa, xm)
        alone.close()
        mixed.close()
    finally:
        dG.free()


def test_server_plus_times_fp64_and_bad_source(grb):
    """Another semiring (PLUS_TIMES fp64 on a weighted graph: C(r,:) = A(src,:)
    outside V) and the host checks of submit."""
    rng = np.random.default_rng(1)
    G = W.rmat_sym(10, 8, seed=2)
    G.x = rng.random(G.nnz)
    dG = DevCsr(G, 11)
    srv = capi.Server(dG.view, (1, 1, 11), workers=2)
    try:
        with pytest.raises(capi.GrBError):
            srv.submit(np.array([G.nrows], dtype=np.int64))
        src = np.array([0, 5, 17], dtype=np.int64)
        rid = srv.submit(src)
        res = srv.wait(rid)
        p, j, x = srv.fetch(rid, res)
        F = W.Csr(3, G.nrows, np.arange(4, dtype=np.int64), src, np.ones(3))
        ref = O.saxpy3("plus_times_fp64", F, G, M=F, mask_comp=True)
        assert_csr_equal(W.Csr(3, G.ncols, p, j, x), ref, fp_tol=0.0, what="plus_times")
    finally:
        srv.close()
        dG.free()


def test_server_drive_poisson(grb):
    """The C++ load generator submits on a Poisson schedule; every request
    completes, in its own batch or in a mixed one."""
    G = W.rmat_sym(11, 16, seed=9)
    dG = DevCsr(G, 1)
    srv = capi.Server(dG.view, ANY_SECONDI, workers=2)
    try:
        reqs = [s for s, _ in _requests(G.nrows, 200, seed=8, vis_frac=0.0)]
        arrive = np.cumsum(np.random.default_rng(2).exponential(50.0, len(reqs)))
        res, wall = W.server_drive(srv, reqs, arrive)
        assert all(r.status == 0 for r in res)
        assert wall * 1e6 >= arrive[-1]
        assert sum(r.k for r in res) == sum(len(s) for s in reqs)
        assert srv.stats()["completed"] == len(reqs)
    finally:
        srv.close()
        dG.free()


def _oracle_bfs(G, src, max_steps):
    """Levels of C<!V> = F*A (any_secondi int64) with V += C, F = C until the
    frontier is empty or max_steps levels ran; returns (last C, V with parents,
    levels run).  The oracle's saxpy3 witness is the last writer in k order,
    as on the device."""
    k = len(src)
    F = W.Csr(k, G.nrows, np.arange(k + 1, dtype=np.int64), np.asarray(src, dtype=np.int64))
    parents = [{int(s): int(s)} for s in src]
    V = F
    C = None
    steps = 0
    while True:
        C = O.saxpy3("any_secondi_int64", F, G, M=V, mask_comp=True)
        steps += 1
        for r in range(k):
            for q in range(C.p[r], C.p[r + 1]):
                parents[r][int(C.j[q])] = int(C.x[q])
        vp = np.zeros(k + 1, dtype=np.int64)
        vj, vx = [], []
        for r in range(k):
            keys = sorted(parents[r])
            vj += keys
            vx += [parents[r][c] for c in keys]
            vp[r + 1] = len(vj)
        V = W.Csr(k, G.nrows, vp, np.array(vj, dtype=np.int64), np.array(vx, dtype=np.int64))
        if C.nnz == 0 or (max_steps > 0 and steps >= max_steps):
            return C, V, steps
        F = C


@pytest.mark.parametrize("workers", [1, 2])
def test_server_multilevel_bfs_matches_oracle(grb, workers):
    """BFS requests of several levels (and until the frontier is empty) with
    their frontier and visited set kept in HBM between levels, mixed with
    one-level requests: the last level and the visited set with parents
    equal the oracle's level-by-level loop."""
    print(f"[server] multilevel bfs workers={workers}", flush=True)
    G = W.rmat_sym(12, 8, seed=5)
    dG = DevCsr(G, 1)
    rng = np.random.default_rng(21 + workers)
    srv = capi.Server(dG.view, ANY_SECONDI, workers=workers, max_batch_reqs=16)
    try:
        reqs = []
        for t in range(24):
            k = int(rng.integers(1, 9))
            src = np.sort(rng.choice(G.nrows, size=k, replace=False)).astype(np.int64)
            steps = [1, 2, 3, 0][t % 4]
            reqs.append((src, steps, srv.submit(src, max_steps=steps, keep_visited=True)))
        for src, steps, rid in reqs:
            res = srv.wait(rid)
            assert res.status == 0, capi.last_error()
            C, V, nst = _oracle_bfs(G, src, steps)
            assert res.steps == nst, (steps, res.steps, nst)
            p, j, x = srv.fetch(rid, res)
            assert_csr_equal(W.Csr(len(src), G.ncols, p, j, x), C, what=f"last level of {rid}")
            vp, vj, vx = srv.fetch_visited(rid, res)
            assert_csr_equal(W.Csr(len(src), G.ncols, vp, vj, vx), V, what=f"visited set of {rid}")
            srv.release(rid)
        st = srv.stats()
        assert st["levels"] >= sum(1 for _ in reqs)
    finally:
        srv.close()
        dG.free()


def test_server_out_of_memory_splits_and_requeues(grb):
    """An allocation forced to fail in the middle of a batch (the count-down
    injector) sends the batch back to the queue with a halved cap; every
    request still equals the oracle.  No worker threads, so the failure lands
    in a known batch."""
    G = W.rmat_sym(11, 16, seed=9)
    dG = DevCsr(G, 1)
    L = capi.lib()
    reqs = [s for s, _ in _requests(G.nrows, 40, seed=13, vis_frac=0.0)]
    srv = capi.Server(dG.view, ANY_SECONDI, workers=0, max_batch_reqs=64)
    try:
        ids = [srv.submit(s) for s in reqs]
        L.gm_debug_fail_after(3)         # the 4th allocation of the first batch fails
        try:
            first = srv.step()
        finally:
            L.gm_debug_fail_after(-1)
        assert first == 40               # admitted whole, then queued again
        assert srv.stats()["oom_splits"] == 1
        sizes = []
        while True:
            n = srv.step()
            if n == 0:
                break
            sizes.append(n)
        assert sizes[0] == 20            # the next admission takes half
        assert sum(sizes) == 40
        for src, rid in zip(reqs, ids):
            res = srv.wait(rid)
            assert res.status == 0
            p, j, x = srv.fetch(rid, res)
            k = len(src)
            F = W.Csr(k, G.nrows, np.arange(k + 1, dtype=np.int64), src)
            ref = O.saxpy3("any_secondi_int64", F, G, M=F, mask_comp=True)
            assert_csr_equal(W.Csr(k, G.ncols, p, j, x), ref, what=f"request {rid}")
            srv.release(rid)
    finally:
        srv.close()
        dG.free()


def test_server_rejects_bad_visited_sets(grb):
    G = W.rmat_sym(8, 4, seed=1)
    dG = DevCsr(G, 1)
    srv = capi.Server(dG.view, ANY_SECONDI, workers=1)
    try:
        src = np.array([1, 2], dtype=np.int64)
        bad = [
            (np.array([1, 1, 2]), np.array([3, 4])),           # vis_p[0] != 0
            (np.array([0, 2, 1]), np.array([3, 4])),           # not monotone
            (np.array([0, 1, 2]), np.array([3, G.nrows])),     # column out of range
            (np.array([0, 2, 2]), np.array([5, 4])),           # row not sorted
            (np.array([0, 2, 2]), np.array([4, 4])),           # duplicate column
        ]
        for vp, vj in bad:
            with pytest.raises(capi.GrBError):
                srv.submit(src, (vp, vj))
        rid = srv.submit(src, (np.array([0, 1, 2]), np.array([1, 2])))
        assert srv.wait(rid).status == 0
    finally:
        srv.close()
        dG.free()
