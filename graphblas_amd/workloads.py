"""Deterministic synthetic inputs for the GrB_mxm workloads.

Wraps graphblas_amd/gen/gm_gen.c (counter-based splitmix64 draws, so every
graph is a pure function of its parameters).  Matrices are returned as
``Csr`` objects holding numpy arrays (int64 row pointers / column indices).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_GEN_SO = os.path.join(_HERE, "lib", "libgm_gen.so")
_gen = None


def _lib():
    global _gen
    if _gen is None:
        if not os.path.exists(_GEN_SO):
            raise RuntimeError(f"{_GEN_SO} is not built; run __graft_entry__.build()")
        g = ctypes.CDLL(_GEN_SO)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        g.gm_gen_rmat_sym_n.restype = i64
        g.gm_gen_rmat_sym_n.argtypes = [i64, i64, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_uint64, P, P]
        g.gm_gen_rmat_sym.restype = i64
        g.gm_gen_rmat_sym.argtypes = [ctypes.c_int, i64, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_uint64, P, P]
        g.gm_gen_triangle.restype = i64
        g.gm_gen_triangle.argtypes = [i64, P, P, ctypes.c_int, P, P]
        g.gm_gen_sym_weights_i32.restype = None
        g.gm_gen_sym_weights_i32.argtypes = [i64, P, P, ctypes.c_uint64, ctypes.c_int32,
                                             ctypes.c_int32, P]
        g.gm_gen_er.restype = None
        g.gm_gen_er.argtypes = [i64, i64, i64, ctypes.c_uint64, P, P, P]
        g.gm_gen_unif.restype = ctypes.c_double
        g.gm_gen_unif.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        g.gm_gen_rand64.restype = ctypes.c_uint64
        g.gm_gen_rand64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        _gen = g
    return _gen


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Csr:
    """A CSR matrix on the host: p[nrows+1], j[nnz] (int64), x values or None."""

    nrows: int
    ncols: int
    p: np.ndarray
    j: np.ndarray
    x: Optional[np.ndarray] = None
    iso: bool = False

    @property
    def nnz(self) -> int:
        return int(self.p[self.nrows])

    def with_values(self, x: Optional[np.ndarray], iso: bool = False) -> "Csr":
        return Csr(self.nrows, self.ncols, self.p, self.j, x, iso)

    def to_scipy(self, dtype=None):
        import scipy.sparse as sp
        if self.x is None:
            data = np.ones(self.nnz, dtype=dtype or np.float64)
        elif self.iso:
            data = np.full(self.nnz, self.x[0], dtype=dtype or self.x.dtype)
        else:
            data = self.x.astype(dtype) if dtype is not None else self.x
        return sp.csr_matrix((data, self.j, self.p), shape=(self.nrows, self.ncols))

    @staticmethod
    def from_scipy(m, keep_values=True) -> "Csr":
        m = m.tocsr()
        m.sort_indices()
        return Csr(m.shape[0], m.shape[1], m.indptr.astype(np.int64),
                   m.indices.astype(np.int64),
                   np.ascontiguousarray(m.data) if keep_values else None)


def rmat_sym(scale: int, ef: int, a=0.57, b=0.19, c=0.19, seed=1) -> Csr:
    """Symmetrised, deduplicated, loop-free RMAT graph (pattern only)."""
    g = _lib()
    n = 1 << scale
    Ap = np.empty(n + 1, dtype=np.int64)
    Aj = np.empty(2 * ef * n, dtype=np.int64)
    nnz = g.gm_gen_rmat_sym(scale, ef, a, b, c, seed, _ptr(Ap), _ptr(Aj))
    if nnz < 0:
        raise MemoryError("rmat generator out of memory")
    Aj = Aj[:nnz].copy()
    return Csr(n, n, Ap, Aj)


def rmat_sym_n(n: int, ef: int, a=0.57, b=0.19, c=0.19, seed=1) -> Csr:
    """RMAT on any n (scale ceil(log2 n), samples with an endpoint >= n dropped),
    symmetrised, deduplicated, loop-free (pattern only)."""
    g = _lib()
    Ap = np.empty(n + 1, dtype=np.int64)
    Aj = np.empty(max(2 * ef * n, 1), dtype=np.int64)
    nnz = g.gm_gen_rmat_sym_n(n, ef, a, b, c, seed, _ptr(Ap), _ptr(Aj))
    if nnz < 0:
        raise MemoryError("rmat generator out of memory")
    return Csr(n, n, Ap, Aj[:nnz].copy())


def triangle(A: Csr, lower: bool) -> Csr:
    """Strict lower (L) or upper (U) triangle of a matrix with sorted rows."""
    g = _lib()
    Lp = np.empty(A.nrows + 1, dtype=np.int64)
    Lj = np.empty(max(A.nnz, 1), dtype=np.int64)
    nnz = g.gm_gen_triangle(A.nrows, _ptr(A.p), _ptr(A.j), 1 if lower else 0, _ptr(Lp), _ptr(Lj))
    return Csr(A.nrows, A.ncols, Lp, Lj[:nnz].copy())


def sym_weights_i32(A: Csr, seed=7, lo=1, hi=100) -> np.ndarray:
    g = _lib()
    x = np.empty(max(A.nnz, 1), dtype=np.int32)
    g.gm_gen_sym_weights_i32(A.nrows, _ptr(A.p), _ptr(A.j), seed, lo, hi, _ptr(x))
    return x[:A.nnz]


def er(nrows: int, ncols: int, per_row: int, seed=1, values=True) -> Csr:
    """Erdos-Renyi: exactly per_row distinct uniform columns per row."""
    g = _lib()
    Ap = np.empty(nrows + 1, dtype=np.int64)
    Aj = np.empty(nrows * per_row, dtype=np.int64)
    Ax = np.empty(nrows * per_row, dtype=np.float64) if values else None
    g.gm_gen_er(nrows, ncols, per_row, seed, _ptr(Ap), _ptr(Aj), _ptr(Ax))
    return Csr(nrows, ncols, Ap, Aj, Ax)


def tricount_inputs(scale: int, ef: int = 16, seed=1):
    """(A, L, U) for the triangle-count workload C<L> = L*U'."""
    A = rmat_sym(scale, ef, seed=seed)
    return A, triangle(A, True), triangle(A, False)


def unif(seed: int, ctr: int) -> float:
    return _lib().gm_gen_unif(seed, ctr)


def rand64(seed: int, ctr: int) -> int:
    return _lib().gm_gen_rand64(seed, ctr)


def loguniform_int(seed: int, ctr: int, lo: int, hi: int) -> int:
    """Integer log-uniform in [lo, hi]."""
    u = unif(seed, ctr)
    v = int(math.floor(math.exp(math.log(lo) + u * (math.log(hi + 1) - math.log(lo)))))
    return max(lo, min(hi, v))


# ---------------------------------------------------------------------------
# batches of uneven graphs
# ---------------------------------------------------------------------------

BATCH_SEMIRINGS = ("plus_pair_int32", "min_plus_int32")


@dataclass
class GraphSpec:
    n: int
    ef: int
    semiring: str        # one of BATCH_SEMIRINGS
    seed: int            # structure
    wseed: int           # int32 edge weights in [1, 100]


def batch_specs(ngraph: int, seed: int = 1, n_lo=512, n_hi=65536, ef_lo=4, ef_hi=32):
    """ngraph RMAT requests: n log-uniform in [n_lo, n_hi], edge factor
    uniform in [ef_lo, ef_hi], semiring plus_pair int32 or min_plus int32 with
    probability 1/2 each; per-graph seeds from counter draws."""
    out = []
    for g in range(ngraph):
        n = loguniform_int(seed, 4 * g, n_lo, n_hi)
        ef = ef_lo + min(ef_hi - ef_lo, int(unif(seed, 4 * g + 1) * (ef_hi - ef_lo + 1)))
        sr = BATCH_SEMIRINGS[0] if unif(seed, 4 * g + 2) < 0.5 else BATCH_SEMIRINGS[1]
        gs = rand64(seed, 4 * g + 3)
        out.append(GraphSpec(n, ef, sr, gs, gs ^ 0x5DEECE66D))
    return out


def batch_graph_host(spec: GraphSpec) -> Csr:
    """One batch graph built on the host (the device generator's reference)."""
    A = rmat_sym_n(spec.n, spec.ef, seed=spec.seed)
    A.x = sym_weights_i32(A, seed=spec.wseed)
    return A


_gendev = None


def _devlib():
    global _gendev
    if _gendev is None:
        from . import capi
        capi.lib()          # the kernel library first (the generator links it)
        so = os.path.join(_HERE, "lib", "libgm_gendev.so")
        if not os.path.exists(so):
            raise RuntimeError(f"{so} is not built; run __graft_entry__.build()")
        d = ctypes.CDLL(so)
        P = ctypes.c_void_p
        d.gm_gen_rmat_batch_dev.restype = ctypes.c_int
        d.gm_gen_rmat_batch_dev.argtypes = [ctypes.c_int64, P, P, P, P, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                            ctypes.c_int32, P, P, ctypes.POINTER(P),
                                            ctypes.POINTER(P), ctypes.POINTER(P)]
        d.gm_server_drive.restype = ctypes.c_int
        d.gm_server_drive.argtypes = [P, ctypes.c_int64, P, P, P, P, P]
        _gendev = d
    return _gendev


def server_drive(server, reqs, arrive_us=None):
    """Open-loop run of reqs (lists of sources) through a capi.Server: request
    r is submitted arrive_us[r] microseconds after the start (all at once when
    None) by the C++ load generator.  Returns (results, wall seconds)."""
    from . import capi
    d = _devlib()
    rp = np.zeros(len(reqs) + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r) for r in reqs])
    src = np.ascontiguousarray(np.concatenate(reqs) if reqs else np.zeros(1), dtype=np.int64)
    arr = None if arrive_us is None else np.ascontiguousarray(arrive_us, dtype=np.float64)
    out = (capi.GmServerResult * max(len(reqs), 1))()
    wall = ctypes.c_double()
    rc = d.gm_server_drive(server.h, len(reqs), _ptr(rp), _ptr(src), _ptr(arr), out,
                           ctypes.byref(wall))
    if rc != 0:
        raise RuntimeError(f"gm_server_drive failed ({rc}): {capi.last_error()}")
    return [out[r] for r in range(len(reqs))], wall.value * 1e-6


class DeviceBatch:
    """Batch graphs generated in HBM (same graphs as batch_graph_host).

    views[g] is graph g's gm_csr (int32 weights); free() releases the arrays."""

    def __init__(self, specs, a=0.57, b=0.19, c=0.19):
        from . import capi
        d = _devlib()
        ng = len(specs)
        n = np.array([s.n for s in specs], dtype=np.int64)
        ef = np.array([s.ef for s in specs], dtype=np.int64)
        seed = np.array([s.seed for s in specs], dtype=np.uint64)
        wseed = np.array([s.wseed for s in specs], dtype=np.uint64)
        self.pbase = np.zeros(ng + 1, dtype=np.int64)
        self.nzbase = np.zeros(ng + 1, dtype=np.int64)
        P, J, X = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        rc = d.gm_gen_rmat_batch_dev(ng, _ptr(n), _ptr(ef), _ptr(seed), _ptr(wseed), a, b, c,
                                     1, 100, _ptr(self.pbase), _ptr(self.nzbase),
                                     ctypes.byref(P), ctypes.byref(J), ctypes.byref(X))
        if rc != 0:
            raise RuntimeError(f"gm_gen_rmat_batch_dev failed ({rc})")
        self.P, self.J, self.X = P.value, J.value, X.value
        self.specs = specs
        self.views = []
        for g, s in enumerate(specs):
            self.views.append(capi.GmCsr(s.n, s.n, self.P + 8 * int(self.pbase[g]),
                                         self.J + 8 * int(self.nzbase[g]),
                                         self.X + 4 * int(self.nzbase[g]), 0, 6))

    def nnz(self, g: int) -> int:
        return int(self.nzbase[g + 1] - self.nzbase[g])

    def download(self, g: int) -> Csr:
        from . import capi
        L = capi.lib()
        s = self.specs[g]
        nz = self.nnz(g)
        p = np.empty(s.n + 1, dtype=np.int64)
        j = np.empty(max(nz, 1), dtype=np.int64)
        x = np.empty(max(nz, 1), dtype=np.int32)
        L.gm_memcpy_d2h(p.ctypes.data, self.P + 8 * int(self.pbase[g]), p.nbytes)
        if nz:
            L.gm_memcpy_d2h(j.ctypes.data, self.J + 8 * int(self.nzbase[g]), 8 * nz)
            L.gm_memcpy_d2h(x.ctypes.data, self.X + 4 * int(self.nzbase[g]), 4 * nz)
        return Csr(s.n, s.n, p, j[:nz], x[:nz])

    def free(self):
        from . import capi
        L = capi.lib()
        for ptr in (self.P, self.J, self.X):
            if ptr:
                L.gm_free(ptr)
        self.P = self.J = self.X = None
