"""Python wrapper of the CPU oracle (oracle/gm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / CPU baseline.  The product
never imports it.

The oracle restates the reference's dot3 (Source/mxm/template/
GB_AxB_dot3_template.c:111-165 + GB_AxB_dot_cij.c:516-652) and saxpy3
(coarse-task accumulation of GB_AxB_saxpy3_coarseHash_phase5.c:22-93)
methods.  Function headers in gm_oracle.c / gm_oracle_template.c cite the
reference lines each part follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

# semiring name -> (code, ztype numpy dtype, value-type of A/B or None if pattern)
SEMIRINGS = {
    "plus_pair_int64": (0, np.int64, None),
    "plus_pair_int32": (1, np.int32, None),
    "any_pair_bool": (2, np.uint8, None),
    "plus_times_fp64": (3, np.float64, np.float64),
    "min_plus_int32": (4, np.int32, np.int32),
    "any_secondi_int64": (5, np.int64, None),
    "plus_times_int64": (6, np.int64, np.int64),
    "min_plus_int64": (7, np.int64, np.int64),
    "max_plus_int32": (8, np.int32, np.int32),
    "min_plus_fp64": (9, np.float64, np.float64),
}


class _OrCsr(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int64), ("ncols", ctypes.c_int64),
                ("p", ctypes.c_void_p), ("j", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("iso", ctypes.c_int32), ("xsize", ctypes.c_int32)]


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        P = ctypes.c_void_p
        L.oracle_dot3.restype = ctypes.c_int64
        L.oracle_dot3.argtypes = [ctypes.c_int, P, ctypes.c_int, P, P, P, P, P]
        L.oracle_saxpy3.restype = ctypes.c_int64
        L.oracle_saxpy3.argtypes = [ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P, P, P, P, P,
                                    ctypes.c_int]
        L.oracle_sum_i64.restype = ctypes.c_int64
        L.oracle_sum_i64.argtypes = [P, ctypes.c_int64]
        L.oracle_neumaier_sum.restype = ctypes.c_double
        L.oracle_neumaier_sum.argtypes = [P, ctypes.c_int64]
        L.oracle_saxpy3_plus_times_fp64_dd.restype = ctypes.c_int64
        L.oracle_saxpy3_plus_times_fp64_dd.argtypes = [P, P, P, P, P]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _csr(m, keep):
    """Build the C struct; `keep` collects arrays that must outlive the call."""
    p = np.ascontiguousarray(m.p, dtype=np.int64)
    j = np.ascontiguousarray(m.j, dtype=np.int64)
    x = None if m.x is None else np.ascontiguousarray(m.x)
    keep += [p, j, x]
    s = _OrCsr(m.nrows, m.ncols, p.ctypes.data, j.ctypes.data,
               None if x is None else x.ctypes.data, 1 if m.iso else 0,
               0 if x is None else x.dtype.itemsize)
    keep.append(s)
    return ctypes.byref(s)


def dot3(semiring: str, M, A, B, mask_struct=True):
    """C<M> = A*B' (CSR rows of A and B); returns (Cp, Ci, Cx) compact CSR."""
    from graphblas_amd.workloads import Csr
    code, zt, _ = SEMIRINGS[semiring]
    keep = []
    mnz = M.nnz
    Cp = np.empty(M.nrows + 1, dtype=np.int64)
    Ci = np.empty(max(mnz, 1), dtype=np.int64)
    Cx = np.empty(max(mnz, 1), dtype=zt)
    nz = lib().oracle_dot3(code, _csr(M, keep), 1 if mask_struct else 0, _csr(A, keep),
                           _csr(B, keep), _ptr(Cp), _ptr(Ci), _ptr(Cx))
    if nz < 0:
        raise RuntimeError(f"oracle_dot3 failed ({nz})")
    return Csr(M.nrows, M.ncols, Cp, Ci[:nz].copy(), Cx[:nz].copy())


def saxpy3(semiring: str, A, B, M=None, mask_comp=False, mask_struct=True):
    """C<M>/C<!M>/C = A*B by saxpy3; returns compact CSR with sorted rows."""
    from graphblas_amd.workloads import Csr
    code, zt, _ = SEMIRINGS[semiring]
    keep = []
    m = A.nrows
    Cp = np.zeros(m + 1, dtype=np.int64)
    Mref = None if M is None else _csr(M, keep)
    Aref, Bref = _csr(A, keep), _csr(B, keep)
    tot = lib().oracle_saxpy3(code, Mref, int(mask_comp), int(mask_struct), Aref, Bref,
                              _ptr(Cp), None, None, 0)
    if tot < 0:
        raise RuntimeError(f"oracle_saxpy3 failed ({tot})")
    counts = Cp[:m].copy()
    Cp[0] = 0
    np.cumsum(counts, out=Cp[1:])
    Ci = np.empty(max(tot, 1), dtype=np.int64)
    Cx = np.empty(max(tot, 1), dtype=zt)
    lib().oracle_saxpy3(code, Mref, int(mask_comp), int(mask_struct), Aref, Bref,
                        _ptr(Cp), _ptr(Ci), _ptr(Cx), 1)
    return Csr(m, B.ncols, Cp, Ci[:tot].copy(), Cx[:tot].copy())


def saxpy3_plus_times_fp64_dd(A, B, C):
    """Compensated (double-double) values of C = A*B under plus_times fp64
    on the pattern of C (sorted rows, e.g. from saxpy3())."""
    keep = []
    Cp = np.ascontiguousarray(C.p, dtype=np.int64)
    Ci = np.ascontiguousarray(C.j, dtype=np.int64)
    Cx = np.empty(max(len(Ci), 1), dtype=np.float64)
    rc = lib().oracle_saxpy3_plus_times_fp64_dd(_csr(A, keep), _csr(B, keep), _ptr(Cp), _ptr(Ci),
                                                 _ptr(Cx))
    if rc != 0:
        raise RuntimeError("oracle_saxpy3_plus_times_fp64_dd: pattern of C misses a product")
    return Cx[:len(Ci)]


def neumaier_sum(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().oracle_neumaier_sum(_ptr(x), x.size)


def set_threads(n: int):
    lib().oracle_set_threads(n)


def num_threads() -> int:
    return lib().oracle_num_threads()


# ---------------------------------------------------------------------------
# steps around the product (small inputs: plain Python over the entries)
# ---------------------------------------------------------------------------

def transpose(A):
    """CSR of A' with sorted rows: the column-major order of A's entries
    (the reference's GB_transpose, Source/transpose/GB_transpose.c, builds
    the same matrix by bucket sort)."""
    from graphblas_amd.workloads import Csr
    rows = np.repeat(np.arange(A.nrows, dtype=np.int64), np.diff(A.p))
    order = np.lexsort((rows, A.j))          # by column, then by row
    Tj = rows[order]
    counts = np.bincount(A.j, minlength=A.ncols) if A.nnz else np.zeros(A.ncols, dtype=np.int64)
    Tp = np.zeros(A.ncols + 1, dtype=np.int64)
    Tp[1:] = np.cumsum(counts)
    Tx = None
    if A.x is not None:
        Tx = A.x[order] if not getattr(A, "iso", False) else A.x
    return Csr(A.ncols, A.nrows, Tp, Tj, Tx)


def _accum_op(op: int, c, t, dtype):
    """z = accum(c, t) for the reference's built-in opcodes (PLUS=1 .. LAND=11;
    the semantics of Source/builtin/include/GB_binop*.h for one type).
    Integer arithmetic wraps; bool PLUS/MAX are OR, TIMES/MIN are AND, MINUS is XOR."""
    if dtype == np.uint8:
        a, b = bool(c), bool(t)
        r = {1: a or b, 2: a != b, 3: a and b, 4: a and b, 5: a or b, 6: a, 7: b, 8: b, 9: True,
             10: a or b, 11: a and b}[op]
        return np.uint8(r)
    if dtype == np.float64:
        a, b = float(c), float(t)

        def fmin(x, y):
            return y if x != x else (x if y != y else min(x, y))

        def fmax(x, y):
            return y if x != x else (x if y != y else max(x, y))
        r = {1: lambda: a + b, 2: lambda: a - b, 3: lambda: a * b, 4: lambda: fmin(a, b),
             5: lambda: fmax(a, b), 6: lambda: a, 7: lambda: b, 8: lambda: b, 9: lambda: 1.0,
             10: lambda: float(a != 0 or b != 0), 11: lambda: float(a != 0 and b != 0)}[op]()
        return np.float64(r)
    bits = np.dtype(dtype).itemsize * 8
    a, b = int(c), int(t)
    r = {1: a + b, 2: a - b, 3: a * b, 4: min(a, b), 5: max(a, b), 6: a, 7: b, 8: b, 9: 1,
         10: int(a != 0 or b != 0), 11: int(a != 0 and b != 0)}[op]
    r &= (1 << bits) - 1
    if r >= 1 << (bits - 1):
        r -= 1 << bits
    return dtype(r)


def accum_mask(C, M, T, mask_comp=False, mask_struct=True, replace=False, op=0, dtype=np.int64):
    """C<M, replace> = accum(C, T), the reference's GB_accum_mask
    (Source/mask/GB_accum_mask.c) with the GraphBLAS spec's rules: Z = T
    (no accum) or the union of C and T with accum on the intersection; inside
    the mask C takes Z, outside C keeps its entries unless replace.  C, M may
    be None."""
    from graphblas_amd.workloads import Csr

    def entries(A, iso=False):
        d = {}
        if A is None:
            return d
        for i in range(A.nrows):
            for p in range(A.p[i], A.p[i + 1]):
                d[(i, int(A.j[p]))] = None if A.x is None else A.x[0 if iso else p]
        return d
    ce = entries(C, getattr(C, "iso", False) if C is not None else False)
    te = entries(T, getattr(T, "iso", False))
    me = entries(M, getattr(M, "iso", False)) if M is not None else None
    out = {}
    for key in set(ce) | set(te):
        if me is None:
            allow = True
        else:
            mij = key in me and (mask_struct or me[key] is None or me[key] != 0)
            allow = (not mij) if mask_comp else mij
        cin, tin = key in ce, key in te
        if allow:
            if op:
                if cin and tin:
                    out[key] = _accum_op(op, ce[key], te[key], dtype)
                elif cin or tin:
                    out[key] = ce[key] if cin else te[key]
            elif tin:
                out[key] = te[key]
        elif cin and not replace:
            out[key] = ce[key]
    m = T.nrows
    keys = sorted(out)
    p = np.zeros(m + 1, dtype=np.int64)
    for (i, _) in keys:
        p[i + 1] += 1
    p = np.cumsum(p)
    j = np.array([k[1] for k in keys], dtype=np.int64)
    x = np.array([out[k] for k in keys], dtype=dtype)
    return Csr(m, T.ncols, p, j, x)
