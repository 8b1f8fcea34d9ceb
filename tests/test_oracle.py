"""Pins the CPU oracle (oracle/gm_oracle.c) before it is trusted as the checker.

The reference holds no golden vectors for this path (its tests regenerate
random inputs and compare against a dense brute force, Test/GB_spec_mxm.m:
99-114, or MATLAB's A*B at 1e-12, Test/test246.m).  Its library does build
from a scratch copy with its own CMake configure step (which generates
Config/GB_config.h and compiles hundreds of files), but this work does not
run the reference's build system or write stand-ins for its generated
headers, so the reference library is not available as a checker here.  Each
oracle function is checked instead against what the mathematics fixes:

* dense brute force on small non-square inputs (a transposed operand, a
  dropped term or a wrong index changes the result),
* closed-form triangle counts (K_n: C(n,3); cycle C_n: 0; wheel W_n: n-1),
  trace(A^3)/6 on RMAT graphs,
* scipy's sparse product (a library routine) for plus_times,
* Neumaier-compensated sums for the fp64 error bound,
* the committed fixtures under tests/golden/ (made by tests/golden/make_golden.py
  from brute force).
"""
import json
import math
import os

import numpy as np
import scipy.sparse as sp
import pytest

import pyoracle as O
from graphblas_amd import workloads as W
from graphblas_amd.workloads import Csr
from helpers import power_law_csr, rand_csr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def dense(m: Csr, fill=0.0, dtype=np.float64):
    D = np.full((m.nrows, m.ncols), fill, dtype=dtype)
    P = np.zeros((m.nrows, m.ncols), dtype=bool)
    for i in range(m.nrows):
        for p in range(m.p[i], m.p[i + 1]):
            D[i, m.j[p]] = 1 if m.x is None else m.x[p]
            P[i, m.j[p]] = True
    return D, P


def brute_dot3(name, M, A, B, mask_struct=True):
    """C(i,j) for (i,j) in M: (+)_k A(i,k) (x) B(j,k), by triple loop."""
    Ad, Ap = dense(A)
    Bd, Bp = dense(B)
    out = {}
    for i in range(M.nrows):
        for p in range(M.p[i], M.p[i + 1]):
            j = M.j[p]
            if not mask_struct and M.x is not None and M.x[p] == 0:
                continue
            ks = [k for k in range(A.ncols) if Ap[i, k] and Bp[j, k]]
            if not ks:
                continue
            out[(i, j)] = _fold(name, [(Ad[i, k], Bd[j, k], k) for k in ks])
    return out


def brute_saxpy(name, A, B, M=None, comp=False, mask_struct=True):
    Ad, Ap = dense(A)
    Bd, Bp = dense(B)
    Md = Mp = None
    if M is not None:
        Md, Mp = dense(M)
    out = {}
    for i in range(A.nrows):
        for j in range(B.ncols):
            if M is not None:
                mv = Mp[i, j] and (mask_struct or Md[i, j] != 0)
                if mv == comp:
                    continue
            ks = [k for k in range(A.ncols) if Ap[i, k] and Bp[k, j]]
            if not ks:
                continue
            out[(i, j)] = _fold(name, [(Ad[i, k], Bd[k, j], k) for k in ks], saxpy=True)
    return out


def _fold(name, terms, saxpy=False):
    if name.startswith("plus_pair"):
        return len(terms)
    if name == "any_pair_bool":
        return 1
    if name == "any_secondi_int64":
        # dot3 keeps the first match (min k), saxpy3's coarse update keeps
        # the last writer in k order (max k)
        return max(k for _, _, k in terms) if saxpy else min(k for _, _, k in terms)
    if name.startswith("plus_times"):
        if name.endswith("fp64"):
            return math.fsum(a * b for a, b, _ in terms)
        return sum(int(a) * int(b) for a, b, _ in terms)
    if name.startswith("min_plus"):
        return min(a + b for a, b, _ in terms)
    if name.startswith("max_plus"):
        return max(a + b for a, b, _ in terms)
    raise ValueError(name)


def as_dict(C: Csr):
    return {(i, int(C.j[p])): C.x[p] for i in range(C.nrows) for p in range(C.p[i], C.p[i + 1])}


def check_equal(got: dict, ref: dict, fp=False):
    assert set(got) == set(ref), f"pattern differs: {len(set(got) ^ set(ref))} entries"
    for key, v in ref.items():
        if fp:
            assert abs(got[key] - v) <= 1e-13 * max(abs(v), 1e-300), (key, got[key], v)
        else:
            assert int(got[key]) == int(v), (key, got[key], v)


VALS = {"plus_pair_int64": None, "plus_pair_int32": None, "any_pair_bool": None,
        "any_secondi_int64": None, "plus_times_fp64": "fp64", "plus_times_int64": "int64",
        "min_plus_int32": "int32", "min_plus_int64": "int64", "max_plus_int32": "int32",
        "min_plus_fp64": "fp64"}


@pytest.mark.parametrize("name", list(VALS))
def test_dot3_vs_brute_force(name):
    rng = np.random.default_rng(abs(hash(name)) % 2 ** 31)
    A = rand_csr(23, 17, 0.3, rng, VALS[name])
    B = rand_csr(19, 17, 0.3, rng, VALS[name])
    M = rand_csr(23, 19, 0.4, rng, "int64")
    C = O.dot3(name, M, A, B)
    check_equal(as_dict(C), brute_dot3(name, M, A, B), fp=name.endswith("fp64"))


def test_dot3_valued_mask_vs_brute_force():
    rng = np.random.default_rng(5)
    A = rand_csr(20, 15, 0.3, rng, "int32")
    B = rand_csr(18, 15, 0.3, rng, "int32")
    M = rand_csr(20, 18, 0.5, rng, "int64")
    M.x[rng.random(M.nnz) < 0.5] = 0
    C = O.dot3("min_plus_int32", M, A, B, mask_struct=False)
    check_equal(as_dict(C), brute_dot3("min_plus_int32", M, A, B, mask_struct=False))


def test_dot3_unbalanced_lists_binary_trim():
    # one list much longer than the other exercises the binary-trim branch
    rng = np.random.default_rng(9)
    A = rand_csr(10, 400, 0.9, rng, "fp64")
    B = rand_csr(12, 400, 0.02, rng, "fp64")
    M = rand_csr(10, 12, 1.0, rng, "int64")
    C = O.dot3("plus_times_fp64", M, A, B)
    check_equal(as_dict(C), brute_dot3("plus_times_fp64", M, A, B), fp=True)


@pytest.mark.parametrize("name", list(VALS))
def test_saxpy3_vs_brute_force(name):
    rng = np.random.default_rng(abs(hash(name + "s")) % 2 ** 31)
    A = rand_csr(21, 16, 0.25, rng, VALS[name])
    B = rand_csr(16, 19, 0.25, rng, VALS[name])
    C = O.saxpy3(name, A, B)
    check_equal(as_dict(C), brute_saxpy(name, A, B), fp=name.endswith("fp64"))


@pytest.mark.parametrize("comp", [False, True])
@pytest.mark.parametrize("struct", [True, False])
def test_saxpy3_masks_vs_brute_force(comp, struct):
    rng = np.random.default_rng(31 + comp + 2 * struct)
    for name in ["plus_pair_int64", "min_plus_int32", "any_secondi_int64"]:
        A = rand_csr(18, 14, 0.3, rng, VALS[name])
        B = rand_csr(14, 20, 0.3, rng, VALS[name])
        M = rand_csr(18, 20, 0.3, rng, "int64")
        if not struct:
            M.x[rng.random(M.nnz) < 0.5] = 0
        C = O.saxpy3(name, A, B, M=M, mask_comp=comp, mask_struct=struct)
        check_equal(as_dict(C), brute_saxpy(name, A, B, M, comp, struct))


def _graph(n, edges) -> Csr:
    adj = [set() for _ in range(n)]
    for u, v in edges:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    p = np.zeros(n + 1, dtype=np.int64)
    for i in range(n):
        p[i + 1] = p[i] + len(adj[i])
    j = np.concatenate([np.array(sorted(a), dtype=np.int64) for a in adj]) if p[-1] else \
        np.zeros(0, dtype=np.int64)
    return Csr(n, n, p, j)


def _tricount(A: Csr) -> int:
    L, U = W.triangle(A, True), W.triangle(A, False)
    return int(O.dot3("plus_pair_int64", L, L, U).x.sum())


@pytest.mark.parametrize("n", [3, 4, 7, 12, 30])
def test_tricount_closed_forms(n):
    K = _graph(n, [(u, v) for u in range(n) for v in range(u + 1, n)])
    assert _tricount(K) == math.comb(n, 3)
    if n > 3:
        C = _graph(n, [(u, (u + 1) % n) for u in range(n)])
        assert _tricount(C) == 0
        # wheel: hub 0 plus the cycle 1..n-1
        Wh = _graph(n, [(0, u) for u in range(1, n)] +
                    [(u, u + 1 if u + 1 < n else 1) for u in range(1, n)])
        assert _tricount(Wh) == (n - 1 if n - 1 > 3 else 4 if n == 4 else n - 1)


@pytest.mark.parametrize("scale", [6, 8, 10])
def test_tricount_trace_a3(scale):
    A = W.rmat_sym(scale, 8, seed=scale)
    D = np.zeros((A.nrows, A.nrows))
    for i in range(A.nrows):
        D[i, A.j[A.p[i]:A.p[i + 1]]] = 1
    tri = int(round(np.trace(D @ D @ D) / 6))
    assert _tricount(A) == tri


def test_saxpy3_plus_times_vs_scipy_and_fsum():
    A, B = W.er(2000, 2000, 8, seed=3), W.er(2000, 2000, 8, seed=4)
    C = O.saxpy3("plus_times_fp64", A, B)
    R = (A.to_scipy() @ B.to_scipy()).tocsr()
    R.sort_indices()
    assert np.array_equal(R.indptr, C.p) and np.array_equal(R.indices, C.j)
    assert np.max(np.abs(R.data - C.x) / np.abs(R.data)) <= 1e-12
    # error against a compensated sum of the exact products, row 0..199
    Bs = B.to_scipy().tocsr()
    worst = 0.0
    for i in range(200):
        terms = {}
        for pa in range(A.p[i], A.p[i + 1]):
            k = A.j[pa]
            for pb in range(Bs.indptr[k], Bs.indptr[k + 1]):
                terms.setdefault(Bs.indices[pb], []).append(A.x[pa] * Bs.data[pb])
        for q in range(C.p[i], C.p[i + 1]):
            exact = math.fsum(terms[C.j[q]])
            worst = max(worst, abs(C.x[q] - exact) / abs(exact))
    assert worst <= 64 * np.finfo(np.float64).eps


def test_compensated_saxpy3_is_correctly_rounded():
    """The double-double oracle equals the exact sum of the exact products
    (rational arithmetic), rounded to double, within half an ulp; a dropped
    term or a product left in double would fail it on these ill-scaled rows."""
    from fractions import Fraction
    rng = np.random.default_rng(5)
    A = power_law_csr(80, 10, rng, values="fp64")
    B = power_law_csr(80, 10, rng, values="fp64")
    # wide exponent range and mixed signs, so plain double sums round
    A.x = A.x * np.exp2(rng.integers(-30, 30, A.nnz)) * rng.choice([-1.0, 1.0], A.nnz)
    B.x = B.x * np.exp2(rng.integers(-30, 30, B.nnz))
    C = O.saxpy3("plus_times_fp64", A, B)
    dd = O.saxpy3_plus_times_fp64_dd(A, B, C)
    worst, plain_worst = 0.0, 0.0
    for i in range(A.nrows):
        acc = {}
        for pa in range(A.p[i], A.p[i + 1]):
            k = A.j[pa]
            for pb in range(B.p[k], B.p[k + 1]):
                jj = int(B.j[pb])
                acc[jj] = acc.get(jj, Fraction(0)) + Fraction(A.x[pa]) * Fraction(B.x[pb])
        for q in range(C.p[i], C.p[i + 1]):
            exact = acc[int(C.j[q])]
            if exact == 0:
                assert dd[q] == 0.0
                continue
            worst = max(worst, float(abs(Fraction(dd[q]) - exact) / abs(exact)))
            plain_worst = max(plain_worst, float(abs(Fraction(C.x[q]) - exact) / abs(exact)))
    assert worst <= 2.0 ** -53 * (1 + 1e-9), worst
    # the check is not vacuous: the sequential sum is visibly worse here
    assert plain_worst > 4 * worst


def test_saxpy3_ascending_k_order_bits():
    """The oracle's fp64 sum is the sequential ascending-k sum (no FMA)."""
    rng = np.random.default_rng(1)
    A = power_law_csr(60, 6, rng, values="fp64")
    B = power_law_csr(60, 6, rng, values="fp64")
    C = O.saxpy3("plus_times_fp64", A, B)
    for i in range(A.nrows):
        acc = {}
        for pa in range(A.p[i], A.p[i + 1]):
            k = A.j[pa]
            for pb in range(B.p[k], B.p[k + 1]):
                t = A.x[pa] * B.x[pb]
                jj = int(B.j[pb])
                acc[jj] = t if jj not in acc else acc[jj] + t
        for q in range(C.p[i], C.p[i + 1]):
            assert C.x[q] == acc[int(C.j[q])]


def test_golden_fixtures():
    path = os.path.join(GOLDEN, "golden_mxm.json")
    with open(path) as f:
        cases = json.load(f)
    assert len(cases) >= 6
    for case in cases:
        def mk(d, key):
            x = None if d.get("x") is None else np.array(d["x"], dtype=VAL_DT[case["semiring"]])
            if key == "M":
                x = None if d.get("x") is None else np.array(d["x"], dtype=np.int64)
            return Csr(d["nrows"], d["ncols"], np.array(d["p"], dtype=np.int64),
                       np.array(d["j"], dtype=np.int64), x)
        A, B = mk(case["A"], "A"), mk(case["B"], "B")
        M = mk(case["M"], "M") if case.get("M") else None
        if case["method"] == "dot3":
            C = O.dot3(case["semiring"], M, A, B, mask_struct=case["mask_struct"])
        else:
            C = O.saxpy3(case["semiring"], A, B, M=M, mask_comp=case["mask_comp"],
                         mask_struct=case["mask_struct"])
        exp = case["C"]
        assert C.p.tolist() == exp["p"], case["name"]
        assert C.j.tolist() == exp["j"], case["name"]
        if case["semiring"].endswith("fp64"):
            assert np.allclose(C.x, exp["x"], rtol=1e-13, atol=0), case["name"]
        else:
            assert [int(v) for v in C.x] == exp["x"], case["name"]


VAL_DT = {"plus_pair_int64": np.int64, "plus_pair_int32": np.int32, "any_pair_bool": np.uint8,
          "any_secondi_int64": np.int64, "plus_times_fp64": np.float64,
          "plus_times_int64": np.int64, "min_plus_int32": np.int32, "min_plus_int64": np.int64,
          "max_plus_int32": np.int32, "min_plus_fp64": np.float64}


def test_planted_mistake_is_caught():
    """A transposed operand (a plausible oracle mistake) fails the brute force."""
    rng = np.random.default_rng(77)
    A = rand_csr(9, 9, 0.4, rng, "fp64")
    B = rand_csr(9, 9, 0.4, rng, "fp64")
    M = rand_csr(9, 9, 0.6, rng, "int64")
    wrong = as_dict(O.dot3("plus_times_fp64", M, B, A))  # operands swapped
    ref = brute_dot3("plus_times_fp64", M, A, B)
    with pytest.raises(AssertionError):
        check_equal(wrong, ref, fp=True)


# ---------------------------------------------------------------------------
# transpose and accum/mask oracles
# ---------------------------------------------------------------------------

def _tiny(rows, ncols, dtype=np.int64):
    """CSR from a list of {col: value} dicts."""
    p = [0]
    j, x = [], []
    for r in rows:
        for c in sorted(r):
            j.append(c)
            x.append(r[c])
        p.append(len(j))
    return W.Csr(len(rows), ncols, np.array(p, dtype=np.int64), np.array(j, dtype=np.int64),
                 np.array(x, dtype=dtype))


def _dense(A, fill=None):
    out = {}
    for i in range(A.nrows):
        for q in range(A.p[i], A.p[i + 1]):
            out[(i, int(A.j[q]))] = A.x[q].item() if A.x is not None else True
    return out


def test_transpose_oracle_matches_scipy_and_is_an_involution():
    rng = np.random.default_rng(5)
    for (m, n, d) in [(1, 1, 1.0), (7, 13, 0.3), (40, 25, 0.1), (30, 30, 0.0)]:
        sm = sp.random(m, n, density=d, format="csr", random_state=rng)
        sm.sort_indices()
        A = W.Csr(m, n, sm.indptr.astype(np.int64), sm.indices.astype(np.int64), sm.data.copy())
        T = O.transpose(A)
        ref = sm.T.tocsr()
        ref.sort_indices()
        assert np.array_equal(T.p, ref.indptr) and np.array_equal(T.j, ref.indices)
        assert np.array_equal(T.x, ref.data)
        TT = O.transpose(T)
        assert np.array_equal(TT.p, A.p) and np.array_equal(TT.j, A.j) and np.array_equal(TT.x, A.x)


def test_accum_mask_oracle_hand_cases():
    C = _tiny([{0: 1, 2: 3}, {1: 5}], 3)
    T = _tiny([{0: 10, 1: 20}, {2: 30}], 3)
    M = _tiny([{0: 1, 1: 1}, {0: 1, 2: 1}], 3)
    # C<M> += T
    got = _dense(O.accum_mask(C, M, T, op=1))
    assert got == {(0, 0): 11, (0, 1): 20, (0, 2): 3, (1, 1): 5, (1, 2): 30}
    # C<M, replace> = T
    got = _dense(O.accum_mask(C, M, T, replace=True))
    assert got == {(0, 0): 10, (0, 1): 20, (1, 2): 30}
    # C<!M> min= T: only (0,2) and (1,1) are writable, T has nothing there
    got = _dense(O.accum_mask(C, M, T, mask_comp=True, op=4))
    assert got == {(0, 0): 1, (0, 2): 3, (1, 1): 5}
    # no mask, SECOND: the union with T winning on the intersection
    got = _dense(O.accum_mask(C, None, T, op=7))
    assert got == {(0, 0): 10, (0, 1): 20, (0, 2): 3, (1, 1): 5, (1, 2): 30}
    # valued mask: a zero entry does not allow
    Mv = _tiny([{0: 0, 1: 1}, {0: 1, 2: 1}], 3)
    got = _dense(O.accum_mask(C, Mv, T, mask_struct=False, op=1))
    assert got == {(0, 0): 1, (0, 1): 20, (0, 2): 3, (1, 1): 5, (1, 2): 30}
    # int32 wrap-around and fp64 / bool semantics
    Ci = _tiny([{0: 2 ** 31 - 1}], 1, np.int32)
    Ti = _tiny([{0: 1}], 1, np.int32)
    assert _dense(O.accum_mask(Ci, None, Ti, op=1, dtype=np.int32)) == {(0, 0): -2 ** 31}
    Cf = _tiny([{0: 1.5}], 1, np.float64)
    Tf = _tiny([{0: 2.0}], 1, np.float64)
    assert _dense(O.accum_mask(Cf, None, Tf, op=3, dtype=np.float64)) == {(0, 0): 3.0}
    assert _dense(O.accum_mask(Cf, None, Tf, op=2, dtype=np.float64)) == {(0, 0): -0.5}
    Cb = _tiny([{0: 1, 1: 0}], 2, np.uint8)
    Tb = _tiny([{0: 1, 1: 1}], 2, np.uint8)
    assert _dense(O.accum_mask(Cb, None, Tb, op=2, dtype=np.uint8)) == {(0, 0): 0, (0, 1): 1}
    assert _dense(O.accum_mask(Cb, None, Tb, op=3, dtype=np.uint8)) == {(0, 0): 1, (0, 1): 0}
