"""CPU-only checks of the drop-in boundary: the library loads and exports every
function the public headers declare, global handles exist, and the host
logic (descriptor decode, semiring decode, method selection, argument
validation, branch predicates) behaves like the reference's.  No call here
touches a GPU.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from graphblas_amd import capi
from graphblas_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("GraphBLAS.h", "gm_kernels.h")]


def declared_functions(path):
    src = open(path).read()
    src = re.sub(r"//.*", "", src)
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                         src, flags=re.M):
        name = m.group(1)
        if name in ("if", "for", "while", "switch", "return", "sizeof", "defined"):
            continue
        line = src[m.start():src.find("\n", m.start())]
        if line.lstrip().startswith("#"):
            continue
        names.add(name)
    return names


@pytest.mark.parametrize("header", HEADERS)
def test_library_exports_declared_functions(header):
    L = capi.lib()
    names = declared_functions(header)
    assert len(names) > 10
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_global_handles_exist():
    for name in list(capi.SEMIRINGS) + ["GrB_DESC_" + d for d in capi.DESCRIPTORS] + \
            list(capi.GRB_TYPE_NAMES.values()) + ["GrB_PLUS_MONOID_INT64", "GrB_LOR",
                                                  "GxB_SECONDI_INT64"]:
        assert capi.handle(name), name


def test_semiring_decode():
    L = capi.lib()
    cases = {
        "GxB_PLUS_PAIR_INT64": (1, 3, 8), "GxB_PLUS_PAIR_INT32": (1, 3, 6),
        "GxB_ANY_PAIR_BOOL": (4, 3, 1), "GxB_LOR_PAIR_BOOL": (4, 3, 1),  # renamed to ANY_PAIR
        "GrB_PLUS_TIMES_SEMIRING_FP64": (1, 1, 11), "GrB_MIN_PLUS_SEMIRING_INT32": (2, 2, 6),
        "GxB_ANY_SECONDI_INT64": (4, 4, 8), "GxB_ANY_FIRSTJ_INT64": (4, 4, 8),
    }
    for name, trip in cases.items():
        s = capi.GmSemiring()
        assert L.GB_semiring_codes(capi.handle(name), ctypes.byref(s))
        assert (s.add_op, s.mult_op, s.ztype) == trip, name


def test_flop_rule_for_masked_saxpy():
    """C<M>=A*B goes to dot3 on B' only when the flop count is well above
    the dot3 work estimate nnz(M)*(mean |A(i,:)| + mean |B(:,j)|) + 16 nnz(B)."""
    pref = capi.lib().GB_mxm_prefer_dot3
    n = 1 << 20
    # Erdos-Renyi 8 per row, mask 8 per row: 2^26 flops vs 2^23*16 + 16*2^23
    assert not pref(8 * 8 * n, 8 * n, 8 * n, n, 8 * n, n)
    # triangle count C<L>=L*L, L with 16 per row and heavy columns: 5e10 flops
    assert pref(50_000_000_000, 16 * n, 16 * n, n, 16 * n, n)
    # boundary: flops exactly 4x the estimate is not enough
    est = 100 * (2.0 + 3.0) + 16 * 300
    assert not pref(int(4 * est), 100, 200, 100, 300, 100)
    assert pref(int(4 * est) + 1, 100, 200, 100, 300, 100)
    # an empty mask or no flops: saxpy3
    assert not pref(10 ** 12, 0, 10, 10, 10, 10)
    assert not pref(0, 10, 10, 10, 10, 10)


def test_dot2_rule():
    """GB_AxB_dot2_control restated: (anz, bnz, anvec, bnvec, avlen)."""
    ctl = capi.lib().GB_AxB_dot2_control
    assert ctl(5000, 4000, 8, 10, 5000)                 # 80 pairs <= 100: C tiny
    assert not ctl(5000, 4000, 100, 100, 5000)          # 10000 pairs > anz + bnz: C large
    assert ctl(2e6, 1e6, 10, 20, 1e4)                   # anz + bnz > 10000 * cnz
    assert not ctl(3000, 3000, 50, 60, 100)             # 3000 pairs, none of the rules
    # A' implicit: row degree < 0.125 and column degree > 1200
    assert ctl(130000, 200000, 100, 300, 2e6)
    assert not ctl(130000, 200000, 100, 300, 1e6)       # row degree 0.13


def test_method_selection():
    sel = capi.lib().GB_mxm_select
    # (mask, comp, a_tran, b_tran) -> 1 dot3, 2 saxpy3 (B' transposed first)
    assert sel(True, False, False, True, 0) == 1       # C<M>=A*B'  (tricount, ST1)
    assert sel(True, True, False, True, 0) == 2        # C<!M>=A*B': transpose B, saxpy3
    assert sel(False, False, False, True, 0) == 2      # C=A*B': transpose B, saxpy3
    assert sel(False, False, False, False, 0) == 2     # C=A*B
    assert sel(True, False, False, False, 0) == 2      # C<M>=A*B
    assert sel(True, True, False, False, 0) == 2       # C<!M>=A*B (BFS, RSC)
    assert sel(True, False, True, False, 0) == 2       # C<M>=A'*B: transpose A, saxpy3
    assert sel(True, False, True, True, 0) == 1        # C<M>=A'*B': transpose A, dot3
    assert sel(True, False, False, False, 7083) == 1   # C<M>=A*B, GxB_AxB_DOT: transpose B, dot3
    assert sel(True, True, False, False, 7083) == 2    # C<!M>: dot3 does not take it


def test_branch_predicates_host_only():
    L = capi.lib()
    s = capi.GmSemiring(1, 3, 8)
    A = capi.GmCsr(10, 5, 1, 1, None, 0, 1)
    B = capi.GmCsr(7, 5, 1, 1, None, 0, 1)
    M = capi.GmCsr(10, 7, 1, 1, None, 0, 8)
    assert L.gm_dot3_branch(ctypes.byref(M), True, ctypes.byref(A), ctypes.byref(B),
                            ctypes.byref(s))
    Mbad = capi.GmCsr(10, 8, 1, 1, None, 0, 8)
    assert not L.gm_dot3_branch(ctypes.byref(Mbad), True, ctypes.byref(A), ctypes.byref(B),
                                ctypes.byref(s))
    bad = capi.GmSemiring(1, 9, 8)
    assert not L.gm_dot3_branch(ctypes.byref(M), True, ctypes.byref(A), ctypes.byref(B),
                                ctypes.byref(bad))
    # valued semiring needs values of the semiring's type
    pt = capi.GmSemiring(1, 1, 11)
    assert not L.gm_dot3_branch(ctypes.byref(M), True, ctypes.byref(A), ctypes.byref(B),
                                ctypes.byref(pt))
    A2 = capi.GmCsr(5, 7, 1, 1, None, 0, 1)
    B2 = capi.GmCsr(7, 9, 1, 1, None, 0, 1)
    assert L.gm_saxpy3_branch(None, False, True, ctypes.byref(A2), ctypes.byref(B2),
                              ctypes.byref(s))
    assert not L.gm_saxpy3_branch(None, False, True, ctypes.byref(B2), ctypes.byref(A2),
                                  ctypes.byref(s))


def _mat(tname, nr, nc):
    return capi.Matrix(tname, nr, nc)


def test_grb_mxm_argument_errors_without_gpu():
    capi.init()
    L = capi.lib()
    A = _mat("int64", 4, 5)
    B = _mat("int64", 6, 5)
    C = _mat("int64", 4, 6)
    M = _mat("int64", 4, 6)
    sr = "GxB_PLUS_PAIR_INT64"
    # dimension mismatch: A*B needs ncols(A) == nrows(B)
    assert capi.mxm(C, None, sr, A, B) == capi.GrB_DIMENSION_MISMATCH
    # NULL semiring
    assert L.GrB_mxm(C.h, None, None, None, A.h, B.h, None) == capi.GrB_NULL_POINTER
    # NULL output
    assert L.GrB_mxm(None, None, None, capi.handle(sr), A.h, B.h, None) == capi.GrB_NULL_POINTER
    # result type must be C's type
    Cf = _mat("fp64", 4, 6)
    assert capi.mxm(Cf, M, sr, A, B, "ST1") == capi.GrB_NOT_IMPLEMENTED
    # an accumulator of another type would need a typecast
    acc = capi.handle("GrB_PLUS_FP64")
    assert L.GrB_mxm(C.h, M.h, acc, capi.handle(sr), A.h, B.h,
                     capi.desc("ST1")) == capi.GrB_NOT_IMPLEMENTED
    # an A' operand reaches the device layer (a transpose there): without a
    # device that fails loudly, with one the empty product succeeds
    At = _mat("int64", 5, 4)
    expect = capi.GrB_SUCCESS if L.gm_device_count() > 0 else capi.GrB_PANIC
    assert capi.mxm(C, M, sr, At, B, "ST0T1") == expect
    # mask dimension
    M2 = _mat("int64", 4, 7)
    assert capi.mxm(C, M2, sr, A, B, "ST1") == capi.GrB_DIMENSION_MISMATCH
    # the error string is attached to C
    s = ctypes.c_char_p()
    L.GrB_Matrix_error(ctypes.byref(s), C.h)
    assert b"dimension" in s.value.lower() or b"mask" in s.value.lower()


def test_descriptor_set_semantics():
    L = capi.lib()
    d = ctypes.c_void_p()
    assert L.GrB_Descriptor_new(ctypes.byref(d)) == 0
    assert L.GrB_Descriptor_set(d, 1, 2) == 0       # MASK, COMP
    assert L.GrB_Descriptor_set(d, 1, 4) == 0       # MASK, STRUCTURE (accumulates)
    assert L.GrB_Descriptor_set(d, 3, 3) == 0       # INP1, TRAN
    assert L.GrB_Descriptor_set(d, 0, 3) == capi.GrB_INVALID_VALUE
    # predefined descriptors are read-only
    assert L.GrB_Descriptor_set(capi.desc("ST1"), 0, 1) == capi.GrB_INVALID_VALUE
    assert L.GrB_Descriptor_free(ctypes.byref(d)) == 0
    assert d.value is None


def test_semiring_new_domain_check():
    L = capi.lib()
    s = ctypes.c_void_p()
    rc = L.GrB_Semiring_new(ctypes.byref(s), capi.handle("GrB_PLUS_MONOID_FP64"),
                            capi.handle("GrB_TIMES_INT64"))
    assert rc == capi.GrB_DOMAIN_MISMATCH
    rc = L.GrB_Semiring_new(ctypes.byref(s), capi.handle("GrB_MIN_MONOID_INT64"),
                            capi.handle("GrB_PLUS_INT64"))
    assert rc == 0
    c = capi.GmSemiring()
    assert L.GB_semiring_codes(s, ctypes.byref(c)) and (c.add_op, c.mult_op, c.ztype) == (2, 2, 8)
    assert L.GrB_Semiring_free(ctypes.byref(s)) == 0
    assert s.value is None


def test_freeing_builtins_keeps_them():
    """Freeing a built-in operator, monoid, semiring or descriptor changes
    nothing and leaves the handle as it was (the reference frees and clears
    only allocated objects, Source/op/GB_Op_free.c,
    Source/semiring/GrB_Semiring_free.c), even when the handle passed is the
    library's own global."""
    L = capi.lib()
    capi.init()
    for fn, name in (("GrB_BinaryOp_free", "GrB_PLUS_INT64"),
                     ("GrB_Monoid_free", "GrB_PLUS_MONOID_INT64"),
                     ("GrB_Semiring_free", "GxB_PLUS_PAIR_INT64"),
                     ("GrB_Descriptor_free", "GrB_DESC_ST1")):
        before = capi.handle(name)
        copy = ctypes.c_void_p(before)
        assert getattr(L, fn)(ctypes.byref(copy)) == 0
        assert copy.value == before, fn
        glob = ctypes.c_void_p.in_dll(L, name)
        assert getattr(L, fn)(ctypes.byref(glob)) == 0
        assert capi.handle(name) == before, fn
    # the semiring and descriptor still work after their globals were freed
    c = capi.GmSemiring()
    assert L.GB_semiring_codes(ctypes.c_void_p(capi.handle("GxB_PLUS_PAIR_INT64")), ctypes.byref(c))
    assert L.GrB_Descriptor_set(capi.desc("ST1"), 0, 1) == capi.GrB_INVALID_VALUE


def test_type_sizes_and_options():
    L = capi.lib()
    capi.init()
    for name, size in (("GrB_BOOL", 1), ("GrB_INT8", 1), ("GrB_UINT16", 2), ("GrB_INT32", 4),
                       ("GrB_UINT32", 4), ("GrB_INT64", 8), ("GrB_FP32", 4), ("GrB_FP64", 8)):
        s1, s2 = ctypes.c_size_t(), ctypes.c_size_t()
        t = ctypes.c_void_p(capi.handle(name))
        assert L.GxB_Type_size(ctypes.byref(s1), t) == 0 and s1.value == size, name
        assert L.GrB_Type_get_size(ctypes.byref(s2), t) == 0 and s2.value == size, name
    assert L.GxB_Type_size(None, ctypes.c_void_p(capi.handle("GrB_FP64"))) == capi.GrB_NULL_POINTER
    A = _mat("int32", 3, 4)
    t = ctypes.c_void_p()
    assert L.GxB_Matrix_type(ctypes.byref(t), A.h) == 0
    assert t.value == capi.handle("GrB_INT32")
    v = ctypes.c_int32()
    GxB_BURBLE = 7019
    assert L.GxB_Global_Option_get_INT32(GxB_BURBLE, ctypes.byref(v)) == 0
    old = v.value
    try:
        assert L.GxB_Global_Option_set_INT32(GxB_BURBLE, 1) == 0
        assert L.GxB_Global_Option_get_INT32(GxB_BURBLE, ctypes.byref(v)) == 0 and v.value == 1
    finally:
        L.GxB_Global_Option_set_INT32(GxB_BURBLE, old)
    assert L.GxB_Global_Option_get_INT32(12345, ctypes.byref(v)) == capi.GrB_INVALID_VALUE


def test_generators_are_well_formed():
    A = W.rmat_sym(9, 8, seed=4)
    n = A.nrows
    assert A.p[0] == 0 and np.all(np.diff(A.p) >= 0)
    for i in range(n):
        row = A.j[A.p[i]:A.p[i + 1]]
        assert np.all(np.diff(row) > 0)         # sorted, no duplicates
        assert not np.any(row == i)              # no self-loops
    S = A.to_scipy()
    assert (S != S.T).nnz == 0                   # symmetric
    E = W.er(1000, 777, 8, seed=3)
    assert np.all(np.diff(E.p) == 8)
    for i in range(0, 1000, 37):
        row = E.j[E.p[i]:E.p[i + 1]]
        assert len(set(row.tolist())) == 8 and row.max() < 777 and np.all(np.diff(row) > 0)
    assert np.all((E.x >= 0) & (E.x < 1))
    # deterministic
    A2 = W.rmat_sym(9, 8, seed=4)
    assert np.array_equal(A.p, A2.p) and np.array_equal(A.j, A2.j)


def test_no_gpu_in_container_reports_cleanly():
    """Without a device, a compute call fails loudly (GrB_PANIC), never falls back."""
    L = capi.lib()
    if L.gm_device_count() > 0:
        pytest.skip("a GPU is present")
    capi.init()
    A = _mat("int64", 4, 4)
    C = _mat("int64", 4, 4)
    rc = capi.mxm(C, None, "GxB_PLUS_PAIR_INT64", A, A)
    assert rc == capi.GrB_PANIC


def test_tuning_file_parse(tmp_path):
    """gm_load_tuning accepts the autotuner's key/value lines and rejects
    unknown keys and missing files (host only, no GPU needed)."""
    good = tmp_path / "t.txt"
    good.write_text("# tuned\ndot3.thread_max 16\ndot3.order 2\nsaxpy3.block_max 1024\n")
    try:
        capi.load_tuning(str(good))
        bad = tmp_path / "b.txt"
        bad.write_text("dot3.nonsense 3\n")
        with pytest.raises(capi.GrBError):
            capi.load_tuning(str(bad))
        with pytest.raises(capi.GrBError):
            capi.load_tuning(str(tmp_path / "missing.txt"))
    finally:
        capi.reset_tuning()


def test_vector_objects_without_gpu():
    """Vector handles, sizes and argument errors need no device."""
    capi.init()
    L = capi.lib()
    v = capi.Vector("int64", 10)
    n = ctypes.c_uint64()
    assert L.GrB_Vector_size(ctypes.byref(n), v.h) == 0 and n.value == 10
    assert v.nvals == 0
    A = _mat("int64", 10, 10)
    assert L.GrB_vxm(None, None, None, capi.handle("GxB_PLUS_PAIR_INT64"), v.h, A.h, None) == \
        capi.GrB_NULL_POINTER
    w = capi.Vector("int64", 7)
    # u' (1 x 10) * A (10 x 10) cannot go into a vector of size 7
    assert capi.vxm(w, None, "GxB_PLUS_PAIR_INT64", v, A) == capi.GrB_DIMENSION_MISMATCH


def test_deserialize_rejects_bad_blobs():
    """Header and structure checks of GrB_Matrix_deserialize run on the host
    (blobs in the reference's layout, built by tests/blobs.py)."""
    import blobs as BL
    capi.init()
    L = capi.lib()
    out = ctypes.c_void_p()

    def des(b, t=None):
        buf = np.frombuffer(b, dtype=np.uint8).copy()
        return L.GrB_Matrix_deserialize(ctypes.byref(out), capi.handle(t) if t else None,
                                        buf.ctypes.data, buf.nbytes)

    # a 2 x 3 int64 matrix with one entry, but its column out of range
    arrs = BL.csr_arrays([0, 1, 1], [5], [7])
    assert des(BL.blob(3, 2, 8, 8, arrs, nvals=1)) == capi.GrB_INVALID_OBJECT
    good = BL.blob(3, 2, 8, 8, BL.csr_arrays([0, 1, 1], [2], [7]), nvals=1)
    # a header shorter than the layout's, a truncated blob, a size field that
    # disagrees with the blob
    assert des(good[:100]) == capi.GrB_INVALID_OBJECT
    assert des(good[:-8]) == capi.GrB_INVALID_OBJECT
    assert des(good + b"\0") == capi.GrB_INVALID_OBJECT
    # a type code outside the reference's, and a type size that disagrees
    assert des(BL.blob(3, 2, 99, 8, BL.csr_arrays([0, 1, 1], [2], [7]), nvals=1)) == \
        capi.GrB_INVALID_OBJECT
    assert des(BL.blob(3, 2, 8, 4, BL.csr_arrays([0, 1, 1], [2], [7]), nvals=1)) == \
        capi.GrB_INVALID_OBJECT
    # a type that differs from the blob's
    assert des(good, "GrB_FP64") == capi.GrB_DOMAIN_MISMATCH
    # row pointers that do not end at nvals, columns out of order
    assert des(BL.blob(3, 2, 8, 8, BL.csr_arrays([0, 1, 2], [2], [7]), nvals=1)) == \
        capi.GrB_INVALID_OBJECT
    assert des(BL.blob(3, 2, 8, 8, BL.csr_arrays([0, 2, 2], [2, 1], [7, 8]), nvals=2)) == \
        capi.GrB_INVALID_OBJECT
    # ZSTD blocks (the reference's default compression) are not decoded
    arrs = BL.csr_arrays([0, 1, 1], [2], [7])
    zs = BL.blob(3, 2, 8, 8, arrs, nvals=1, methods={"p": BL.ZSTD},
                 blocks={"p": [b"\x28\xb5\x2f\xfd" + bytes(20)]})
    assert des(zs) == capi.GrB_NOT_IMPLEMENTED
    # an LZ4 block that decodes to fewer bytes than the array's length
    bad = BL.blob(3, 2, 8, 8, arrs, nvals=1, methods={"i": BL.LZ4},
                  blocks={"i": [BL.lz4_literals(arrs["i"][:4])]})
    assert des(bad) == capi.GrB_INVALID_OBJECT
    # a back-reference before the start of the output
    assert des(BL.blob(3, 2, 8, 8, arrs, nvals=1, methods={"i": BL.LZ4},
                       blocks={"i": [bytes([0x14, 1, 9, 0])]})) == capi.GrB_INVALID_OBJECT
    # an entry count whose byte sizes wrap past 2^64 to the arrays' real
    # 8 bytes (row pointers ending at that count): rejected before any entry
    # is read
    big = (1 << 61) + 1
    assert des(BL.blob(3, 1, 8, 8, BL.csr_arrays([0, big], [2], [7]), nvals=big)) == \
        capi.GrB_INVALID_OBJECT
