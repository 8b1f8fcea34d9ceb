"""GrB_Matrix_serialize / deserialize in the reference's blob layout
(Source/serialize/GB_serialize.c, GB_deserialize.c and
GB_deserialize_from_blob.c of the reference): blobs this library writes are
read back by an independent parser (tests/blobs.py) and checked field by
field; blobs built by that module in every sparsity, orientation and block
encoding the reader takes (uncompressed, LZ4 literals, LZ4 back-references,
several blocks per array) are deserialized and compared with the matrix they
encode."""
import ctypes

import numpy as np
import pytest

import blobs as BL
from graphblas_amd import capi

pytestmark = pytest.mark.gpu


def _deserialize(b, tname, nrows, ncols):
    L = capi.lib()
    buf = np.frombuffer(b, dtype=np.uint8).copy()
    out = ctypes.c_void_p()
    rc = L.GrB_Matrix_deserialize(ctypes.byref(out), None, buf.ctypes.data, buf.nbytes)
    assert rc == 0, capi.last_error()
    M = capi.Matrix.__new__(capi.Matrix)
    M.h, M.type_name, M.nrows, M.ncols = out, tname, nrows, ncols
    return M.to_csr()


def _dense_to_csr(D, present):
    p = np.zeros(D.shape[0] + 1, dtype=np.int64)
    p[1:] = np.cumsum(present.sum(axis=1))
    r, c = np.nonzero(present)
    return p, c.astype(np.int64), D[r, c]


def _check(got, p, j, x):
    assert np.array_equal(got.p, p)
    assert np.array_equal(got.j, j)
    assert np.array_equal(np.asarray(got.x), np.asarray(x))


@pytest.fixture
def mat():
    rng = np.random.default_rng(12)
    nrows, ncols = 37, 53
    present = rng.random((nrows, ncols)) < 0.15
    present[5] = False                       # an empty row
    D = rng.integers(-100, 100, size=(nrows, ncols)).astype(np.int64)
    return nrows, ncols, present, D


def test_written_blob_layout(grb, mat):
    """The library's blob: the reference's header, one uncompressed block per
    array, sparse by row, the reference's element type name."""
    nrows, ncols, present, D = mat
    p, j, x = _dense_to_csr(D, present)
    from graphblas_amd.workloads import Csr
    A = capi.Matrix.from_csr(Csr(nrows, ncols, p, j, x), "int64")
    L = capi.lib()
    size = ctypes.c_uint64()
    assert L.GrB_Matrix_serializeSize(ctypes.byref(size), A.h) == 0
    buf = np.zeros(size.value, dtype=np.uint8)
    n = ctypes.c_uint64(size.value)
    assert L.GrB_Matrix_serialize(buf.ctypes.data, ctypes.byref(n), A.h) == 0
    assert n.value == size.value
    h = BL.parse(buf.tobytes())
    assert h["size"] == n.value and h["code"] == 8 and h["typesize"] == 8
    assert (h["vlen"], h["vdim"], h["nvec"], h["nvals"]) == (ncols, nrows, nrows, len(j))
    assert h["nonempty"] == int((np.diff(p) > 0).sum())
    assert h["sic"] == 4 * 2 and h["version"] == 9003001
    assert h["nblocks"] == (1, 0, 0, 1, 1) and all(m == BL.NONE for m in h["methods"])
    assert np.array_equal(np.frombuffer(h["arrays"]["p"], np.int64), p)
    assert np.array_equal(np.frombuffer(h["arrays"]["i"], np.int64), j)
    assert np.array_equal(np.frombuffer(h["arrays"]["x"], np.int64), x)
    assert h["names"] == [b"", b"GrB_INT64"]
    _check(_deserialize(buf.tobytes(), "int64", nrows, ncols), p, j, x)


@pytest.mark.parametrize("form", ["sparse", "hyper", "bitmap", "full", "sparse_csc",
                                  "bitmap_csc", "iso", "lz4_literals", "lz4_blocks"])
def test_reference_layout_blobs(grb, mat, form):
    nrows, ncols, present, D = mat
    if form == "full":
        present = np.ones_like(present)
    p, j, x = _dense_to_csr(D, present)
    nv = len(j)
    if form in ("sparse", "iso", "lz4_literals", "lz4_blocks"):
        xs = np.array([7], np.int64) if form == "iso" else x
        arrs = BL.csr_arrays(p, j, xs)
        kw = dict(nvals=nv, iso=(form == "iso"))
        if form == "lz4_literals":
            kw["methods"] = {k: BL.LZ4 for k in "pix"}
            kw["blocks"] = {k: [BL.lz4_literals(arrs[k])] for k in "pix"}
        if form == "lz4_blocks":
            # the column indices in three blocks (GB_PARTITION ranges), the
            # values in one block of back-references over a repeated pattern
            ib = arrs["i"]
            cuts = [BL.partition(len(ib), k, 3) for k in range(3)]
            kw["methods"] = {"i": BL.LZ4, "x": BL.LZ4}
            blk, data = BL.lz4_repeat(np.array([3, -4], np.int64).tobytes(), len(arrs["x"]))
            arrs["x"] = data
            x = np.frombuffer(data, np.int64)
            kw["blocks"] = {"i": [BL.lz4_literals(ib[a:b]) for a, b in cuts], "x": [blk]}
        b = BL.blob(ncols, nrows, 8, 8, arrs, **kw)
        if form == "iso":
            x = np.full(nv, 7, np.int64)
    elif form == "hyper":
        rows = np.nonzero(np.diff(p))[0]
        ph = np.concatenate([[0], p[rows + 1]]).astype(np.int64)
        arrs = {"p": ph.tobytes(), "h": rows.astype(np.int64).tobytes(),
                "i": j.astype(np.int64).tobytes(), "x": x.tobytes()}
        b = BL.blob(ncols, nrows, 8, 8, arrs, sparsity=1, nvec=len(rows), nvals=nv)
    elif form in ("bitmap", "full"):
        arrs = {"x": D.tobytes()}
        if form == "bitmap":
            arrs["b"] = present.astype(np.int8).tobytes()
        b = BL.blob(ncols, nrows, 8, 8, arrs, sparsity=4 if form == "bitmap" else 8, nvals=nv)
    elif form == "sparse_csc":
        pc, jc, xc = _dense_to_csr(D.T.copy(), present.T.copy())
        b = BL.blob(nrows, ncols, 8, 8, BL.csr_arrays(pc, jc, xc), csc=True, nvals=nv)
    elif form == "bitmap_csc":
        arrs = {"x": D.T.copy().tobytes(), "b": present.T.astype(np.int8).tobytes()}
        b = BL.blob(nrows, ncols, 8, 8, arrs, sparsity=4, csc=True, nvals=nv)
    _check(_deserialize(b, "int64", nrows, ncols), p, j, x)
