"""Blobs in the reference's serialized-matrix layout, built independently of
the library for its tests (Source/serialize/GB_serialize.c:300-390 and
GB_serialize.h:106-140 of the reference; the block decoding
GB_deserialize_from_blob.c), and an LZ4 block writer that emits the two kinds
of sequences the LZ4 block format has (literals, and literals followed by a
back-reference)."""
import struct

import numpy as np

NONE, LZ4, ZSTD = -1, 1000, 3000
TYPE_NAME = {1: "GrB_BOOL", 6: "GrB_INT32", 8: "GrB_INT64", 11: "GrB_FP64"}


def lz4_literals(data: bytes) -> bytes:
    """One LZ4 sequence holding all of data as literals (a valid block)."""
    n = len(data)
    out = bytearray()
    if n < 15:
        out.append(n << 4)
    else:
        out.append(15 << 4)
        r = n - 15
        while r >= 255:
            out.append(255)
            r -= 255
        out.append(r)
    out += data
    return bytes(out)


def lz4_repeat(unit: bytes, total: int) -> bytes:
    """An LZ4 block decoding to `unit` repeated to `total` bytes: the unit as
    literals, one overlapping back-reference of distance len(unit), then the
    last five bytes as literals."""
    data = (unit * (total // len(unit) + 2))[:total]
    u = len(unit)
    assert total >= u + 4 + 5
    mlen = total - u - 5
    out = bytearray()
    ml = mlen - 4
    out.append((min(u, 15) << 4) | min(ml, 15))
    if u >= 15:
        r = u - 15
        while r >= 255:
            out.append(255)
            r -= 255
        out.append(r)
    out += unit
    out += struct.pack("<H", u)
    if ml >= 15:
        r = ml - 15
        while r >= 255:
            out.append(255)
            r -= 255
        out.append(r)
    out += lz4_literals(data[total - 5:])
    return bytes(out), data


def partition(n, k, nb):
    """Byte range of block k of nb equal parts of n bytes (GB_PARTITION)."""
    a = 0 if k == 0 else int((k * float(n)) / nb)
    b = n if k == nb - 1 else int(((k + 1) * float(n)) / nb)
    return a, b


def blob(vlen, vdim, typecode, typesize, arrays, sparsity=2, iso=False, csc=False, nvec=None,
         nvals=0, methods=None, blocks=None, name=b"", version=9003001):
    """arrays: dict of raw bytes for 'p', 'h', 'b', 'i', 'x'; blocks: dict of
    lists of already-encoded blocks (default: one uncompressed block)."""
    keys = ["p", "h", "b", "i", "x"]
    methods = methods or {}
    blocks = blocks or {}
    raw = {k: arrays.get(k, b"") for k in keys}
    enc = {}
    for k in keys:
        if k in blocks:
            enc[k] = blocks[k]
        else:
            enc[k] = [raw[k]] if raw[k] else []
    hdr_rest = bytearray()
    hdr_rest += struct.pack("<i", version)
    nvec = vdim if nvec is None else nvec
    nonempty = 0
    hdr_rest += struct.pack("<qqqqqq", vlen, vdim, nvec, nonempty, nvals, typesize)
    hdr_rest += struct.pack("<qqqqq", *[len(raw[k]) for k in keys])
    hdr_rest += struct.pack("<ff", 0.0625, 0.04)
    hdr_rest += struct.pack("<ii", 15, 4 * sparsity + (2 if iso else 0) + (1 if csc else 0))
    for k in keys:
        hdr_rest += struct.pack("<ii", len(enc[k]), methods.get(k, NONE))
    ends = bytearray()
    for k in keys:
        e = 0
        for blk in enc[k]:
            e += len(blk)
            ends += struct.pack("<q", e)
    body = b"".join(b"".join(enc[k]) for k in keys)
    tail = name + b"\0" + TYPE_NAME.get(typecode, "").encode() + b"\0"
    size = 8 + 4 + len(hdr_rest) + len(ends) + len(body) + len(tail)
    return struct.pack("<Qi", size, typecode) + bytes(hdr_rest) + bytes(ends) + body + tail


def parse(b: bytes):
    """Header fields and uncompressed arrays of a blob (for blobs this test
    suite reads back from the library)."""
    s = 0
    size, code, version = struct.unpack_from("<Qii", b, s); s += 16
    vlen, vdim, nvec, nonempty, nvals, typesize = struct.unpack_from("<qqqqqq", b, s); s += 48
    lens = struct.unpack_from("<qqqqq", b, s); s += 40
    hs, bs = struct.unpack_from("<ff", b, s); s += 8
    sc, sic = struct.unpack_from("<ii", b, s); s += 8
    nbm = struct.unpack_from("<10i", b, s); s += 40
    nbs = nbm[0::2]
    meths = nbm[1::2]
    ends = []
    for nb in nbs:
        ends.append(struct.unpack_from(f"<{nb}q", b, s) if nb else ())
        s += 8 * nb
    arrays = {}
    for k, ln, e in zip(["p", "h", "b", "i", "x"], lens, ends):
        arrays[k] = b[s:s + ln]
        s += e[-1] if e else 0
    rest = b[s:]
    return dict(size=size, code=code, version=version, vlen=vlen, vdim=vdim, nvec=nvec,
                nonempty=nonempty, nvals=nvals, typesize=typesize, lens=lens, sic=sic,
                nblocks=nbs, methods=meths, arrays=arrays, names=rest.split(b"\0")[:2])


def csr_arrays(p, j, x):
    return {"p": np.asarray(p, np.int64).tobytes(), "i": np.asarray(j, np.int64).tobytes(),
            "x": np.asarray(x).tobytes()}
