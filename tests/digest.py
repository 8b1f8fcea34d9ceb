"""Order-independent digests of CSR results, for comparing results too large
to keep as fixtures: per band of rows, the entry count and the wrapping sum
of a 64-bit mix of (row, column, value bits) over the band's entries.  One
changed, missing or extra entry changes its band's digest (up to 2^-64)."""
from __future__ import annotations

import numpy as np

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)
_M4 = np.uint64(0xD6E8FEB86659FD93)


def entry_mix(i: np.ndarray, j: np.ndarray, xbits: np.ndarray) -> np.ndarray:
    """splitmix64-style finaliser of one entry's (i, j, value bits)."""
    with np.errstate(over="ignore"):
        h = (i.astype(np.uint64) * _M1) ^ ((j.astype(np.uint64) + np.uint64(0x632BE59BD9B4E019)) * _M2)
        h ^= xbits.astype(np.uint64) * _M3
        h ^= h >> np.uint64(31)
        h *= _M4
        h ^= h >> np.uint64(32)
    return h


def value_bits(x: np.ndarray) -> np.ndarray:
    """The value array's bits as uint64 (sign-extended for signed integers)."""
    x = np.asarray(x)
    if x.dtype == np.float64:
        return x.view(np.uint64)
    return x.astype(np.int64).view(np.uint64)


def band_digest(p: np.ndarray, j: np.ndarray, x: np.ndarray, nbands: int, row0: int = 0,
                nrows_total: int = None):
    """[(nnz, digest)] for nbands equal bands of rows of a CSR (rows sorted).
    row0 offsets the row ids (for a slice of a larger matrix)."""
    n = len(p) - 1 if nrows_total is None else nrows_total
    out = []
    for b in range(nbands):
        r0, r1 = (n * b) // nbands - row0, (n * (b + 1)) // nbands - row0
        last = len(p) - 1
        r0, r1 = min(max(r0, 0), last), min(max(r1, 0), last)
        lo, hi = int(p[r0]), int(p[r1])
        if hi <= lo:
            out.append((0, 0))
            continue
        rows = np.repeat(np.arange(r0, r1, dtype=np.int64) + row0, np.diff(p[r0:r1 + 1]))
        h = entry_mix(rows, j[lo:hi], value_bits(x[lo:hi]))
        with np.errstate(over="ignore"):
            d = int(np.add.reduce(h, dtype=np.uint64))
        out.append((hi - lo, d))
    return out
