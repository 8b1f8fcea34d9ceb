"""Writes tests/golden/fullsize_rmat22.json: the oracle's result of the
triangle count C<L> = L*U' (plus_pair int64) on RMAT scale 22, edge factor
16, a/b/c = 0.57/0.19/0.19, seed 1 (graphblas_amd.workloads.tricount_inputs),
as nnz(C), sum(C) and per-band digests (tests/digest.py) over 64 bands of
rows.  The full C (41.8M entries) is too large to commit; the GPU test
compares its own C band by band with these digests and a band bit-exactly
with the live oracle.  Run from the repository root (about a minute on 8
cores):  python tests/golden/make_fullsize.py
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import pyoracle as O                      # noqa: E402
from graphblas_amd import workloads as W  # noqa: E402
from digest import band_digest            # noqa: E402

NBANDS = 64


def main():
    t0 = time.time()
    _, L, U = W.tricount_inputs(22, 16, seed=1)
    C = O.dot3("plus_pair_int64", L, L, U)
    bands = band_digest(C.p, C.j, C.x, NBANDS)
    out = {"workload": "C<L>=L*U' plus_pair int64, RMAT scale 22, ef 16, a/b/c 0.57/0.19/0.19, "
                       "seed 1, symmetrised, deduplicated, no self-loops, no presort",
           "n": L.nrows, "nnz_L": L.nnz, "nnz_C": int(C.p[-1]), "sum_C": int(C.x.sum()),
           "nbands": NBANDS, "bands": [[nz, str(d)] for nz, d in bands],
           "made_by": "tests/golden/make_fullsize.py (CPU oracle)"}
    path = os.path.join(ROOT, "tests", "golden", "fullsize_rmat22.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(f"wrote {path}: nnz(C)={out['nnz_C']} sum={out['sum_C']} in {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
