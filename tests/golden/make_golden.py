"""Writes tests/golden/golden_mxm.json: small GrB_mxm cases with expected C.

Expected outputs come from a dense brute-force triple loop (the method of the
reference's spec oracle, Test/GB_spec_mxm.m:99-114), computed independently
of the C oracle; for ANY_SECONDI the dot3 witness is the smallest common k
(first match of the reference's merge) and the saxpy3 witness the largest k
(last writer of its coarse-task update).  Inputs are seeded random matrices.
Run: python tests/golden/make_golden.py
"""
import json
import math
import os

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_mxm.json")


def rand(nr, nc, dens, rng, vals):
    P = rng.random((nr, nc)) < dens
    p = [0]
    j, x = [], []
    for i in range(nr):
        cols = np.nonzero(P[i])[0]
        j += cols.tolist()
        if vals == "fp64":
            x += [float(v) for v in rng.random(len(cols))]
        elif vals == "int32":
            x += [int(v) for v in rng.integers(1, 101, len(cols))]
        elif vals == "mask":
            x += [int(v) for v in rng.integers(0, 2, len(cols))]
        p.append(len(j))
    return {"nrows": nr, "ncols": nc, "p": p, "j": j, "x": x if vals else None}


def get(m, i, k):
    row = m["j"][m["p"][i]:m["p"][i + 1]]
    if k in row:
        q = m["p"][i] + row.index(k)
        return True, (1 if m["x"] is None else m["x"][q])
    return False, None


def fold(sr, terms, saxpy):
    if sr == "plus_pair_int64":
        return len(terms)
    if sr == "any_secondi_int64":
        ks = [k for _, _, k in terms]
        return max(ks) if saxpy else min(ks)
    if sr == "plus_times_fp64":
        return math.fsum(a * b for a, b, _ in terms)
    if sr == "min_plus_int32":
        return min(a + b for a, b, _ in terms)
    raise ValueError(sr)


def expected(case):
    A, B, M = case["A"], case["B"], case.get("M")
    sr, saxpy = case["semiring"], case["method"] == "saxpy3"
    nr = A["nrows"]
    nc = B["ncols"] if saxpy else B["nrows"]
    inner = A["ncols"]
    p, j, x = [0], [], []
    for i in range(nr):
        for jj in range(nc):
            if M is not None:
                there, mv = get(M, i, jj)
                ok = there and (case["mask_struct"] or mv != 0)
                if ok == case.get("mask_comp", False):
                    continue
            terms = []
            for k in range(inner):
                ea, a = get(A, i, k)
                eb, b = get(B, k, jj) if saxpy else get(B, jj, k)
                if ea and eb:
                    terms.append((a, b, k))
            if terms:
                j.append(jj)
                x.append(fold(sr, terms, saxpy))
        p.append(len(j))
    return {"p": p, "j": j, "x": x}


def main():
    rng = np.random.default_rng(20240812)
    cases = []
    for sr, vals in [("plus_pair_int64", None), ("plus_times_fp64", "fp64"),
                     ("min_plus_int32", "int32"), ("any_secondi_int64", None)]:
        # dot3: C<M> = A*B'
        A = rand(14, 11, 0.35, rng, vals)
        B = rand(12, 11, 0.35, rng, vals)
        M = rand(14, 12, 0.5, rng, "mask")
        for ms in (True, False):
            c = {"name": f"dot3_{sr}_struct{int(ms)}", "method": "dot3", "semiring": sr,
                 "A": A, "B": B, "M": M, "mask_struct": ms, "mask_comp": False}
            c["C"] = expected(c)
            cases.append(c)
        # saxpy3: C = A*B, C<M>, C<!M>
        A = rand(13, 10, 0.3, rng, vals)
        B = rand(10, 15, 0.3, rng, vals)
        M = rand(13, 15, 0.4, rng, "mask")
        for mode in ("none", "mask", "comp"):
            c = {"name": f"saxpy3_{sr}_{mode}", "method": "saxpy3", "semiring": sr, "A": A,
                 "B": B, "M": None if mode == "none" else M, "mask_struct": True,
                 "mask_comp": mode == "comp"}
            c["C"] = expected(c)
            cases.append(c)
    with open(OUT, "w") as f:
        json.dump(cases, f, separators=(",", ":"))
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
