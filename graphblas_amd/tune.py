"""Autotuner for the bin cutoffs and index thresholds of dot3 and saxpy3.

Coordinate descent over the fields of gm_dot3_params / gm_saxpy3_params on a
sample of a workload: starting from the built-in values, each field in turn
takes every candidate value while the others stay fixed, and the fastest
setting whose results are identical to the built-in setting's is kept; two
sweeps.  The result is written as "key value" lines that gm_load_tuning (and
gm_init, through the environment variable GM_TUNING_FILE) read.

    python -m graphblas_amd.tune --workload batched --out tuning.txt

Workloads: batched (a sample of the 4096-graph C<A>=A*A' batch, one
gm_batch_dot3 per semiring), tricount (C<L>=L*U' on an RMAT graph),
er (C=A*B plus_times fp64 on Erdos-Renyi rows).
"""
from __future__ import annotations

import argparse
import ctypes
import time

import numpy as np

from . import capi
from . import workloads as W

DOT3_SPACE = {
    "thread_max": [4, 8, 16, 32, 64],
    "team16_max": [128, 256, 512, 1024],
    "wave_max": [2048, 4096, 8192],
    "bitmap_min_len": [64, 128, 256, 512, -1],
    "order": [1, 2],
    "row_min": [16, 64, 256, 1024],
}
DOT3_BUILTIN = {"thread_max": 8, "team16_max": 512, "wave_max": 4096, "bitmap_min_len": 128,
                "order": 0, "row_min": -1}
SAXPY3_SPACE = {
    "wave_small_max": [16, 32, 64],
    "wave_large_max": [128, 256, 512],
    "block_max": [512, 1024, 2048],
    "mode": [1, 2],
}
SAXPY3_BUILTIN = {"wave_small_max": 64, "wave_large_max": 512, "block_max": 2048, "mode": 0}


def _digest(L, out):
    """(nvals, sum of indices, bytes of the values) of a result, to compare settings."""
    n = out.nvals
    p = np.empty(out.nrows + 1, dtype=np.int64)
    L.gm_memcpy_d2h(p.ctypes.data, out.p, p.nbytes)
    j = np.empty(max(n, 1), dtype=np.int64)
    if n:
        L.gm_memcpy_d2h(j.ctypes.data, out.j, 8 * n)
    zs = {1: 1, 6: 4, 8: 8, 11: 8}[out.type]
    nx = 1 if out.iso else n
    x = np.empty(max(nx * zs, 1), dtype=np.uint8)
    if nx:
        L.gm_memcpy_d2h(x.ctypes.data, out.x, nx * zs)
    return (n, hash(p.tobytes()), hash(j[:n].tobytes()), hash(x[:nx * zs].tobytes()))


def _free(L, out):
    for ptr in (out.p, out.j, out.x):
        L.gm_free(ptr)


class _Batched:
    """A sample of the 4096-graph batch, generated in HBM."""

    def __init__(self, nreq, seed):
        self.specs = W.batch_specs(nreq, seed=seed)
        self.db = W.DeviceBatch(self.specs)
        self.groups = {}
        for name, trip in (("plus_pair_int32", (1, 3, 6)), ("min_plus_int32", (2, 2, 6))):
            idx = [g for g, s in enumerate(self.specs) if s.semiring == name]
            reqs = (capi.GmDot3Request * max(len(idx), 1))()
            for r, g in enumerate(idx):
                v = self.db.views[g]
                reqs[r] = capi.GmDot3Request(v, v, v)
            self.groups[name] = (trip, idx, reqs)

    def run(self, prm):
        L = capi.lib()
        digests = []
        for name, (trip, idx, reqs) in self.groups.items():
            if not idx:
                continue
            rb = np.zeros(len(idx) + 1, dtype=np.int64)
            cb = np.zeros(len(idx) + 1, dtype=np.int64)
            out = capi.GmCsrOut()
            sr = capi.GmSemiring(*trip)
            rc = L.gm_batch_dot3(ctypes.byref(out), rb.ctypes.data, cb.ctypes.data, reqs, len(idx),
                                 False, ctypes.byref(sr), ctypes.byref(prm))
            if rc != 0:
                raise RuntimeError(capi.last_error())
            digests.append(_digest(L, out) if self.check else None)
            _free(L, out)
        return digests

    def close(self):
        self.db.free()


class _Tricount:
    def __init__(self, scale, seed):
        _, Lm, Um = W.tricount_inputs(scale, 16, seed=seed)
        self.dL, self.dU = capi.DeviceCsr(Lm, 1), capi.DeviceCsr(Um, 1)

    def run(self, prm):
        L = capi.lib()
        out = capi.GmCsrOut()
        sr = capi.GmSemiring(1, 3, 8)
        rc = L.gm_dot3(ctypes.byref(out), ctypes.byref(self.dL.view), True, ctypes.byref(self.dL.view),
                       ctypes.byref(self.dU.view), ctypes.byref(sr), ctypes.byref(prm))
        if rc != 0:
            raise RuntimeError(capi.last_error())
        d = _digest(capi.lib(), out) if self.check else None
        _free(L, out)
        return [d]

    def close(self):
        self.dL.free()
        self.dU.free()


class _ER:
    def __init__(self, logn, seed):
        n = 1 << logn
        self.dA = capi.DeviceCsr(W.er(n, n, 8, seed=seed), 11)
        self.dB = capi.DeviceCsr(W.er(n, n, 8, seed=seed + 1), 11)

    def run(self, prm):
        L = capi.lib()
        out = capi.GmCsrOut()
        sr = capi.GmSemiring(1, 1, 11)
        rc = L.gm_saxpy3(ctypes.byref(out), None, False, True, ctypes.byref(self.dA.view),
                         ctypes.byref(self.dB.view), ctypes.byref(sr), ctypes.byref(prm))
        if rc != 0:
            raise RuntimeError(capi.last_error())
        d = _digest(L, out) if self.check else None
        _free(L, out)
        return [d]

    def close(self):
        self.dA.free()
        self.dB.free()


def _time(wl, prm, reps):
    L = capi.lib()
    wl.check = True
    dig = wl.run(prm)           # warm-up, and the results of this setting
    wl.check = False
    L.gm_synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        wl.run(prm)
        L.gm_synchronize()
        best = min(best, time.perf_counter() - t0)
    return best, dig


def tune(wl, space, builtin, make, reps=3, sweeps=2, log=print):
    cur = {k: 0 for k in space}             # 0: built-in value
    base_t, base_dig = _time(wl, make(cur), reps)
    log(f"built-in: {base_t * 1e3:.2f} ms")
    best_t = base_t
    for sweep in range(sweeps):
        for key, cands in space.items():
            for v in cands:
                if v == cur[key] or (cur[key] == 0 and v == builtin[key]):
                    continue
                trial = dict(cur)
                trial[key] = v
                t, dig = _time(wl, make(trial), reps)
                ok = dig == base_dig
                log(f"  sweep {sweep} {key}={v}: {t * 1e3:.2f} ms{'' if ok else ' (results differ: rejected)'}")
                if ok and t < best_t * 0.98:
                    best_t, cur = t, trial
    log(f"tuned: {best_t * 1e3:.2f} ms ({base_t / best_t:.2f}x the built-in setting)")
    return cur, base_t, best_t


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", choices=["batched", "tricount", "er"], default="batched")
    ap.add_argument("--requests", type=int, default=512, help="batched: graphs in the sample")
    ap.add_argument("--scale", type=int, default=18, help="tricount: RMAT scale; er: log2 n")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    capi.init()
    if args.workload == "er":
        wl = _ER(args.scale, 1)
        space, builtin, prefix = SAXPY3_SPACE, SAXPY3_BUILTIN, "saxpy3"

        def make(d):
            return capi.GmSaxpy3Params(d["wave_small_max"], d["wave_large_max"], d["block_max"],
                                       d["mode"])
    else:
        wl = _Batched(args.requests, 7) if args.workload == "batched" else _Tricount(args.scale, 3)
        space, builtin, prefix = DOT3_SPACE, DOT3_BUILTIN, "dot3"

        def make(d):
            return capi.GmDot3Params(d["thread_max"], d["team16_max"], d["wave_max"],
                                     d["bitmap_min_len"], d["order"], d["row_min"])
    try:
        cur, t0, t1 = tune(wl, space, builtin, make, reps=args.reps,
                           log=lambda m: print(m, flush=True))
    finally:
        wl.close()
    with open(args.out, "w") as f:
        f.write(f"# autotuned on workload {args.workload}: {t0 * 1e3:.3f} ms -> {t1 * 1e3:.3f} ms\n")
        for k, v in cur.items():
            if v != 0:
                f.write(f"{prefix}.{k} {v}\n")
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
