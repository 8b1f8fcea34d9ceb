"""Benchmark of GrB_mxm on one MI355X.

Default workload (one step = one GrB_mxm call through the C API, inputs
already resident in HBM):

    C = A*B, plus_times fp64 (saxpy3), A and B n = 2^20 square CSR with
    exactly 8 distinct uniform columns per row, values uniform [0,1).

value = semiring multiplies per second over the timed steps (2^26 mults per
call, from a host pass over the inputs).  The roofline object covers the
dominant kernel (the saxpy3 numeric kernel for 8-entry rows), timed with HIP
events on the library stream over the timed region: achieved = algorithmic
bytes of one launch / its average duration.  cpu_baseline times the CPU
oracle (a plain-C OpenMP restatement of the reference's saxpy3) on the host.

Other workloads (--workload): dot3-rmat14 / dot3-rmat22 (C<L> = L*U',
plus_pair int64, the triangle count).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X published HBM3E bandwidth

# HBM bytes per launch of the dominant kernel(s), measured by two rocprofv3
# counter runs (FETCH_SIZE, WRITE_SIZE; each scaled by what the counter reads
# for a 1 GiB device copy) and kept as text summaries beside the kernel trace
PROFILE_SUMMARY = {"saxpy3-er20": "profiles/sx_fused_saxpy3_er20.txt",
                   "dot3-rmat22": "profiles/dot3_tricount_rmat22.txt"}


def measured_traffic(workload):
    """(bytes per launch, summary file) from the committed counter summary, or (None, None)."""
    rel = PROFILE_SUMMARY.get(workload)
    if rel is None or not os.path.exists(os.path.join(ROOT, rel)):
        return None, None
    for line in open(os.path.join(ROOT, rel)):
        if line.startswith("traffic_bytes_per_launch="):
            return float(line.split("=", 1)[1]), rel
    return None, None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="saxpy3-er20",
                    choices=["saxpy3-er20", "dot3-rmat14", "dot3-rmat22"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="budget for the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the timed steps, copy MIB MiB device-to-device once (a "
                         "known-size copy for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def saxpy3_alg_bytes(A, B, nnzC, sa=8, sb=8, sc=8):
    """Algorithmic bytes of C=A*B (single-pass gather model): read A once,
    for every A(i,k) read B's row pointers and row k once, write C once."""
    rowlen_b = np.diff(B.p)
    flops = int(rowlen_b[A.j].sum())
    return (8 * (A.nrows + 1) + (8 + sa) * A.nnz + 16 * A.nnz + (8 + sb) * flops
            + 8 * (A.nrows + 1) + (8 + sc) * nnzC), flops


def dot3_alg_bytes(M, A, B, nnzC, sa=0, sb=0, sc=8):
    """Algorithmic bytes of C<M>=A*B'.  Merge model: both lists of every mask
    entry streamed once.  Probe model (the lower bound used for the roofline):
    only the shorter list of every entry is read (its elements are looked up
    in the longer row's index), plus M once and C written once."""
    la, lb = np.diff(A.p), np.diff(B.p)
    rows = np.repeat(np.arange(M.nrows), np.diff(M.p))
    a, b = la[rows], lb[M.j]
    fixed = 8 * (M.nrows + 1) + 8 * M.nnz + 8 * (M.nrows + 1) + (8 + sc) * nnzC
    merge = int(fixed + (16 + (8 + sa) * a + (8 + sb) * b).sum())
    probe = int(fixed + (16 + (8 + max(sa, sb)) * np.minimum(a, b)).sum())
    return merge, probe


def timed_steps(capi, fn, steps, warmup):
    L = capi.lib()
    for _ in range(warmup):
        fn()
    L.gm_synchronize()
    capi.prof_reset()
    capi.prof_enable(True)
    L.gm_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    L.gm_synchronize()
    t1 = time.perf_counter()
    capi.prof_enable(False)
    kern = capi.prof_get()
    return (t1 - t0) * 1e3 / steps, kern


def cpu_baseline(fn, budget_s, unit_per_call):
    import pyoracle as O
    fn()  # warm-up (allocations, page faults)
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or reps >= 200:
            break
    return unit_per_call * reps / el, O.num_threads(), reps, el


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build measures "
              "the single-GPU path, multi-GPU sharding is a design note.", file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from graphblas_amd import capi
    from graphblas_amd import workloads as W

    capi.init()
    dev = capi.device_info()
    if args.copy_calibration > 0:
        L = capi.lib()
        nb = args.copy_calibration << 20
        src, dst = L.gm_malloc(nb), L.gm_malloc(nb)
        L.gm_memset(src, 1, nb)
        L.gm_synchronize()
        L.gm_memcpy_d2d(dst, src, nb)
        L.gm_synchronize()
        L.gm_free(src)
        L.gm_free(dst)
    out = {"higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "data": "synthetic", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "metric": "semiring mults/s for GrB_mxm",
           "unit": "mults/s"}

    if args.workload == "saxpy3-er20":
        n = 1 << 20
        A, B = W.er(n, n, 8, seed=1), W.er(n, n, 8, seed=2)
        Ad, Bd = capi.Matrix.from_csr(A, "fp64"), capi.Matrix.from_csr(B, "fp64")
        C = capi.Matrix("fp64", n, n)

        def step():
            capi.mxm_checked(C, None, "GrB_PLUS_TIMES_SEMIRING_FP64", Ad, Bd)

        step()
        nnzC = C.nvals
        balg, mults = saxpy3_alg_bytes(A, B, nnzC)
        ms, kern = timed_steps(capi, step, args.steps, args.warmup)
        dom = max((k for k in kern if k.startswith("sx_")), key=lambda k: kern[k][1])
        out.update({"dtype": "f64",
                    "config": {"workload": "saxpy3 C=A*B plus_times fp64, Erdos-Renyi n=2^20, "
                                           "8 nnz/row, seeded", "n": n, "nnz_per_row": 8,
                               "nnz_C": nnzC, "method": "saxpy3 (single-pass fused kernel, LDS sort-compress)"}})

        def cpu_fn():
            import pyoracle as O
            O.saxpy3("plus_times_fp64", A, B)
        cpu_unit = mults
        cpu_sample = "full workload (ER 2^20 x 8, plus_times fp64), symbolic+numeric, repeated"
    else:
        scale = 14 if args.workload == "dot3-rmat14" else 22
        Ag, L, U = W.tricount_inputs(scale, 16, seed=1)
        Ld, Ud = capi.Matrix.from_csr(L, "bool"), capi.Matrix.from_csr(U, "bool")
        C = capi.Matrix("int64", L.nrows, L.nrows)

        def step():
            C.clear()      # a fresh, empty C for every call
            capi.mxm_checked(C, Ld, "GxB_PLUS_PAIR_INT64", Ld, Ud, "ST1")

        step()
        ntri = capi.reduce_sum_int64(C)
        nnzC = C.nvals
        mults = int(ntri)      # plus_pair: every multiply is one common neighbour
        balg_merge, balg = dot3_alg_bytes(L, L, U, nnzC)
        ms, kern = timed_steps(capi, step, args.steps, args.warmup)
        # dot3 spreads one call over several kernels: the roofline covers
        # all of them together, with the probe-model bytes
        dom = "dot3 (all kernels)"
        kern[dom] = (args.steps, sum(v[1] for k, v in kern.items() if k.startswith("dot3")))
        out["alg_bytes_merge_model"] = balg_merge
        out.update({"dtype": "int64",
                    "config": {"workload": f"dot3 C<L>=L*U' plus_pair int64 (triangle count), "
                                           f"RMAT scale {scale}, ef 16, seeded, no presort",
                               "n": 1 << scale, "nnz_L": L.nnz, "triangles": int(ntri),
                               "nnz_C": nnzC, "method": "dot3"}})

        def cpu_fn():
            import pyoracle as O
            O.dot3("plus_pair_int64", L, L, U)
        cpu_unit = mults
        cpu_sample = f"full workload (RMAT scale {scale} tricount dot3), repeated"
        if scale == 22:
            cpu_sample = "rows 0..n/64 of the scale-22 mask (bounded sample)"
            nr = L.nrows // 64
            Ls = W.Csr(nr, L.ncols, L.p[:nr + 1].copy(), L.j[:L.p[nr]].copy())
            sub_ref = None

            def cpu_fn():  # noqa: F811
                import pyoracle as O
                return O.dot3("plus_pair_int64", Ls, Ls, U)
            import pyoracle as O
            sub_ref = O.dot3("plus_pair_int64", Ls, Ls, U)
            cpu_unit = int(sub_ref.x.sum())

    launches, tot_ms = kern.get(dom, (0, 0.0))
    avg_ms = tot_ms / max(launches, 1)
    achieved = balg / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None
    out["value"] = mults / (ms * 1e-3)
    out["ms_per_step"] = ms
    out["mults_per_call"] = mults
    out["alg_bytes_per_call"] = balg
    out["alg_gbs_per_call"] = balg / (ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(args.workload)
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                       "traffic_source": traffic_src,
                       "traffic_gbs": (traffic / (avg_ms * 1e-3) / 1e9) if traffic and avg_ms > 0 else None,
                       "kernel": dom, "kernel_avg_ms": avg_ms,
                       "kernel_alg_bytes": balg}
    out["kernels_ms_per_step"] = {k: v[1] / args.steps for k, v in sorted(kern.items())}
    out["device"] = dev
    if not args.no_cpu:
        v, cores, reps, el = cpu_baseline(cpu_fn, args.cpu_seconds, cpu_unit)
        out["cpu_baseline"] = {"value": v, "unit": "mults/s", "cores": cores, "kind": "port",
                               "sample": f"{cpu_sample}: {reps} calls in {el:.1f} s"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
