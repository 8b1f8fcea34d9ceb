"""Benchmark of GrB_mxm on one MI355X.

Default workload (one step = one GrB_mxm call through the C API, inputs
already resident in HBM): the triangle count C<L> = L*U' with the plus_pair
int64 semiring (GrB_DESC_ST1, the dot3 method) on an RMAT scale-22 graph
(n = 4,194,304, edge factor 16, a/b/c/d = 0.57/0.19/0.19/0.05, symmetrised,
deduplicated, no self-loops, no degree presort).

value = semiring multiplies per second over the timed steps (one multiply
per common neighbour: sum(C), the triangle count).  The roofline object
covers all dot3 kernels of a call together, timed with HIP events on the
library stream over the timed region: achieved = algorithmic bytes of one
call (probe model: the shorter list of every mask entry read once, M read
once, C written once) / the kernels' summed duration.  The result of the
first call is checked against the committed fixture
tests/golden/fullsize_rmat22.json (nnz(C), sum(C)).  cpu_baseline times the
CPU oracle (a plain-C OpenMP restatement of the reference's dot3) on a
bounded sample of the same mask (its first n/64 rows, which hold the hub
rows).

Other workloads (--workload): saxpy3-er20 (C = A*B plus_times fp64, A and B
n = 2^20 with 8 uniform columns per row); dot3-rmat14; saxpy3-tricount-rmat22
(the same count as C<L> = L*L with the GxB_AxB_SAXPY hint);
mxm-tricount-rmat22 (C<L> = L*L with no hint: GrB_mxm's flop rule runs dot3
on L'); batched-4096; served-rmat20.
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
    ap.add_argument("--workload", default="dot3-rmat22",
                    choices=["saxpy3-er20", "dot3-rmat14", "dot3-rmat22", "served-rmat20",
                             "batched-4096", "saxpy3-tricount-rmat22", "mxm-tricount-rmat22"])
    ap.add_argument("--requests", type=int, default=4096,
                    help="served workload: number of requests")
    ap.add_argument("--rate", type=float, default=10000.0,
                    help="served workload: Poisson arrival rate in requests/s")
    ap.add_argument("--tuning", default=os.path.join("graphblas_amd", "tuning", "mi355x_batched.txt"),
                    help="batched workload: autotuned bin cutoffs (gm_load_tuning); '' for none")
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


def cpu_model() -> str:
    """The host CPU's model name (lscpu's 'Model name', from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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


def served_requests(n, nreq, seed):
    """nreq BFS-step frontiers on an n-vertex graph: k log-uniform in [1, 256]
    distinct uniform sources per request."""
    from graphblas_amd import workloads as W
    rng = np.random.default_rng(seed)
    reqs = []
    for r in range(nreq):
        k = W.loguniform_int(seed, r, 1, 256)
        reqs.append(np.sort(rng.choice(n, size=k, replace=False)).astype(np.int64))
    return reqs


def run_server(capi, W, view, reqs, rate, workers, max_batch, budget, seed=5):
    """Drive reqs through a fresh server (Poisson arrivals at the given rate,
    or all at once when rate <= 0) with the C++ load generator; returns
    (wall seconds, results, server stats)."""
    srv = capi.Server(view, (4, 4, 8), workers=workers, max_batch_reqs=max_batch,
                      flop_budget=budget)
    arrive = None
    if rate > 0:
        rng = np.random.default_rng(seed)
        arrive = np.cumsum(rng.exponential(1e6 / rate, len(reqs)))
    res, wall = W.server_drive(srv, reqs, arrive)
    st = srv.stats()
    srv.close()
    return wall, res, st


def fetch_all(capi, view, reqs, workers, max_batch, budget):
    """Every request's rows (untimed; for comparing two configurations)."""
    srv = capi.Server(view, (4, 4, 8), workers=workers, max_batch_reqs=max_batch,
                      flop_budget=budget)
    ids = [srv.submit(s) for s in reqs]
    rows = []
    for i in ids:
        r = srv.wait(i)
        rows.append(srv.fetch(i, r))
        srv.release(i)
    srv.close()
    return rows


def bench_served(args, capi, W):
    """Served frontier expansion on a resident RMAT scale-20 graph."""
    G = W.rmat_sym(20, 16, seed=1)
    dG = capi.DeviceCsr(G, 1)
    deg = np.diff(G.p)
    reqs = served_requests(G.nrows, args.requests, seed=11)
    flops = np.array([int(deg[s].sum()) for s in reqs], dtype=np.int64)
    total_flops = int(flops.sum())
    budget = 1 << 22
    # every result is the same served (mixed into batches) and alone
    rows_srv = fetch_all(capi, dG.view, reqs, 2, 256, budget)
    rows_pl = fetch_all(capi, dG.view, reqs, 1, 1, budget)
    same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
               for a, b in zip(rows_srv, rows_pl))
    del rows_srv, rows_pl
    # warm-up, then saturation throughput: every request offered at once
    run_server(capi, W, dG.view, reqs[:256], 0, 2, 256, budget)
    t_srv, r_srv, st_srv = run_server(capi, W, dG.view, reqs, 0, 2, 256, budget)
    t_pl, r_pl, st_pl = run_server(capi, W, dG.view, reqs, 0, 1, 1, budget)
    sat = len(reqs) / t_srv
    rate = args.rate
    # open loop: Poisson arrivals at the fixed rate
    t_po, r_po, st_po = run_server(capi, W, dG.view, reqs, rate, 2, 256, budget)
    t_pp, r_pp, st_pp = run_server(capi, W, dG.view, reqs, rate, 1, 1, budget)

    def pct(res):
        lat = np.array([r.latency_us for r in res]) * 1e-3
        return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    p50, p99 = pct(r_po)
    q50, q99 = pct(r_pp)
    # roofline of the served kernels: the saxpy3 single-pass gather model of
    # every request (F's rows and entries, 16 B of graph row pointers and
    # 8 B per neighbour of every source, C written as index + int64 value)
    # over the kernel time of one profiled saturation pass
    nnzc = np.array([r.nvals for r in r_srv], dtype=np.int64)
    ks = np.array([len(s) for s in reqs], dtype=np.int64)
    srv_bytes = int((8 * (ks + 1) + 8 * ks + 16 * ks + 8 * flops + 8 * (ks + 1) + 16 * nnzc).sum())
    capi.prof_reset()
    capi.prof_enable(True)
    run_server(capi, W, dG.view, reqs, 0, 2, 256, budget)
    capi.lib().gm_synchronize()
    capi.prof_enable(False)
    kern = capi.prof_get()
    ktot = sum(v[1] for v in kern.values())
    roof = None
    if ktot > 0:
        ach = srv_bytes / (ktot * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": None,
                "kernel": "all kernels of a saturation pass (summed event times; batches "
                          "overlap on the device, so this understates the achieved rate)",
                "kernel_ms": ktot, "kernel_alg_bytes": srv_bytes,
                "kernels_ms": {k: v[1] for k, v in sorted(kern.items())}}
    out = {"metric": "served frontier expansion: semiring mults/s (saturation) with p50/p99 ms "
                     "at a fixed Poisson rate",
           "unit": "mults/s", "dtype": "int64",
           "config": {"workload": "served BFS steps C<!V>=F*A, any_secondi int64, RMAT scale 20 "
                                  "ef 16 bool resident in HBM; k log-uniform [1,256] sources per "
                                  "request, V = F (first step)",
                      "n": G.nrows, "nnz": G.nnz, "requests": len(reqs),
                      "flop_budget": budget, "workers": 2, "max_batch_reqs": 256},
           "value": total_flops / t_srv, "ms_per_step": t_srv * 1e3 / len(reqs),
           "requests_per_s": sat,
           "plain": {"requests_per_s": len(reqs) / t_pl, "mults_per_s": total_flops / t_pl,
                     "what": "one saxpy3 launch sequence per request (batch size 1, one worker)"},
           "served_vs_plain": (len(reqs) / t_srv) / (len(reqs) / t_pl),
           "results_identical_served_vs_plain": bool(same),
           "mean_batch": st_srv["mean_batch"],
           "poisson": {"rate_per_s": rate, "p50_ms": p50, "p99_ms": p99,
                       "achieved_per_s": len(reqs) / t_po, "mean_batch": st_po["mean_batch"],
                       "max_in_flight": st_po["max_in_flight"],
                       "plain_p50_ms": q50, "plain_p99_ms": q99,
                       "plain_achieved_per_s": len(reqs) / t_pp},
           "total_mults": total_flops,
           "roofline": roof}
    if not args.no_cpu:
        # oracle on a stacked sample of the requests (C<!F> = F*A)
        nr = max(1, len(reqs) // 8)
        srcs = np.concatenate(reqs[:nr])
        F = W.Csr(len(srcs), G.nrows, np.arange(len(srcs) + 1, dtype=np.int64), srcs)
        unit = int(flops[:nr].sum())

        def cpu_fn():
            import pyoracle as O
            O.saxpy3("any_secondi_int64", F, G, M=F, mask_comp=True)
        v, cores, reps, el = cpu_baseline(cpu_fn, args.cpu_seconds, unit)
        out["cpu_baseline"] = {"value": v, "unit": "mults/s", "cores": cores, "kind": "port",
                               "cpu_model": cpu_model(),
                               "sample": f"first {nr} requests stacked into one C<!F>=F*A: "
                                         f"{reps} calls in {el:.1f} s"}
    else:
        out["cpu_baseline"] = None
    dG.free()
    return out


BATCH_SR = {"plus_pair_int32": (1, 3, 6), "min_plus_int32": (2, 2, 6)}


def bench_batched(args, capi, W):
    """4096 uneven RMAT graphs, C<A> = A*A' each (dot3), plus_pair int32 or
    min_plus int32: the batched path (one gm_batch_dot3 per semiring) against
    the plain path (one gm_dot3 per request)."""
    import ctypes
    L = capi.lib()
    tuning = None
    if args.tuning:
        tuning = os.path.join(ROOT, args.tuning) if not os.path.isabs(args.tuning) else args.tuning
        capi.load_tuning(tuning)
    nreq = args.requests
    specs = W.batch_specs(nreq, seed=1)
    t0 = time.perf_counter()
    db = W.DeviceBatch(specs)
    t_gen = time.perf_counter() - t0
    groups = {name: [g for g, s in enumerate(specs) if s.semiring == name] for name in BATCH_SR}
    mask_struct = False                      # C<A>: the mask is A's (nonzero) values

    def plain(check):
        sums = {}
        for g, s in enumerate(specs):
            out = capi.GmCsrOut()
            sr = capi.GmSemiring(*BATCH_SR[s.semiring])
            v = db.views[g]
            rc = L.gm_dot3(ctypes.byref(out), ctypes.byref(v), mask_struct, ctypes.byref(v),
                           ctypes.byref(v), ctypes.byref(sr), None)
            if rc != 0:
                raise RuntimeError(f"gm_dot3 request {g}: {capi.last_error()}")
            if check:
                sums[g] = stacked_sums(L, out, 0, s.n, 0)
            for ptr in (out.p, out.j, out.x):
                L.gm_free(ptr)
        return sums

    def batched(check):
        sums = {}
        for name, idx in groups.items():
            if not idx:
                continue
            reqs = (capi.GmDot3Request * len(idx))()
            for r, g in enumerate(idx):
                v = db.views[g]
                reqs[r] = capi.GmDot3Request(v, v, v)
            rb = np.zeros(len(idx) + 1, dtype=np.int64)
            cb = np.zeros(len(idx) + 1, dtype=np.int64)
            out = capi.GmCsrOut()
            sr = capi.GmSemiring(*BATCH_SR[name])
            rc = L.gm_batch_dot3(ctypes.byref(out), rb.ctypes.data, cb.ctypes.data, reqs, len(idx),
                                 mask_struct, ctypes.byref(sr), None)
            if rc != 0:
                raise RuntimeError(f"gm_batch_dot3 {name}: {capi.last_error()}")
            if check:
                for r, g in enumerate(idx):
                    sums[g] = stacked_sums(L, out, int(rb[r]), int(rb[r + 1]), int(cb[r]))
            for ptr in (out.p, out.j, out.x):
                L.gm_free(ptr)
        return sums

    # untimed passes: results of both paths, compared request by request
    ref = plain(True)
    got = batched(True)
    same = all(ref[g] == got[g] for g in range(nreq))
    # total semiring multiplies: common neighbours over all mask entries;
    # algorithmic bytes of every request (probe model, as dot3_alg_bytes)
    mults = 0
    alg_bytes = 0
    for g in range(nreq):
        out = capi.GmCsrOut()
        sr = capi.GmSemiring(1, 3, 8)
        v = db.views[g]
        rc = L.gm_dot3(ctypes.byref(out), ctypes.byref(v), True, ctypes.byref(v), ctypes.byref(v),
                       ctypes.byref(sr), None)
        if rc != 0:
            raise RuntimeError(capi.last_error())
        r = ctypes.c_int64()
        L.gm_reduce_scalar(ctypes.byref(r), 1, 8, out.x, out.nvals, out.iso)
        mults += r.value
        mn, ln = ctypes.c_int64(), ctypes.c_int64()
        if L.gm_dot3_work(ctypes.byref(mn), ctypes.byref(ln), ctypes.byref(v), ctypes.byref(v),
                          ctypes.byref(v)) != 0:
            raise RuntimeError(capi.last_error())
        nm = int(db.nzbase[g + 1] - db.nzbase[g])
        sab = 4 if specs[g].semiring == "min_plus_int32" else 0
        # M: row pointers, indices, int32 mask values; C: row pointers,
        # indices, int32 values; per entry 16 B of row pointers + the shorter list
        alg_bytes += (8 * (specs[g].n + 1) + 12 * nm + 8 * (specs[g].n + 1) + 12 * out.nvals
                      + 16 * nm + (8 + sab) * mn.value)
        for ptr in (out.p, out.j, out.x):
            L.gm_free(ptr)
    # both legs timed the same way, without the per-kernel events; the
    # kernel split comes from a separate profiled pass of each leg
    L.gm_synchronize()
    t0 = time.perf_counter()
    plain(False)
    L.gm_synchronize()
    t_plain = time.perf_counter() - t0
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        batched(False)
    L.gm_synchronize()
    t_b = (time.perf_counter() - t0) / steps
    capi.prof_reset()
    capi.prof_enable(True)
    plain(False)
    L.gm_synchronize()
    capi.prof_enable(False)
    plain_kernel_ms = sum(v[1] for v in capi.prof_get().values())
    capi.prof_reset()
    capi.prof_enable(True)
    batched(False)
    L.gm_synchronize()
    capi.prof_enable(False)
    kern = capi.prof_get()
    steps_prof = 1
    nnz = int(db.nzbase[-1])
    ktot = sum(v[1] for v in kern.values()) / steps_prof
    ach = alg_bytes / (ktot * 1e-3) / 1e9 if ktot > 0 else 0.0
    ns = np.array([s.n for s in specs])
    out = {"metric": "batched uneven mxm: semiring mults/s, batched vs plain one-call-per-request",
           "unit": "mults/s", "dtype": "int32",
           "config": {"workload": f"{nreq} RMAT graphs (n log-uniform [512, 65536], edge factor "
                                  "uniform [4, 32], a/b/c = 0.57/0.19/0.19, int32 weights [1,100]);"
                                  " C<A>=A*A' by dot3 with plus_pair int32 or min_plus int32 (1/2 "
                                  "each)", "requests": nreq, "nnz_total": nnz,
                      "n_total": int(ns.sum()),
                      "per_semiring": {k: len(v) for k, v in groups.items()},
                      "method": "gm_batch_dot3 (block-diagonal stacking, one binned dot3 per "
                                "semiring) vs gm_dot3 per request"},
           "value": mults / t_b, "ms_per_step": t_b * 1e3, "steps": steps,
           "total_mults": int(mults),
           "plain": {"ms": t_plain * 1e3, "kernel_ms": plain_kernel_ms,
                     "mults_per_s": mults / t_plain, "requests_per_s": nreq / t_plain},
           "batched_vs_plain": t_plain / t_b,
           "results_identical_batched_vs_plain": bool(same),
           "generate_ms": t_gen * 1e3,
           "tuning": os.path.relpath(tuning, ROOT) if tuning else None,
           "kernels_ms_per_step": {k: v[1] / steps_prof for k, v in sorted(kern.items())},
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "all kernels of one batched step (summed event times)",
                        "kernel_ms": ktot, "kernel_alg_bytes": int(alg_bytes)}}
    if not args.no_cpu:
        import pyoracle as O
        hosts, units = [], 0
        for g in range(nreq):
            if specs[g].n <= 4096:
                A = W.batch_graph_host(specs[g])
                c = O.dot3("plus_pair_int64", A, A, A)
                hosts.append((specs[g].semiring, A))
                units += int(c.x.sum())
            if len(hosts) >= 64:
                break

        def cpu_fn():
            for name, A in hosts:
                O.dot3(name, A, A, A, mask_struct=False)
        v, cores, reps, el = cpu_baseline(cpu_fn, args.cpu_seconds, units)
        out["cpu_baseline"] = {"value": v, "unit": "mults/s", "cores": cores, "kind": "port",
                               "cpu_model": cpu_model(),
                               "sample": f"the first {len(hosts)} requests with n <= 4096, one "
                                         f"oracle dot3 each: {reps} passes in {el:.1f} s"}
    else:
        out["cpu_baseline"] = None
    db.free()
    return out


def stacked_sums(L, out, r0, r1, col_base):
    """(nnz, sum of column indices - col_base, wrapped int32 sum of values) of
    rows [r0, r1) of a CSR result in HBM."""
    import ctypes
    pp = np.zeros(2, dtype=np.int64)
    L.gm_memcpy_d2h(pp.ctypes.data, out.p + 8 * r0, 8)
    L.gm_memcpy_d2h(pp.ctypes.data + 8, out.p + 8 * r1, 8)
    lo, hi = int(pp[0]), int(pp[1])
    nz = hi - lo
    sj, sx = ctypes.c_int64(0), ctypes.c_int32(0)
    if nz > 0:
        L.gm_reduce_scalar(ctypes.byref(sj), 1, 8, out.j + 8 * lo, nz, 0)
        L.gm_reduce_scalar(ctypes.byref(sx), 1, 6, out.x + (0 if out.iso else 4 * lo), nz, out.iso)
    return nz, sj.value - nz * col_base, sx.value


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
    if args.workload in ("served-rmat20", "batched-4096"):
        if args.workload == "served-rmat20":
            out.update(bench_served(args, capi, W))
        else:
            out.update(bench_batched(args, capi, W))
        out["device"] = dev
        print(json.dumps(out))
        return

    if args.workload in ("saxpy3-tricount-rmat22", "mxm-tricount-rmat22"):
        # the triangle count as C<L> = L*L (plus_pair int64, structural mask),
        # the same count as the dot3 form C<L> = L*U'.  saxpy3-...: the
        # GxB_AxB_SAXPY hint keeps the masked saxpy3 method; mxm-...: no hint,
        # GrB_mxm's flop rule picks the method (dot3 on L' here)
        forced = args.workload.startswith("saxpy3")
        Ag, L, U = W.tricount_inputs(22, 16, seed=1)
        Ld = capi.Matrix.from_csr(L, "bool")
        C = capi.Matrix("int64", L.nrows, L.nrows)
        dsc = capi.Descriptor(mask_struct=True, axb=capi.GxB_AxB_SAXPY if forced else 0)

        def step():
            C.clear()
            capi.mxm_checked(C, Ld, "GxB_PLUS_PAIR_INT64", Ld, Ld, dsc)

        step()
        ntri = capi.reduce_sum_int64(C)
        nnzC = C.nvals
        sx_bytes, flops = saxpy3_alg_bytes(L, L, nnzC, sa=0, sb=0, sc=8)
        if forced:
            balg = sx_bytes
        else:
            # dot3's probe model on (L, L, L'), plus the transpose of L
            # (row pointers and indices read once and written once)
            balg = dot3_alg_bytes(L, L, U, nnzC)[1] + 2 * (8 * (L.nrows + 1) + 8 * L.nnz)
        ms, kern = timed_steps(capi, step, args.steps, args.warmup)
        dsc.free()
        pre = ("sx_",) if forced else ("dot3_", "transpose_")
        dom = ("saxpy3" if forced else "dot3 + transpose") + " (all kernels)"
        kern[dom] = (args.steps, sum(v[1] for k, v in kern.items() if k.startswith(pre)))
        # the multiplies the masked product needs (each one a common
        # neighbour), whatever the method: the same count as dot3-rmat22
        mults = int(ntri)
        out.update({"dtype": "int64",
                    "config": {"workload": ("saxpy3" if forced else "GrB_mxm") +
                                           " C<L>=L*L plus_pair int64 (triangle count), "
                                           "RMAT scale 22, ef 16, seeded",
                               "n": L.nrows, "nnz_L": L.nnz, "triangles": int(ntri),
                               "nnz_C": nnzC, "flops_unmasked": int(flops),
                               "method": "saxpy3 (masked, hash per row; GxB_AxB_SAXPY)" if forced
                               else "flop rule: dot3 on L' (transpose + masked dot product)"}})
        nr = L.nrows // 64
        Ls = W.Csr(nr, L.ncols, L.p[:nr + 1].copy(), L.j[:L.p[nr]].copy())
        cpu_sample = "rows 0..n/64 of C<L>=L*L (bounded sample)"
        cpu_unit = 0
        if not args.no_cpu:
            import pyoracle as O
            cpu_unit = int(O.saxpy3("plus_pair_int64", Ls, L, M=Ls).x.sum())

        def cpu_fn():
            import pyoracle as O
            O.saxpy3("plus_pair_int64", Ls, L, M=Ls)
    elif args.workload == "saxpy3-er20":
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
        if scale == 22:
            # the first call's result against the oracle's committed fixture
            with open(os.path.join(ROOT, "tests", "golden", "fullsize_rmat22.json")) as f:
                gold = json.load(f)
            ok = (int(ntri) == gold["sum_C"] and int(nnzC) == gold["nnz_C"])
            out["check"] = {"against": "tests/golden/fullsize_rmat22.json (oracle's nnz(C), sum(C))",
                            "ok": bool(ok)}
            if not ok:
                raise SystemExit(f"dot3-rmat22 result differs from the fixture: sum {ntri} "
                                 f"(want {gold['sum_C']}), nnz {nnzC} (want {gold['nnz_C']})")
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
                               "cpu_model": cpu_model(),
                               "sample": f"{cpu_sample}: {reps} calls in {el:.1f} s"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
