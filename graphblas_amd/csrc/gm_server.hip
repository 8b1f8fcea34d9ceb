// gm_server.hip: continuous-batching request loop for frontier expansion.
//
// A server owns one device-resident graph A (CSR in HBM).  A request is a
// multi-source breadth-first search of one or more levels: F has one row per
// source (F(r, s_r) = 1) and V, the visited set, is the request's own k x n
// matrix (by default V = F).  Each level is the step C<!V> = F*A; with the
// ANY_SECONDI semiring C(r,i) is the parent of every newly reached vertex i
// (the reference's BFS step, GraphBLAS/@GrB/bfs.m:140,162).  Between levels
// the request's state stays in HBM: the next frontier is C and the visited
// set becomes V + C (its values are the parents: the source itself for the
// source row, -1 for visited entries the caller supplied).
//
// Worker threads (one HIP stream each, default priority) take pending
// requests in arrival order and admit them into a batch while the batch's
// workspace estimate stays within the byte budget -- the lesser of the
// configured budget and the HBM still free (driver-free plus the arena's
// cached blocks) -- and its flop upper bound (sum over its frontier entries
// of deg) within the flop budget; a single request larger than either runs
// alone.  The admitted requests' frontiers and visited sets are stacked on
// the device from their HBM slices (no host copy of F or V after submit) and
// one saxpy3 call computes the level for all of them.  A request with more
// levels to run goes back to the front of the queue, so it meets the
// requests that arrived meanwhile in the next batch (continuous batching at
// level granularity), and several workers' batches overlap on the device.
// A batch that runs out of device memory is queued again and the next
// admission takes at most half as many requests (the cap doubles back after
// each batch that succeeds); a single request that does not fit fails with
// GM_OUT_OF_MEMORY.

#include "gm_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>
#include <cstring>

namespace gm
{

static double now_us()
{
    using namespace std::chrono;
    return (double) duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

// a device CSR whose row slices belong to one or more requests; freed with
// the last request that refers to it
struct SvHold
{
    gm_csr_out M;
    SvHold() { memset(&M, 0, sizeof(M)); }
    ~SvHold()
    {
        gm_free(M.p);
        gm_free(M.j);
        gm_free(M.x);
    }
};

// rows [row0, row0 + k) of a held CSR
struct SvSlice
{
    std::shared_ptr<SvHold> h;
    int64_t row0 = 0;
    int64_t nnz = 0;
};

struct SvRequest
{
    int64_t id = 0;
    std::vector<int64_t> src;
    int64_t k = 0;
    SvSlice F, V, C;                  // frontier, visited set, last level's result
    // before the first level F and V are the host arrays below (F: one entry
    // per source; V: the visited rows with their values), uploaded with the
    // rest of the batch that first runs the request; later levels read the
    // HBM slices above
    std::vector<int64_t> hvp, hvj;
    std::vector<uint8_t> hvx;
    int64_t flops = 0;                // flop bound of the next level
    int64_t flops_total = 0;          // flop bounds of the levels run
    int64_t max_steps = 1;            // levels to run (<= 0: until the frontier is empty)
    int64_t steps = 0;                // levels done
    bool keep_visited = false;
    double t_arrive = 0, t_start = -1, t_done = 0;
    int state = 0;                    // 0 pending, 1 running, 2 finished
    int rc = 0;
    int64_t batch_id = 0, batch_size = 0;
    bool needs_visited() const { return keep_visited || max_steps <= 0 || max_steps > 1; }
};

// per-row flop bound of the next level: sum of deg(j) over the row's entries
// (one wavefront per row)
__global__ __launch_bounds__(256) void k_sv_rowflops(const int64_t *__restrict__ Cp,
                                                     const int64_t *__restrict__ Cj,
                                                     const int64_t *__restrict__ Gp, int64_t nrows,
                                                     int64_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    for (int64_t r = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < nrows;
         r += ((int64_t) gridDim.x * blockDim.x) >> 6)
    {
        int64_t f = 0;
        for (int64_t q = Cp[r] + lane; q < Cp[r + 1]; q += 64)
        {
            const int64_t v = Cj[q];
            f += Gp[v + 1] - Gp[v];
        }
        #pragma unroll
        for (int off = 32; off > 0; off >>= 1) f += __shfl_down(f, off, 64);
        if (lane == 0) out[r] = f;
    }
}

// workspace estimate of a request's next level: the stacked frontier and
// visited set, the product allocated at its flop bound (index, value and
// per-row bookkeeping), and the union V + C when the visited set is kept
static int64_t sv_bytes(const SvRequest &r)
{
    const int64_t vn = r.V.nnz;
    int64_t b = 16 * (r.F.nnz + vn) + 24 * r.flops + 64 * r.k;
    if (r.needs_visited()) b += 16 * (vn + r.flops);
    return b;
}

} // namespace gm

using namespace gm;

struct gm_server
{
    gm_csr graph;
    gm_semiring sr;
    gm_server_config cfg;
    std::vector<int64_t> deg;         // host copy of the graph's row lengths
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<SvRequest>> pending;
    std::unordered_map<int64_t, std::shared_ptr<SvRequest>> reqs;
    std::vector<std::thread> workers;
    bool stop = false;
    int64_t cap_reqs = 0;             // requests per batch at most (halved after an
                                      // out-of-memory batch, doubled back after success)
    int64_t next_id = 1, next_batch = 1;
    int64_t completed = 0, batches = 0, batched_reqs = 0, in_flight = 0, max_in_flight = 0;
    int64_t levels = 0, oom_splits = 0;
    double flops_done = 0;
};

// value v of type zt as bytes
static void sv_value(uint8_t *out, int32_t zt, int64_t v)
{
    switch (zt)
    {
        case GM_FP64: { double d = (double) v; memcpy(out, &d, 8); break; }
        case GM_FP32: { float f = (float) v; memcpy(out, &f, 4); break; }
        case GM_INT64: case GM_UINT64: memcpy(out, &v, 8); break;
        case GM_INT32: case GM_UINT32: { int32_t w = (int32_t) v; memcpy(out, &w, 4); break; }
        case GM_INT16: case GM_UINT16: { int16_t w = (int16_t) v; memcpy(out, &w, 2); break; }
        default: out[0] = (uint8_t) (v != 0);
    }
}

// one level of every request in the batch; returns the status of the batch
static int sv_level(gm_server *s, std::vector<std::shared_ptr<SvRequest>> &batch)
{
    const int32_t zt = s->sr.ztype;
    const size_t zs = type_size(zt);
    const bool valued = !(s->sr.mult_op == GM_MULT_PAIR || s->sr.mult_op == GM_MULT_K);
    hipStream_t st = stream();
    bool want_v = false, all_host = true;
    int64_t K = 0;
    // the first-level requests' rows, stacked on the host in batch order and
    // uploaded at once (F: one entry per source, iso one; V with its values)
    std::vector<int64_t> hfp(1, 0), hfj, hvp(1, 0), hvj;
    std::vector<uint8_t> hvx;
    for (auto &r : batch)
    {
        want_v |= r->needs_visited();
        K += r->k;
        if (r->F.h) { all_host = false; continue; }
        for (int64_t t = 0; t < r->k; t++)
        {
            hfj.push_back(r->src[t]);
            hfp.push_back((int64_t) hfj.size());
            hvp.push_back(hvp.back() + (r->hvp[t + 1] - r->hvp[t]));
        }
        hvj.insert(hvj.end(), r->hvj.begin(), r->hvj.end());
        hvx.insert(hvx.end(), r->hvx.begin(), r->hvx.end());
    }
    const int64_t Kh = (int64_t) hfp.size() - 1;
    DevBuf uFp, uFj, uFx, uVp, uVj, uVx;
    if (Kh > 0)
    {
        const int64_t fz = hfp[Kh], vz = hvp[Kh];
        if (!uFp.alloc(sizeof(int64_t) * (Kh + 1)) || !uFj.alloc(sizeof(int64_t) * (fz > 0 ? fz : 1)) ||
            !uFx.alloc(zs > 0 ? zs : 8) || !uVp.alloc(sizeof(int64_t) * (Kh + 1)) ||
            !uVj.alloc(sizeof(int64_t) * (vz > 0 ? vz : 1)) || !uVx.alloc(hvx.size() > 0 ? hvx.size() : 8))
            return set_error(GM_OUT_OF_MEMORY, "server: upload of new requests");
        std::vector<uint8_t> one(zs > 0 ? zs : 1);
        sv_value(one.data(), zt, 1);
        GM_HIP(hipMemcpyAsync(uFp.ptr, hfp.data(), sizeof(int64_t) * (Kh + 1), hipMemcpyHostToDevice, st), "F rows");
        if (fz > 0) GM_HIP(hipMemcpyAsync(uFj.ptr, hfj.data(), sizeof(int64_t) * fz, hipMemcpyHostToDevice, st), "F");
        GM_HIP(hipMemcpyAsync(uFx.ptr, one.data(), zs, hipMemcpyHostToDevice, st), "F value");
        GM_HIP(hipMemcpyAsync(uVp.ptr, hvp.data(), sizeof(int64_t) * (Kh + 1), hipMemcpyHostToDevice, st), "V rows");
        if (vz > 0) GM_HIP(hipMemcpyAsync(uVj.ptr, hvj.data(), sizeof(int64_t) * vz, hipMemcpyHostToDevice, st), "V");
        if (!hvx.empty())
            GM_HIP(hipMemcpyAsync(uVx.ptr, hvx.data(), hvx.size(), hipMemcpyHostToDevice, st), "V values");
        // the pageable sources above must be read before they go out of scope
        GM_HIP(hipStreamSynchronize(st), "server upload sync");
    }
    const int64_t n = s->graph.ncols;
    DevBuf tf, Pf, Jf, Xf, tv, Pv, Jv, Xv;
    gm_csr Fs{}, Vs{};
    if (all_host)
    {
        // every request starts here: the upload is the stacked operand
        Fs.nrows = K; Fs.ncols = n; Fs.p = uFp.as<int64_t>(); Fs.j = uFj.as<int64_t>();
        Fs.x = uFx.ptr; Fs.iso = 1; Fs.type = zt;
        Vs.nrows = K; Vs.ncols = n; Vs.p = uVp.as<int64_t>(); Vs.j = uVj.as<int64_t>();
        Vs.x = uVx.ptr; Vs.iso = 0; Vs.type = zt;
        Xv.ptr = nullptr;
    }
    else
    {
        std::vector<BatchPart> pf, pv;
        int64_t row = 0, hrow = 0;
        for (auto &r : batch)
        {
            if (r->F.h)
            {
                const SvHold &F = *r->F.h, &V = *r->V.h;
                pf.push_back(BatchPart{F.M.p + r->F.row0, F.M.j, F.M.x, r->k, row, 0, 0, F.M.iso, 0});
                pv.push_back(BatchPart{V.M.p + r->V.row0, V.M.j, V.M.x, r->k, row, 0, 0, V.M.iso, 0});
            }
            else
            {
                pf.push_back(BatchPart{uFp.as<int64_t>() + hrow, uFj.as<int64_t>(), uFx.ptr, r->k, row, 0, 0, 1, 0});
                pv.push_back(BatchPart{uVp.as<int64_t>() + hrow, uVj.as<int64_t>(), uVx.ptr, r->k, row, 0, 0, 0, 0});
                hrow += r->k;
            }
            row += r->k;
        }
        GM_TRY(bt_stack(pf, K, n, valued, zt, tf, Pf, Jf, Xf, &Fs));
        GM_TRY(bt_stack(pv, K, n, want_v, zt, tv, Pv, Jv, Xv, &Vs));
    }
    if (!valued) Fs.x = nullptr;
    auto C = std::make_shared<SvHold>();
    GM_TRY(gm_saxpy3(&C->M, &Vs, true, true, &Fs, &s->graph, &s->sr, nullptr));
    std::shared_ptr<SvHold> Vn;
    if (want_v)
    {
        // V + C: the product is outside V (the mask), so this is a merge of
        // disjoint rows; FIRST keeps V's parents on an intersection
        Vn = std::make_shared<SvHold>();
        gm_csr vc = Vs, tc{};
        if (!all_host) vc.x = Xv.ptr;
        vc.iso = 0;
        tc.nrows = C->M.nrows; tc.ncols = C->M.ncols; tc.p = C->M.p; tc.j = C->M.j;
        tc.x = C->M.x; tc.iso = C->M.iso; tc.type = zt;
        GM_TRY(gm_accum_mask(&Vn->M, &vc, nullptr, false, true, false, &tc, 6));
    }
    // the result's row pointers and the next level's flop bound per row
    std::vector<int64_t> hp(K + 1), hf(K > 0 ? K : 1);
    DevBuf rf;
    if (!rf.alloc(sizeof(int64_t) * (K > 0 ? K : 1)))
        return set_error(GM_OUT_OF_MEMORY, "server: flop bound workspace");
    if (K > 0)
    {
        int64_t g = (K * 64 + 255) / 256;
        if (g > (int64_t) cu_count() * 16) g = (int64_t) cu_count() * 16;
        k_sv_rowflops<<<(unsigned) g, 256, 0, st>>>(C->M.p, C->M.j, s->graph.p, K, rf.as<int64_t>());
        GM_HIP(hipGetLastError(), "k_sv_rowflops");
        GM_HIP(hipMemcpyAsync(hf.data(), rf.ptr, sizeof(int64_t) * K, hipMemcpyDeviceToHost, st), "flops");
    }
    GM_HIP(hipMemcpyAsync(hp.data(), C->M.p, sizeof(int64_t) * (K + 1), hipMemcpyDeviceToHost, st),
           "result row pointers");
    GM_HIP(hipStreamSynchronize(st), "server level sync");
    C->M.type = zt;
    if (Vn) Vn->M.type = zt;
    int64_t row = 0;
    for (auto &r : batch)
    {
        r->C = SvSlice{C, row, hp[row + r->k] - hp[row]};
        int64_t f = 0;
        for (int64_t t = 0; t < r->k; t++) f += hf[row + t];
        // the next frontier is this level's result (values included); the
        // visited set grows by it
        r->F = r->C;
        r->flops = f;
        r->hvp = std::vector<int64_t>();
        r->hvj = std::vector<int64_t>();
        r->hvx = std::vector<uint8_t>();
        if (Vn)
        {
            // rows of V + C: V's entries plus C's (disjoint)
            r->V = SvSlice{Vn, row, r->V.nnz + r->C.nnz};
        }
        row += r->k;
    }
    (void) zs;
    return GM_SUCCESS;
}

// Admits one batch from the queue and runs its level.  block: wait for work
// (a worker thread); otherwise return 0 at once when nothing is pending.
// Returns the number of requests in the batch, or -1 when the server stops.
static int64_t sv_run_one(gm_server *s, bool block)
{
    {
        std::vector<std::shared_ptr<SvRequest>> batch;
        // HBM still free for this batch (driver-free plus the arena's cached blocks)
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { (void) hipGetLastError(); free_b = 0; }
        const double avail = (double) free_b + (double) gm_arena_bytes_cached();
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (block) s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
            if (s->pending.empty()) return block ? -1 : 0;     // stop requested and drained
            double budget = (double) s->cfg.mem_budget;
            if (avail > 0 && avail * 0.5 < budget) budget = avail * 0.5;
            // admission in queue order by workspace bytes and flop bound
            int64_t flops = 0;
            double bytes = 0;
            const int64_t cap = std::min<int64_t>(s->cap_reqs, s->cfg.max_batch_reqs);
            while (!s->pending.empty() && (int64_t) batch.size() < cap)
            {
                auto &r = s->pending.front();
                const double rb = (double) sv_bytes(*r);
                if (!batch.empty() && (flops + r->flops > s->cfg.flop_budget || bytes + rb > budget)) break;
                flops += r->flops;
                bytes += rb;
                r->state = 1;
                batch.push_back(r);
                s->pending.pop_front();
            }
            s->in_flight++;
            if (s->in_flight > s->max_in_flight) s->max_in_flight = s->in_flight;
        }
        const double t0 = now_us();
        for (auto &r : batch)
            if (r->t_start < 0) r->t_start = t0;
        std::vector<int64_t> fl(batch.size());
        for (size_t t = 0; t < batch.size(); t++) fl[t] = batch[t]->flops;
        const int rc = sv_level(s, batch);
        const double t1 = now_us();
        bool requeued = false;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            const int64_t bid = s->next_batch++;
            if (rc == GM_OUT_OF_MEMORY && batch.size() > 1)
            {
                // queue the batch again, in order, at the front; the next
                // admissions take at most half of it
                s->cap_reqs = std::max<int64_t>(1, (int64_t) batch.size() / 2);
                s->oom_splits++;
                for (size_t t = batch.size(); t-- > 0;)
                {
                    batch[t]->state = 0;
                    s->pending.push_front(batch[t]);
                }
                requeued = true;
            }
            else
            {
                if (rc == GM_SUCCESS && s->cap_reqs < s->cfg.max_batch_reqs)
                    s->cap_reqs = std::min<int64_t>(s->cfg.max_batch_reqs, 2 * s->cap_reqs);
                s->batches++;
                s->batched_reqs += (int64_t) batch.size();
                s->levels += (rc == GM_SUCCESS) ? (int64_t) batch.size() : 0;
                for (size_t t = batch.size(); t-- > 0;)
                {
                    auto &r = batch[t];
                    r->batch_id = bid;
                    r->batch_size = (int64_t) batch.size();
                    if (rc == GM_SUCCESS)
                    {
                        s->flops_done += (double) fl[t];
                        r->flops_total += fl[t];
                        r->steps++;
                        const bool more = (r->max_steps <= 0 || r->steps < r->max_steps) && r->C.nnz > 0;
                        if (more)
                        {
                            r->state = 0;
                            s->pending.push_front(r);     // next level, ahead of new arrivals
                            requeued = true;
                            continue;
                        }
                    }
                    r->t_done = t1;
                    r->state = 2;
                    r->rc = rc;
                    if (!r->keep_visited && rc == GM_SUCCESS) r->V = SvSlice{};
                    r->F = SvSlice{};
                    s->completed++;
                }
            }
            s->in_flight--;
        }
        if (requeued) s->cv_work.notify_all();
        s->cv_done.notify_all();
        return (int64_t) batch.size();
    }
}

static void sv_worker(gm_server *s)
{
    while (sv_run_one(s, true) >= 0) { }
}

// host checks of a visited set: vis_p[0] = 0, monotone, columns in [0, n)
// and strictly increasing in every row
static int sv_check_visited(const int64_t *vp, const int64_t *vj, int64_t k, int64_t n)
{
    if (vp[0] != 0) return set_error(GM_INVALID_VALUE, "gm_server_submit: vis_p[0] must be 0");
    for (int64_t t = 0; t < k; t++)
    {
        if (vp[t + 1] < vp[t]) return set_error(GM_INVALID_VALUE, "gm_server_submit: vis_p not monotone");
        for (int64_t q = vp[t]; q < vp[t + 1]; q++)
        {
            if (vj[q] < 0 || vj[q] >= n)
                return set_error(GM_INVALID_VALUE, "gm_server_submit: visited vertex out of range");
            if (q > vp[t] && vj[q] <= vj[q - 1])
                return set_error(GM_INVALID_VALUE, "gm_server_submit: visited row %lld not sorted",
                                 (long long) t);
        }
    }
    return GM_SUCCESS;
}

extern "C" {

gm_server *gm_server_create(const gm_csr *graph, const gm_semiring *sr, const gm_server_config *cfg)
{
    if (!graph || !sr || !graph->p) { set_error(GM_INVALID_VALUE, "gm_server_create: NULL"); return nullptr; }
    if (ensure_init() != GM_SUCCESS) return nullptr;
    if (graph->nrows != graph->ncols || !gm_saxpy3_branch(graph, true, true, graph, graph, sr))
    {
        set_error(GM_NO_VALUE, "gm_server_create: needs a square graph and a semiring the saxpy3 kernels implement");
        return nullptr;
    }
    const bool valued = !(sr->mult_op == GM_MULT_PAIR || sr->mult_op == GM_MULT_K);
    if (valued && graph->type != sr->ztype)
    {
        set_error(GM_NO_VALUE, "gm_server_create: the graph's values must have the semiring's type");
        return nullptr;
    }
    gm_server *s = new gm_server();
    s->graph = *graph;
    s->sr = *sr;
    gm_server_config d{2, 256, (int64_t) 1 << 24, 0};
    s->cfg = cfg ? *cfg : d;
    if (s->cfg.workers < 0) s->cfg.workers = d.workers;
    if (s->cfg.max_batch_reqs <= 0) s->cfg.max_batch_reqs = d.max_batch_reqs;
    if (s->cfg.flop_budget <= 0) s->cfg.flop_budget = d.flop_budget;
    if (s->cfg.mem_budget <= 0) s->cfg.mem_budget = device_hbm_bytes() / 4;
    // degrees for admission
    std::vector<int64_t> hp(graph->nrows + 1);
    if (gm_memcpy_d2h(hp.data(), graph->p, sizeof(int64_t) * (graph->nrows + 1)) != GM_SUCCESS)
    {
        delete s;
        return nullptr;
    }
    s->deg.resize(graph->nrows);
    for (int64_t i = 0; i < graph->nrows; i++) s->deg[i] = hp[i + 1] - hp[i];
    s->cap_reqs = s->cfg.max_batch_reqs;
    for (int w = 0; w < s->cfg.workers; w++) s->workers.emplace_back(sv_worker, s);
    return s;
}

int64_t gm_server_submit_bfs(gm_server *s, const int64_t *sources, int64_t k, const int64_t *vis_p,
                             const int64_t *vis_j, const gm_server_request_opts *opts)
{
    if (!s || (!sources && k > 0) || k < 0) return set_error(GM_INVALID_VALUE, "gm_server_submit: bad args");
    if ((vis_p == nullptr) != (vis_j == nullptr))
        return set_error(GM_INVALID_VALUE, "gm_server_submit: vis_p and vis_j go together");
    const int64_t n = (int64_t) s->deg.size();
    auto r = std::make_shared<SvRequest>();
    r->src.assign(sources, sources + k);
    r->k = k;
    int64_t f = 0;
    for (int64_t t = 0; t < k; t++)
    {
        if (sources[t] < 0 || sources[t] >= n)
            return set_error(GM_INVALID_VALUE, "gm_server_submit: source out of range");
        f += s->deg[sources[t]];
    }
    if (vis_p) GM_TRY(sv_check_visited(vis_p, vis_j, k, n));
    r->flops = f;
    if (opts)
    {
        r->max_steps = opts->max_steps;
        r->keep_visited = opts->keep_visited != 0;
    }
    // frontier F (the sources) and visited set V stay on the host until the
    // batch that runs the request's first level uploads them with the others
    const int32_t zt = s->sr.ztype;
    const size_t zs = type_size(zt);
    std::vector<int64_t> fp(k + 1), fj(r->src);
    for (int64_t t = 0; t <= k; t++) fp[t] = t;
    r->F.nnz = k;
    // V's values: the parent, i.e. the source for its own row (positional
    // semirings) or one; -1 for the entries the caller marked visited
    const bool positional = s->sr.mult_op == GM_MULT_K;
    std::vector<int64_t> vp(k + 1), vj;
    std::vector<uint8_t> vx;
    if (vis_p)
    {
        vp.assign(vis_p, vis_p + k + 1);
        vj.assign(vis_j, vis_j + vis_p[k]);
        vx.resize(zs * vj.size());
        for (size_t q = 0; q < vj.size(); q++) sv_value(vx.data() + q * zs, zt, -1);
        for (int64_t t = 0; t < k; t++)
            for (int64_t q = vp[t]; q < vp[t + 1]; q++)
                if (vj[q] == r->src[t]) sv_value(vx.data() + q * zs, zt, positional ? r->src[t] : 1);
    }
    else
    {
        vp = fp;
        vj = fj;
        vx.resize(zs * k);
        for (int64_t t = 0; t < k; t++) sv_value(vx.data() + t * zs, zt, positional ? r->src[t] : 1);
    }
    r->V.nnz = vp[k];
    r->hvp = std::move(vp);
    r->hvj = std::move(vj);
    r->hvx = std::move(vx);
    r->t_arrive = now_us();
    int64_t id;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return set_error(GM_INVALID_VALUE, "gm_server_submit: server stopping");
        id = s->next_id++;
        r->id = id;
        s->reqs[id] = r;
        s->pending.push_back(r);
    }
    s->cv_work.notify_one();
    return id;
}

int64_t gm_server_submit(gm_server *s, const int64_t *sources, int64_t k, const int64_t *vis_p,
                         const int64_t *vis_j)
{
    return gm_server_submit_bfs(s, sources, k, vis_p, vis_j, nullptr);
}

int64_t gm_server_step(gm_server *s)
{
    if (!s) return set_error(GM_INVALID_VALUE, "gm_server_step: NULL");
    const int64_t n = sv_run_one(s, false);
    return n < 0 ? 0 : n;
}

int gm_server_wait(gm_server *s, int64_t id, gm_server_result *out)
{
    if (!s || !out) return set_error(GM_INVALID_VALUE, "gm_server_wait: NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return set_error(GM_INVALID_VALUE, "gm_server_wait: unknown request");
    auto r = it->second;
    if (s->workers.empty() && r->state != 2)
        return set_error(GM_INVALID_VALUE, "gm_server_wait: request not complete (no worker threads: "
                                           "run gm_server_step)");
    s->cv_done.wait(lk, [&r] { return r->state == 2; });
    memset(out, 0, sizeof(*out));
    out->status = r->rc;
    out->k = r->k;
    out->flops = r->flops_total;
    out->latency_us = r->t_done - r->t_arrive;
    out->queue_us = r->t_start - r->t_arrive;
    out->batch_id = r->batch_id;
    out->batch_size = r->batch_size;
    out->nvals = r->C.h ? r->C.nnz : 0;
    out->steps = r->steps;
    out->vnvals = r->V.h ? r->V.nnz : 0;
    return out->status;
}

// rows [row0, row0 + k) of a held CSR to host arrays (row pointers from 0)
static int sv_copy_rows(const SvSlice &sl, int64_t k, int64_t *Cp, int64_t *Cj, void *Cx)
{
    const gm_csr_out &C = sl.h->M;
    GM_TRY(gm_memcpy_d2h(Cp, C.p + sl.row0, sizeof(int64_t) * (k + 1)));
    const int64_t base = Cp[0], nz = Cp[k] - Cp[0];
    for (int64_t t = 0; t <= k; t++) Cp[t] -= base;
    if (Cj && nz > 0) GM_TRY(gm_memcpy_d2h(Cj, C.j + base, sizeof(int64_t) * nz));
    const size_t xs = type_size(C.type);
    if (Cx && nz > 0)
    {
        if (C.iso)
        {
            std::vector<uint8_t> v(xs);
            GM_TRY(gm_memcpy_d2h(v.data(), C.x, xs));
            for (int64_t t = 0; t < nz; t++) memcpy((uint8_t *) Cx + t * xs, v.data(), xs);
        }
        else GM_TRY(gm_memcpy_d2h(Cx, (const uint8_t *) C.x + base * xs, xs * nz));
    }
    return GM_SUCCESS;
}

static std::shared_ptr<SvRequest> sv_done(gm_server *s, int64_t id)
{
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second->state != 2 || it->second->rc != GM_SUCCESS) return nullptr;
    return it->second;
}

int gm_server_fetch(gm_server *s, int64_t id, int64_t *Cp, int64_t *Cj, void *Cx)
{
    if (!s || !Cp) return set_error(GM_INVALID_VALUE, "gm_server_fetch: NULL");
    auto r = sv_done(s, id);
    if (!r || !r->C.h) return set_error(GM_INVALID_VALUE, "gm_server_fetch: request not complete");
    return sv_copy_rows(r->C, r->k, Cp, Cj, Cx);
}

int gm_server_fetch_visited(gm_server *s, int64_t id, int64_t *Vp, int64_t *Vj, void *Vx)
{
    if (!s || !Vp) return set_error(GM_INVALID_VALUE, "gm_server_fetch_visited: NULL");
    auto r = sv_done(s, id);
    if (!r) return set_error(GM_INVALID_VALUE, "gm_server_fetch_visited: request not complete");
    if (!r->V.h)
        return set_error(GM_INVALID_VALUE, "gm_server_fetch_visited: the request did not keep its visited set");
    return sv_copy_rows(r->V, r->k, Vp, Vj, Vx);
}

int gm_server_release(gm_server *s, int64_t id)
{
    if (!s) return GM_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(s->mu);
    s->reqs.erase(id);                 // the held results go with their last request
    return GM_SUCCESS;
}

int gm_server_stats(gm_server *s, gm_server_counters *st)
{
    if (!s || !st) return GM_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(s->mu);
    st->completed = s->completed;
    st->batches = s->batches;
    st->pending = (int64_t) s->pending.size();
    st->max_in_flight = s->max_in_flight;
    st->mean_batch = s->batches ? (double) s->batched_reqs / (double) s->batches : 0.0;
    st->flops = s->flops_done;
    st->levels = s->levels;
    st->oom_splits = s->oom_splits;
    return GM_SUCCESS;
}

void gm_server_destroy(gm_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->reqs.clear();
        s->pending.clear();
    }
    (void) gm_synchronize();
    delete s;
}

} // extern "C"
