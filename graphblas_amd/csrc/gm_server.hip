// gm_server.hip: continuous-batching request loop for frontier expansion.
//
// A server owns one device-resident graph A (CSR in HBM).  Each request is a
// multi-source frontier step C<!V> = F*A: F has one row per source (F(r,s_r)
// = 1) and V, the visited set, is the request's own k x n matrix (by default
// V = F).  With the ANY_SECONDI semiring C(r,i) is the parent s_r of every
// newly reached vertex i (the reference's BFS step, GraphBLAS/@GrB/bfs.m).
//
// Worker threads (one HIP stream each, default priority) take pending
// requests in arrival order and admit them into a batch while the batch's
// flop upper bound -- sum over its sources of deg(s), the bound on the
// symbolic phase's work and on C's size -- stays within the budget (a single
// request larger than the budget runs alone).  The admitted requests are
// stacked by rows and computed by one saxpy3 call; a request that arrives
// while batches are on the GPU joins the next batch as soon as a worker is
// free, so several batches overlap on the device.  Results stay in HBM; the
// caller waits for a request, reads its latency and copies its rows out.

#include "gm_internal.h"

#include <atomic>
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

struct SvBatch
{
    gm_csr_out C;
    int64_t id = 0;
    int64_t size = 0;
    ~SvBatch()
    {
        gm_free(C.p);
        gm_free(C.j);
        gm_free(C.x);
    }
};

struct SvRequest
{
    int64_t id = 0;
    std::vector<int64_t> src;
    std::vector<int64_t> vp, vj;      // optional visited set (k+1, nnz)
    int64_t flops = 0;
    double t_arrive = 0, t_start = 0, t_done = 0;
    int state = 0;                    // 0 pending, 1 running, 2 finished
    int rc = 0;                       // result of the batch's saxpy3 call
    std::shared_ptr<SvBatch> batch;
    int64_t row0 = 0;
    int64_t nvals = 0;                // entries of the request's rows
};

} // namespace gm

using namespace gm;

struct gm_server
{
    gm_csr graph;
    gm_semiring sr;
    gm_server_config cfg;
    void *one = nullptr;              // device value 1 of the operand type (F is iso)
    std::vector<int64_t> deg;         // host copy of the graph's row lengths
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<SvRequest>> pending;
    std::unordered_map<int64_t, std::shared_ptr<SvRequest>> reqs;
    std::vector<std::thread> workers;
    bool stop = false;
    int64_t next_id = 1, next_batch = 1;
    int64_t completed = 0, batches = 0, batched_reqs = 0, in_flight = 0, max_in_flight = 0;
    double flops_done = 0;
};

static void sv_worker(gm_server *s)
{
    while (true)
    {
        std::vector<std::shared_ptr<SvRequest>> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
            if (s->pending.empty()) return;          // stop requested and drained
            // admission by flop upper bound, in arrival order
            int64_t flops = 0;
            while (!s->pending.empty() && (int64_t) batch.size() < s->cfg.max_batch_reqs)
            {
                auto &r = s->pending.front();
                if (!batch.empty() && flops + r->flops > s->cfg.flop_budget) break;
                flops += r->flops;
                r->state = 1;
                batch.push_back(r);
                s->pending.pop_front();
            }
            s->in_flight++;
            if (s->in_flight > s->max_in_flight) s->max_in_flight = s->in_flight;
        }
        const double t0 = now_us();
        // stacked F and V (host), uploaded to HBM
        int64_t K = 0;
        bool any_vis = false;
        for (auto &r : batch) { K += (int64_t) r->src.size(); any_vis |= !r->vp.empty(); }
        std::vector<int64_t> Fp(K + 1), Fj(K > 0 ? K : 1);
        int64_t row = 0;
        for (auto &r : batch)
        {
            r->row0 = row;
            for (int64_t s0 : r->src) { Fp[row] = row; Fj[row] = s0; row++; }
        }
        Fp[K] = K;
        std::vector<int64_t> Vp, Vj;
        if (any_vis)
        {
            Vp.reserve(K + 1);
            Vp.push_back(0);
            for (auto &r : batch)
            {
                const int64_t k = (int64_t) r->src.size();
                for (int64_t t = 0; t < k; t++)
                {
                    if (!r->vp.empty())
                        for (int64_t q = r->vp[t]; q < r->vp[t + 1]; q++) Vj.push_back(r->vj[q]);
                    else
                        Vj.push_back(r->src[t]);
                    Vp.push_back((int64_t) Vj.size());
                }
            }
        }
        int rc = GM_SUCCESS;
        DevBuf dFp, dFj, dVp, dVj;
        if (!dFp.alloc(sizeof(int64_t) * (K + 1)) || !dFj.alloc(sizeof(int64_t) * Fj.size()) ||
            (any_vis && (!dVp.alloc(sizeof(int64_t) * Vp.size()) ||
                         !dVj.alloc(sizeof(int64_t) * (Vj.size() > 0 ? Vj.size() : 1)))))
            rc = GM_OUT_OF_MEMORY;
        if (rc == GM_SUCCESS) rc = gm_memcpy_h2d(dFp.ptr, Fp.data(), sizeof(int64_t) * (K + 1));
        if (rc == GM_SUCCESS) rc = gm_memcpy_h2d(dFj.ptr, Fj.data(), sizeof(int64_t) * K);
        if (rc == GM_SUCCESS && any_vis)
        {
            rc = gm_memcpy_h2d(dVp.ptr, Vp.data(), sizeof(int64_t) * Vp.size());
            if (rc == GM_SUCCESS && !Vj.empty())
                rc = gm_memcpy_h2d(dVj.ptr, Vj.data(), sizeof(int64_t) * Vj.size());
        }
        auto res = std::make_shared<SvBatch>();
        memset(&res->C, 0, sizeof(res->C));
        if (rc == GM_SUCCESS)
        {
            gm_csr F{K, s->graph.nrows, dFp.as<int64_t>(), dFj.as<int64_t>(), s->one, 1, s->graph.type};
            gm_csr V = F;
            if (any_vis) { V.p = dVp.as<int64_t>(); V.j = dVj.as<int64_t>(); V.x = nullptr; V.iso = 0; }
            V.ncols = s->graph.ncols;
            rc = gm_saxpy3(&res->C, &V, true, true, &F, &s->graph, &s->sr, nullptr);
        }
        // the batch's row pointers, once, so that waiting for a request needs
        // no device access
        std::vector<int64_t> hcp;
        if (rc == GM_SUCCESS)
        {
            hcp.resize(K + 1);
            rc = gm_memcpy_d2h(hcp.data(), res->C.p, sizeof(int64_t) * (K + 1));
        }
        const double t1 = now_us();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            res->id = s->next_batch++;
            res->size = (int64_t) batch.size();
            for (auto &r : batch)
            {
                r->t_start = t0;
                r->t_done = t1;
                r->state = 2;
                r->rc = rc;
                if (rc == GM_SUCCESS)
                {
                    r->batch = res;
                    r->nvals = hcp[r->row0 + (int64_t) r->src.size()] - hcp[r->row0];
                }
                s->flops_done += (double) r->flops;
            }
            s->completed += (int64_t) batch.size();
            s->batches++;
            s->batched_reqs += (int64_t) batch.size();
            s->in_flight--;
        }
        s->cv_done.notify_all();
    }
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
    gm_server *s = new gm_server();
    s->graph = *graph;
    s->sr = *sr;
    gm_server_config d{2, 256, (int64_t) 1 << 24, 0};
    s->cfg = cfg ? *cfg : d;
    if (s->cfg.workers <= 0) s->cfg.workers = d.workers;
    if (s->cfg.max_batch_reqs <= 0) s->cfg.max_batch_reqs = d.max_batch_reqs;
    if (s->cfg.flop_budget <= 0) s->cfg.flop_budget = d.flop_budget;
    // degrees for admission
    std::vector<int64_t> hp(graph->nrows + 1);
    if (gm_memcpy_d2h(hp.data(), graph->p, sizeof(int64_t) * (graph->nrows + 1)) != GM_SUCCESS)
    {
        delete s;
        return nullptr;
    }
    s->deg.resize(graph->nrows);
    for (int64_t i = 0; i < graph->nrows; i++) s->deg[i] = hp[i + 1] - hp[i];
    // the iso value 1 of F in the graph's value type
    const size_t ts = type_size(graph->type) ? type_size(graph->type) : 8;
    s->one = gm_malloc(ts);
    uint8_t one[8] = {0};
    switch (graph->type)
    {
        case GM_FP64: { double v = 1.0; memcpy(one, &v, 8); break; }
        case GM_FP32: { float v = 1.0f; memcpy(one, &v, 4); break; }
        default: one[0] = 1;
    }
    if (!s->one || gm_memcpy_h2d(s->one, one, ts) != GM_SUCCESS)
    {
        gm_free(s->one);
        delete s;
        return nullptr;
    }
    for (int w = 0; w < s->cfg.workers; w++) s->workers.emplace_back(sv_worker, s);
    return s;
}

int64_t gm_server_submit(gm_server *s, const int64_t *sources, int64_t k, const int64_t *vis_p,
                         const int64_t *vis_j)
{
    if (!s || (!sources && k > 0) || k < 0) return set_error(GM_INVALID_VALUE, "gm_server_submit: bad args");
    auto r = std::make_shared<SvRequest>();
    r->src.assign(sources, sources + k);
    int64_t f = 0;
    for (int64_t t = 0; t < k; t++)
    {
        if (sources[t] < 0 || sources[t] >= (int64_t) s->deg.size())
            return set_error(GM_INVALID_VALUE, "gm_server_submit: source out of range");
        f += s->deg[sources[t]];
    }
    r->flops = f;
    if (vis_p && vis_j)
    {
        r->vp.assign(vis_p, vis_p + k + 1);
        r->vj.assign(vis_j, vis_j + vis_p[k]);
    }
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

int gm_server_wait(gm_server *s, int64_t id, gm_server_result *out)
{
    if (!s || !out) return set_error(GM_INVALID_VALUE, "gm_server_wait: NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return set_error(GM_INVALID_VALUE, "gm_server_wait: unknown request");
    auto r = it->second;
    s->cv_done.wait(lk, [&r] { return r->state == 2; });
    memset(out, 0, sizeof(*out));
    out->status = r->rc;
    out->k = (int64_t) r->src.size();
    out->flops = r->flops;
    out->latency_us = r->t_done - r->t_arrive;
    out->queue_us = r->t_start - r->t_arrive;
    out->batch_id = r->batch ? r->batch->id : 0;
    out->batch_size = r->batch ? r->batch->size : 0;
    out->nvals = r->nvals;
    return out->status;
}

int gm_server_fetch(gm_server *s, int64_t id, int64_t *Cp, int64_t *Cj, void *Cx)
{
    if (!s || !Cp) return set_error(GM_INVALID_VALUE, "gm_server_fetch: NULL");
    std::shared_ptr<SvRequest> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end() || it->second->state != 2 || it->second->rc != GM_SUCCESS)
            return set_error(GM_INVALID_VALUE, "gm_server_fetch: request not complete");
        r = it->second;
    }
    const int64_t k = (int64_t) r->src.size();
    const gm_csr_out &C = r->batch->C;
    GM_TRY(gm_memcpy_d2h(Cp, C.p + r->row0, sizeof(int64_t) * (k + 1)));
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

int gm_server_release(gm_server *s, int64_t id)
{
    if (!s) return GM_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(s->mu);
    s->reqs.erase(id);                 // the batch result goes when its last request does
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
    gm_free(s->one);
    delete s;
}

} // extern "C"
