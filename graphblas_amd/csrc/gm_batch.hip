// gm_batch.hip: many independent mxm requests in one launch sequence.
//
// dot3 batch: requests C_r<M_r> = A_r * B_r' (one semiring) are stacked
// block-diagonally: the rows of M_r and A_r start at row_base[r], the rows of
// B_r (= the columns of M_r and C_r) at col_base[r], and the inner columns of
// A_r and B_r at inner_base[r].  One dot3 over the stacked operands computes
// every request; no request's entries can meet another's, so each result is
// exactly what the request computes alone.  The persistent dot3 kernels'
// device bins are the shared work queue of all requests.
//
// saxpy3 batch: requests C_r<M_r> = A_r * B with one shared B (a served graph)
// are stacked by rows of A_r (and M_r); the column space is B's.
//
// The stacked result C is returned as one CSR plus the row base of every
// request (and, for dot3, the column base to subtract); gm_batch_split
// materialises per-request CSR matrices on the device if a caller needs them.

#include "gm_internal.h"
#include "gm_block.h"
#include <vector>

namespace gm
{

// nnz of every part (its last row pointer)
__global__ void k_bt_nnz(const BatchPart *parts, int64_t nparts, int64_t *nnz)
{
    for (int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < nparts;
         r += (int64_t) gridDim.x * blockDim.x)
        nnz[r] = parts[r].p[parts[r].nrows] - parts[r].p[0];
}

__global__ void k_bt_set_base(BatchPart *parts, int64_t nparts, const int64_t *ebase)
{
    for (int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < nparts;
         r += (int64_t) gridDim.x * blockDim.x)
        parts[r].ent_base = ebase[r];
}

// part owning stacked row / entry t: last r with base[r] <= t
template <int64_t BatchPart::*FIELD>
__device__ __forceinline__ int64_t bt_find_in(const BatchPart *parts, int64_t lo, int64_t hi, int64_t t)
{
    while (lo < hi)
    {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (parts[mid].*FIELD <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int64_t BatchPart::*FIELD>
__device__ __forceinline__ int64_t bt_find(const BatchPart *parts, int64_t nparts, int64_t t)
{
    return bt_find_in<FIELD>(parts, 0, nparts - 1, t);
}

// A workgroup takes BT_TILE consecutive rows / entries: the parts owning
// the tile's first and last element are searched once, and each element
// searches only between them (usually one part)
constexpr int BT_TILE = 256 * 16;

// stacked row pointers
__global__ void k_bt_rows(const BatchPart *parts, int64_t nparts, int64_t total_rows,
                          int64_t total_nnz, int64_t *P)
{
    __shared__ int64_t s_r[2];
    for (int64_t t0 = (int64_t) blockIdx.x * BT_TILE; t0 < total_rows; t0 += (int64_t) gridDim.x * BT_TILE)
    {
        const int64_t t1 = (t0 + BT_TILE < total_rows) ? t0 + BT_TILE : total_rows;
        if (threadIdx.x < 2)
            s_r[threadIdx.x] = bt_find<&BatchPart::row_base>(parts, nparts, threadIdx.x == 0 ? t0 : t1 - 1);
        __syncthreads();
        const int64_t r0 = s_r[0], r1 = s_r[1];
        __syncthreads();
        for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x)
        {
            const int64_t r = bt_find_in<&BatchPart::row_base>(parts, r0, r1, t);
            const BatchPart &b = parts[r];
            P[t] = b.ent_base + (b.p[t - b.row_base] - b.p[0]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) P[total_rows] = total_nnz;
}

// stacked column indices (+ values of size xs, iso parts expanded)
__global__ void k_bt_ents(const BatchPart *parts, int64_t nparts, int64_t total_nnz, int64_t *J,
                          void *X, int xs)
{
    __shared__ int64_t s_r[2];
    for (int64_t t0 = (int64_t) blockIdx.x * BT_TILE; t0 < total_nnz; t0 += (int64_t) gridDim.x * BT_TILE)
    {
        const int64_t t1 = (t0 + BT_TILE < total_nnz) ? t0 + BT_TILE : total_nnz;
        if (threadIdx.x < 2)
            s_r[threadIdx.x] = bt_find<&BatchPart::ent_base>(parts, nparts, threadIdx.x == 0 ? t0 : t1 - 1);
        __syncthreads();
        const int64_t r0 = s_r[0], r1 = s_r[1];
        __syncthreads();
        for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x)
        {
            const int64_t r = bt_find_in<&BatchPart::ent_base>(parts, r0, r1, t);
            const BatchPart &b = parts[r];
            const int64_t q = b.p[0] + (t - b.ent_base);
            J[t] = b.j[q] + b.col_add;
            if (X)
            {
                const int64_t qs = b.iso ? 0 : q;
                if (xs == 8) ((uint64_t *) X)[t] = ((const uint64_t *) b.x)[qs];
                else if (xs == 4) ((uint32_t *) X)[t] = ((const uint32_t *) b.x)[qs];
                else if (xs == 2) ((uint16_t *) X)[t] = ((const uint16_t *) b.x)[qs];
                else ((uint8_t *) X)[t] = ((const uint8_t *) b.x)[qs];
            }
        }
    }
}

// Stacked copy of parts; arrays allocated here (released by the caller's
// DevBufs).  Returns the stacked gm_csr view.
int bt_stack(const std::vector<BatchPart> &hparts, int64_t total_rows, int64_t ncols,
                    bool values, int32_t type, DevBuf &dparts, DevBuf &P, DevBuf &J, DevBuf &X,
                    gm_csr *out)
{
    hipStream_t st = stream();
    const int64_t np = (int64_t) hparts.size();
    const size_t xs = type_size(type);
    DevBuf nnz;
    if (!dparts.alloc(sizeof(BatchPart) * np) || !nnz.alloc(sizeof(int64_t) * (np + 1)))
        return set_error(GM_OUT_OF_MEMORY, "batch: part table");
    GM_HIP(hipMemcpyAsync(dparts.ptr, hparts.data(), sizeof(BatchPart) * np, hipMemcpyHostToDevice, st),
           "batch part table");
    const unsigned gp = (unsigned) ((np + 255) / 256 < 1024 ? (np + 255) / 256 : 1024);
    k_bt_nnz<<<gp ? gp : 1, 256, 0, st>>>(dparts.as<BatchPart>(), np, nnz.as<int64_t>());
    GM_HIP(hipGetLastError(), "k_bt_nnz");
    GM_TRY(scan_exclusive_i64(nnz.as<int64_t>(), nnz.as<int64_t>(), np));
    k_bt_set_base<<<gp ? gp : 1, 256, 0, st>>>(dparts.as<BatchPart>(), np, nnz.as<int64_t>());
    int64_t *h = (int64_t *) pinned_scratch(sizeof(int64_t));
    GM_HIP(hipMemcpyAsync(h, nnz.as<int64_t>() + np, sizeof(int64_t), hipMemcpyDeviceToHost, st),
           "batch nnz");
    GM_HIP(hipStreamSynchronize(st), "batch nnz sync");
    const int64_t total_nnz = *h;
    if (!P.alloc(sizeof(int64_t) * (total_rows + 1)) || !J.alloc(sizeof(int64_t) * (total_nnz > 0 ? total_nnz : 1)) ||
        (values && !X.alloc(xs * (total_nnz > 0 ? total_nnz : 1))))
        return set_error(GM_OUT_OF_MEMORY, "batch: stacked operand (%lld entries)", (long long) total_nnz);
    {
        ProfScope ps("batch_stack");
        const int64_t cap = (int64_t) cu_count() * 16;
        int64_t g1 = (total_rows + BT_TILE - 1) / BT_TILE;
        if (g1 > cap) g1 = cap;
        k_bt_rows<<<(unsigned) (g1 > 0 ? g1 : 1), 256, 0, st>>>(dparts.as<BatchPart>(), np, total_rows,
                                                              total_nnz, P.as<int64_t>());
        int64_t g2 = (total_nnz + BT_TILE - 1) / BT_TILE;
        if (g2 > cap) g2 = cap;
        if (total_nnz > 0)
            k_bt_ents<<<(unsigned) g2, 256, 0, st>>>(dparts.as<BatchPart>(), np, total_nnz, J.as<int64_t>(),
                                                     values ? X.ptr : nullptr, (int) xs);
        GM_HIP(hipGetLastError(), "batch stack kernels");
    }
    out->nrows = total_rows;
    out->ncols = ncols;
    out->p = P.as<int64_t>();
    out->j = J.as<int64_t>();
    out->x = values ? X.ptr : nullptr;
    out->iso = 0;
    out->type = type;
    return GM_SUCCESS;
}

static bool same_csr(const gm_csr &a, const gm_csr &b)
{
    return a.p == b.p && a.j == b.j && a.x == b.x && a.nrows == b.nrows && a.ncols == b.ncols;
}

__global__ void k_bt_gather(const int64_t *src, const int64_t *idx, int64_t *dst, int64_t n)
{
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t) gridDim.x * blockDim.x)
        dst[t] = src[idx[t]];
}

// positional results (the inner index k) of stacked requests carry the
// request's inner offset: subtract it, row by row
__global__ void k_bt_fix_k(const int64_t *Cp, int64_t *Cx, const int64_t *rbase, const int64_t *ibase,
                           int64_t nparts, int64_t total_rows)
{
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < total_rows;
         t += (int64_t) gridDim.x * blockDim.x)
    {
        int64_t lo = 0, hi = nparts - 1;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (rbase[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const int64_t sub = ibase[lo];
        if (sub == 0) continue;
        for (int64_t q = Cp[t]; q < Cp[t + 1]; q++) Cx[q] -= sub;
    }
}

// per-request split of a stacked result
struct SplitPart
{
    int64_t *p, *j;
    void *x;
    int64_t row_base, nrows, col_sub;
};

__global__ void k_bt_split_rows(const int64_t *Cp, const SplitPart *parts, int64_t nparts,
                                int64_t total_rows)
{
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < total_rows + nparts;
         t += (int64_t) gridDim.x * blockDim.x)
    {
        if (t >= total_rows)
        {
            // the closing row pointer of every part
            const SplitPart &s = parts[t - total_rows];
            s.p[s.nrows] = Cp[s.row_base + s.nrows] - Cp[s.row_base];
            continue;
        }
        int64_t lo = 0, hi = nparts - 1;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (parts[mid].row_base <= t) lo = mid; else hi = mid - 1;
        }
        const SplitPart &s = parts[lo];
        s.p[t - s.row_base] = Cp[t] - Cp[s.row_base];
    }
}

__global__ void k_bt_split_ents(const int64_t *Cp, const int64_t *Ci, const void *Cx, int xs, int iso,
                                const SplitPart *parts, const int64_t *ebase, int64_t nparts,
                                int64_t total_nnz)
{
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < total_nnz;
         t += (int64_t) gridDim.x * blockDim.x)
    {
        int64_t lo = 0, hi = nparts - 1;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (ebase[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const SplitPart &s = parts[lo];
        const int64_t q = t - ebase[lo];
        s.j[q] = Ci[t] - s.col_sub;
        if (Cx && !iso)
        {
            if (xs == 8) ((uint64_t *) s.x)[q] = ((const uint64_t *) Cx)[t];
            else if (xs == 4) ((uint32_t *) s.x)[q] = ((const uint32_t *) Cx)[t];
            else ((uint8_t *) s.x)[q] = ((const uint8_t *) Cx)[t];
        }
    }
}

} // namespace gm

using namespace gm;

extern "C" int gm_batch_dot3(gm_csr_out *C, int64_t *row_base, int64_t *col_base,
                             const gm_dot3_request *reqs, int64_t nreq, bool mask_struct,
                             const gm_semiring *sr, const gm_dot3_params *params)
{
    if (!C || !row_base || !col_base || (!reqs && nreq > 0) || !sr || nreq < 0)
        return set_error(GM_INVALID_VALUE, "gm_batch_dot3: NULL argument");
    GM_TRY(ensure_init());
    // host-side bases; every request must be a valid dot3 on its own
    std::vector<BatchPart> pm, pa, pb;
    pm.reserve(nreq); pa.reserve(nreq); pb.reserve(nreq);
    int64_t rows = 0, bcols = 0, inner = 0;
    bool shared_all = true;
    for (int64_t r = 0; r < nreq; r++)
    {
        const gm_dot3_request &q = reqs[r];
        if (!gm_dot3_branch(&q.M, mask_struct, &q.A, &q.B, sr))
            return set_error(GM_NO_VALUE, "gm_batch_dot3: request %lld not handled", (long long) r);
        row_base[r] = rows;
        col_base[r] = bcols;
        pm.push_back(BatchPart{q.M.p, q.M.j, q.M.x, q.M.nrows, rows, bcols, 0, q.M.iso, 0});
        pa.push_back(BatchPart{q.A.p, q.A.j, q.A.x, q.A.nrows, rows, inner, 0, q.A.iso, 0});
        pb.push_back(BatchPart{q.B.p, q.B.j, q.B.x, q.B.nrows, bcols, inner, 0, q.B.iso, 0});
        if (!(same_csr(q.M, q.A) && same_csr(q.A, q.B))) shared_all = false;
        rows += q.M.nrows;
        bcols += q.B.nrows;
        inner += q.A.ncols;
    }
    row_base[nreq] = rows;
    col_base[nreq] = bcols;
    const bool valued = !(sr->mult_op == GM_MULT_PAIR || sr->mult_op == GM_MULT_K);
    const int32_t mtype = nreq > 0 ? reqs[0].M.type : GM_BOOL;
    const int32_t vtype = nreq > 0 ? reqs[0].A.type : sr->ztype;
    for (int64_t r = 0; r < nreq; r++)
        if (reqs[r].M.type != mtype || (valued && (reqs[r].A.type != vtype || reqs[r].B.type != vtype)))
            return set_error(GM_NO_VALUE, "gm_batch_dot3: requests mix types");
    const bool mvals = !mask_struct;
    if (mvals)
        for (int64_t r = 0; r < nreq; r++)
            if (!reqs[r].M.x) return set_error(GM_NO_VALUE, "gm_batch_dot3: valued mask without values");

    DevBuf tm, Pm, Jm, Xm, ta, Pa, Ja, Xa, tb, Pb, Jb, Xb;
    gm_csr M{}, A{}, B{};
    if (shared_all)
    {
        // every request is C<A> = A*A': stack A once and use it three times
        const bool vals = valued || mvals;
        GM_TRY(bt_stack(pa, rows, inner, vals, vtype, ta, Pa, Ja, Xa, &A));
        M = A;
        M.ncols = bcols;
        if (!mvals) M.x = nullptr;
        B = A;
        B.nrows = bcols;
    }
    else
    {
        GM_TRY(bt_stack(pa, rows, inner, valued, vtype, ta, Pa, Ja, Xa, &A));
        GM_TRY(bt_stack(pm, rows, bcols, mvals, mtype, tm, Pm, Jm, Xm, &M));
        GM_TRY(bt_stack(pb, bcols, inner, valued, vtype, tb, Pb, Jb, Xb, &B));
    }
    if (!valued) { A.x = nullptr; B.x = nullptr; }
    A.ncols = inner;
    B.ncols = inner;
    GM_TRY(gm_dot3(C, &M, mask_struct, &A, &B, sr, params));
    if (sr->mult_op == GM_MULT_K && inner > 0 && nreq > 1)
    {
        // positional values: k of the stacked problem = k of the request + its inner base
        hipStream_t st = stream();
        std::vector<int64_t> ib(nreq);
        int64_t acc = 0;
        for (int64_t r = 0; r < nreq; r++) { ib[r] = acc; acc += reqs[r].A.ncols; }
        DevBuf drb, dib;
        if (!drb.alloc(sizeof(int64_t) * nreq) || !dib.alloc(sizeof(int64_t) * nreq))
            return set_error(GM_OUT_OF_MEMORY, "gm_batch_dot3: positional fix");
        GM_HIP(hipMemcpyAsync(drb.ptr, row_base, sizeof(int64_t) * nreq, hipMemcpyHostToDevice, st), "fix");
        GM_HIP(hipMemcpyAsync(dib.ptr, ib.data(), sizeof(int64_t) * nreq, hipMemcpyHostToDevice, st), "fix");
        int64_t gr = (rows + 255) / 256;
        if (gr > (int64_t) cu_count() * 16) gr = (int64_t) cu_count() * 16;
        k_bt_fix_k<<<(unsigned) (gr > 0 ? gr : 1), 256, 0, st>>>(C->p, (int64_t *) C->x, drb.as<int64_t>(),
                                                               dib.as<int64_t>(), nreq, rows);
        GM_HIP(hipGetLastError(), "k_bt_fix_k");
        GM_HIP(hipStreamSynchronize(st), "fix sync");
    }
    return GM_SUCCESS;
}

extern "C" int gm_batch_saxpy3(gm_csr_out *C, int64_t *row_base, const gm_saxpy3_request *reqs,
                               int64_t nreq, bool mask_comp, bool mask_struct, const gm_csr *B,
                               const gm_semiring *sr, const gm_saxpy3_params *params)
{
    if (!C || !row_base || (!reqs && nreq > 0) || !B || !sr || nreq < 0)
        return set_error(GM_INVALID_VALUE, "gm_batch_saxpy3: NULL argument");
    GM_TRY(ensure_init());
    std::vector<BatchPart> pm, pa;
    int64_t rows = 0;
    const bool masked = nreq > 0 && reqs[0].M.p != nullptr;
    for (int64_t r = 0; r < nreq; r++)
    {
        const gm_saxpy3_request &q = reqs[r];
        const bool has_m = q.M.p != nullptr;
        if (has_m != masked) return set_error(GM_NO_VALUE, "gm_batch_saxpy3: masks on some requests only");
        if (!gm_saxpy3_branch(has_m ? &q.M : nullptr, mask_comp, mask_struct, &q.A, B, sr))
            return set_error(GM_NO_VALUE, "gm_batch_saxpy3: request %lld not handled", (long long) r);
        row_base[r] = rows;
        pa.push_back(BatchPart{q.A.p, q.A.j, q.A.x, q.A.nrows, rows, 0, 0, q.A.iso, 0});
        if (masked) pm.push_back(BatchPart{q.M.p, q.M.j, q.M.x, q.M.nrows, rows, 0, 0, q.M.iso, 0});
        rows += q.A.nrows;
    }
    row_base[nreq] = rows;
    const bool valued = !(sr->mult_op == GM_MULT_PAIR || sr->mult_op == GM_MULT_K);
    DevBuf ta, Pa, Ja, Xa, tm, Pm, Jm, Xm;
    gm_csr A{}, M{};
    GM_TRY(bt_stack(pa, rows, B->nrows, valued, valued ? reqs[0].A.type : GM_BOOL, ta, Pa, Ja, Xa, &A));
    if (masked)
    {
        const bool mvals = !mask_struct;
        GM_TRY(bt_stack(pm, rows, B->ncols, mvals, reqs[0].M.type, tm, Pm, Jm, Xm, &M));
    }
    return gm_saxpy3(C, masked ? &M : nullptr, mask_comp, mask_struct, &A, B, sr, params);
}

extern "C" int gm_batch_split(gm_csr_out *out, const gm_csr_out *C, const int64_t *row_base,
                              const int64_t *col_base, int64_t nreq)
{
    if (!out || !C || !row_base || nreq < 0) return set_error(GM_INVALID_VALUE, "gm_batch_split: NULL");
    GM_TRY(ensure_init());
    hipStream_t st = stream();
    const size_t xs = type_size(C->type);
    // per-request sizes need the stacked row pointers at the bases
    std::vector<int64_t> hb(nreq + 1);
    {
        DevBuf idx, val;
        if (!idx.alloc(sizeof(int64_t) * (nreq + 1)) || !val.alloc(sizeof(int64_t) * (nreq + 1)))
            return set_error(GM_OUT_OF_MEMORY, "gm_batch_split: workspace");
        GM_HIP(hipMemcpyAsync(idx.ptr, row_base, sizeof(int64_t) * (nreq + 1), hipMemcpyHostToDevice, st),
               "split bases");
        k_bt_gather<<<(unsigned) ((nreq + 256) / 256), 256, 0, st>>>(C->p, idx.as<int64_t>(),
                                                                  val.as<int64_t>(), nreq + 1);
        GM_HIP(hipGetLastError(), "k_bt_gather");
        GM_HIP(hipMemcpyAsync(hb.data(), val.ptr, sizeof(int64_t) * (nreq + 1), hipMemcpyDeviceToHost, st),
               "split bases");
        GM_HIP(hipStreamSynchronize(st), "split bases sync");
    }
    std::vector<SplitPart> sp(nreq);
    for (int64_t r = 0; r < nreq; r++)
    {
        const int64_t nr = row_base[r + 1] - row_base[r], nz = hb[r + 1] - hb[r];
        gm_csr_out &o = out[r];
        o.nrows = nr;
        o.ncols = col_base ? (col_base[r + 1] - col_base[r]) : C->ncols;
        o.nvals = nz;
        o.iso = C->iso;
        o.type = C->type;
        o.p = (int64_t *) gm_malloc(sizeof(int64_t) * (nr + 1));
        o.j = (int64_t *) gm_malloc(sizeof(int64_t) * (nz > 0 ? nz : 1));
        o.x = gm_malloc(C->iso ? xs : xs * (nz > 0 ? nz : 1));
        if (!o.p || !o.j || !o.x) return set_error(GM_OUT_OF_MEMORY, "gm_batch_split: output");
        sp[r] = SplitPart{o.p, o.j, o.x, row_base[r], nr, col_base ? col_base[r] : 0};
        if (C->iso)
            GM_HIP(hipMemcpyAsync(o.x, C->x, xs, hipMemcpyDeviceToDevice, st), "split iso value");
    }
    DevBuf dsp, deb;
    if (!dsp.alloc(sizeof(SplitPart) * (nreq > 0 ? nreq : 1)) || !deb.alloc(sizeof(int64_t) * (nreq + 1)))
        return set_error(GM_OUT_OF_MEMORY, "gm_batch_split: tables");
    GM_HIP(hipMemcpyAsync(dsp.ptr, sp.data(), sizeof(SplitPart) * nreq, hipMemcpyHostToDevice, st), "split");
    GM_HIP(hipMemcpyAsync(deb.ptr, hb.data(), sizeof(int64_t) * (nreq + 1), hipMemcpyHostToDevice, st),
           "split");
    const int64_t rows = row_base[nreq];
    const int64_t total = hb[nreq] - hb[0];
    const int64_t cap = (int64_t) cu_count() * 16;
    {
        ProfScope ps("batch_split");
        int64_t g1 = (rows + nreq + 255) / 256;
        if (g1 > cap) g1 = cap;
        if (nreq > 0)
            k_bt_split_rows<<<(unsigned) (g1 > 0 ? g1 : 1), 256, 0, st>>>(C->p, dsp.as<SplitPart>(),
                                                                          nreq, rows);
        int64_t g2 = (total + 255) / 256;
        if (g2 > cap) g2 = cap;
        if (total > 0)
            k_bt_split_ents<<<(unsigned) g2, 256, 0, st>>>(C->p, C->j, C->x, (int) xs, C->iso,
                                                           dsp.as<SplitPart>(), deb.as<int64_t>(),
                                                           nreq, total);
        GM_HIP(hipGetLastError(), "split kernels");
    }
    GM_HIP(hipStreamSynchronize(st), "split sync");
    return GM_SUCCESS;
}
