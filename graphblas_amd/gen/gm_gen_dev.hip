// gm_gen_dev.hip: many RMAT graphs generated directly in HBM.
//
// The same graphs as gm_gen_rmat_sym_n + gm_gen_sym_weights_i32 (gm_gen.c),
// bit for bit: the counter-based draws are a pure function of (seed,
// counter), so each sample e of graph g takes the same quadrant path on the
// GPU as on the host.  Pipeline, over chunks of graphs of at most 2^30
// directed edges:
//
//   1. one thread per sample writes both directions of its edge as a 64-bit
//      key (global row << 16 | column); a dropped sample (self-loop or an
//      endpoint >= n) writes the all-ones key, which sorts last;
//   2. radix sort of the keys (rocPRIM), then unique: rows come out sorted
//      and deduplicated;
//   3. row pointers by binary search of each row's first key; column
//      indices and symmetric int32 weights (a hash of the unordered pair)
//      from the keys.
//
// Columns fit 16 bits because n <= 65536.  The graphs are laid out back to
// back: graph g's CSR is P[pbase[g] .. pbase[g]+n_g] (starting at 0),
// J[nzbase[g] ..] and X[nzbase[g] ..].

#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <stdint.h>
#include <vector>
#include "gm_kernels.h"

namespace
{

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t rand64(uint64_t seed, uint64_t ctr)
{
    return mix64(seed + (ctr + 1) * 0x9E3779B97F4A7C15ULL);
}

__device__ __forceinline__ double unif(uint64_t seed, uint64_t ctr)
{
    return (double) (rand64(seed, ctr) >> 11) * (1.0 / 9007199254740992.0);
}

// last index t in [0, n) with a[t] <= v (a ascending, a[0] <= v)
__device__ __forceinline__ int64_t seg_of(const int64_t *a, int64_t n, int64_t v)
{
    int64_t lo = 0, hi = n - 1;
    while (lo < hi)
    {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (a[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t *a, int64_t n, uint64_t v)
{
    int64_t lo = 0, hi = n;
    while (lo < hi)
    {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

struct ChunkArgs
{
    int64_t ng;             // graphs in the chunk
    const int64_t *n;       // [ng]
    const int32_t *scale;   // [ng]
    const uint64_t *seed;   // [ng]
    const uint64_t *wseed;  // [ng]
    const int64_t *sp;      // [ng+1] sample prefix
    const int64_t *rb;      // [ng+1] global row base (chunk-local)
    const int64_t *pb;      // [ng+1] prefix of n+1 (row-pointer slots)
    double a, ab, abc;
};

__global__ void k_gen_keys(ChunkArgs c, uint64_t *keys, int64_t S)
{
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    for (int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; s < S; s += stride)
    {
        const int64_t g = seg_of(c.sp, c.ng, s);
        const uint64_t e = (uint64_t) (s - c.sp[g]);
        const uint64_t seed = c.seed[g];
        const int sc = c.scale[g];
        int64_t r = 0, q = 0;
        for (int lev = 0; lev < sc; lev++)
        {
            const double x = unif(seed, e * 64 + (uint64_t) lev);
            int bu, bv;
            if (x < c.a)        { bu = 0; bv = 0; }
            else if (x < c.ab)  { bu = 0; bv = 1; }
            else if (x < c.abc) { bu = 1; bv = 0; }
            else                { bu = 1; bv = 1; }
            r = 2 * r + bu;
            q = 2 * q + bv;
        }
        const int64_t n = c.n[g];
        uint64_t k0 = ~0ULL, k1 = ~0ULL;
        if (r != q && r < n && q < n)
        {
            const uint64_t base = (uint64_t) c.rb[g];
            k0 = ((base + (uint64_t) r) << 16) | (uint64_t) q;
            k1 = ((base + (uint64_t) q) << 16) | (uint64_t) r;
        }
        keys[2 * s] = k0;
        keys[2 * s + 1] = k1;
    }
}

// row pointers of every graph of the chunk (n_g + 1 slots each)
__global__ void k_gen_rows(ChunkArgs c, const uint64_t *ukeys, int64_t cnt, int64_t *P)
{
    const int64_t tot = c.pb[c.ng];
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += stride)
    {
        const int64_t g = seg_of(c.pb, c.ng, idx);
        const int64_t t = idx - c.pb[g];
        const uint64_t r0 = (uint64_t) c.rb[g];
        const int64_t first = lower_bound_u64(ukeys, cnt, r0 << 16);
        const int64_t at = lower_bound_u64(ukeys, cnt, (r0 + (uint64_t) t) << 16);
        P[idx] = at - first;
    }
}

__global__ void k_gen_nnz(ChunkArgs c, const uint64_t *ukeys, int64_t cnt, int64_t *nnz)
{
    const int64_t g = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= c.ng) return;
    const int64_t lo = lower_bound_u64(ukeys, cnt, (uint64_t) c.rb[g] << 16);
    const int64_t hi = lower_bound_u64(ukeys, cnt, (uint64_t) c.rb[g + 1] << 16);
    nnz[g] = hi - lo;
}

__global__ void k_gen_vals(ChunkArgs c, const uint64_t *ukeys, int64_t cnt, int64_t *J,
                           int32_t *X, int32_t lo, uint64_t range)
{
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    for (int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += stride)
    {
        const uint64_t key = ukeys[q];
        const int64_t row = (int64_t) (key >> 16);
        const int64_t col = (int64_t) (key & 0xFFFF);
        const int64_t g = seg_of(c.rb, c.ng, row);
        const int64_t i = row - c.rb[g];
        J[q] = col;
        if (X)
        {
            const uint64_t u = (uint64_t) (i < col ? i : col), v = (uint64_t) (i < col ? col : i);
            const uint64_t h = rand64(c.wseed[g], u * 0x100000001B3ULL ^ (v << 1));
            X[q] = lo + (int32_t) (h % range);
        }
    }
}

struct Dev
{
    void *p = nullptr;
    ~Dev() { if (p) gm_free(p); }
    bool alloc(size_t b) { p = gm_malloc(b ? b : 8); return p != nullptr; }
};

#define GEN_HIP(call)                                     \
    do {                                                  \
        if ((call) != hipSuccess) { rc = GM_PANIC; goto done; } \
    } while (0)

} // namespace

extern "C" int gm_gen_rmat_batch_dev(int64_t ng, const int64_t *n, const int64_t *ef,
                                     const uint64_t *seed, const uint64_t *wseed, double a,
                                     double b, double c, int32_t wlo, int32_t whi,
                                     int64_t *pbase, int64_t *nzbase, void **dP, void **dJ,
                                     void **dX)
{
    if (ng <= 0 || !n || !ef || !seed || !pbase || !nzbase || !dP || !dJ) return GM_INVALID_VALUE;
    for (int64_t g = 0; g < ng; g++)
        if (n[g] < 1 || n[g] > 65536 || ef[g] < 0) return GM_INVALID_VALUE;
    if (gm_init() != GM_SUCCESS) return GM_PANIC;
    const int64_t CHUNK = (int64_t) 1 << 30;      // directed edges per chunk
    int64_t cap = 0, rows = 0;
    pbase[0] = 0;
    for (int64_t g = 0; g < ng; g++)
    {
        cap += 2 * ef[g] * n[g];
        pbase[g + 1] = pbase[g] + n[g] + 1;
        rows += n[g];
    }
    (void) rows;
    int rc = GM_SUCCESS;
    hipStream_t st = nullptr;
    void *P = gm_malloc(sizeof(int64_t) * (size_t) pbase[ng]);
    void *J = gm_malloc(sizeof(int64_t) * (size_t) (cap > 0 ? cap : 1));
    void *X = dX ? gm_malloc(sizeof(int32_t) * (size_t) (cap > 0 ? cap : 1)) : nullptr;
    const double ab = a + b, abc = a + b + c;
    const uint64_t range = (uint64_t) (whi - wlo + 1);
    int64_t nz_done = 0;
    nzbase[0] = 0;
    if (!P || !J || (dX && !X)) { rc = GM_OUT_OF_MEMORY; goto done; }
    GEN_HIP(hipDeviceSynchronize());
    GEN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (int64_t g0 = 0; g0 < ng;)
    {
        // chunk [g0, g1)
        int64_t g1 = g0, edges = 0;
        while (g1 < ng && (g1 == g0 || edges + 2 * ef[g1] * n[g1] <= CHUNK))
        {
            edges += 2 * ef[g1] * n[g1];
            g1++;
        }
        const int64_t m = g1 - g0;
        std::vector<int64_t> hn(m), hsp(m + 1), hrb(m + 1), hpb(m + 1);
        std::vector<int32_t> hsc(m);
        std::vector<uint64_t> hseed(m), hwseed(m);
        hsp[0] = hrb[0] = hpb[0] = 0;
        for (int64_t t = 0; t < m; t++)
        {
            const int64_t g = g0 + t;
            hn[t] = n[g];
            int s = 0;
            while (((int64_t) 1 << s) < n[g]) s++;
            hsc[t] = s;
            hseed[t] = seed[g];
            hwseed[t] = wseed ? wseed[g] : 0;
            hsp[t + 1] = hsp[t] + ef[g] * n[g];
            hrb[t + 1] = hrb[t] + n[g];
            hpb[t + 1] = hpb[t] + n[g] + 1;
        }
        const int64_t S = hsp[m], R = hrb[m];
        int rowbits = 0;
        while (((int64_t) 1 << rowbits) < R + 1) rowbits++;
        const int end_bit = 16 + rowbits;
        Dev d_n, d_sc, d_seed, d_wseed, d_sp, d_rb, d_pb, d_k0, d_k1, d_cnt, d_nnz, d_tmp;
        if (!d_n.alloc(8 * m) || !d_sc.alloc(4 * m) || !d_seed.alloc(8 * m) ||
            !d_wseed.alloc(8 * m) || !d_sp.alloc(8 * (m + 1)) || !d_rb.alloc(8 * (m + 1)) ||
            !d_pb.alloc(8 * (m + 1)) || !d_k0.alloc(8 * (size_t) (2 * S)) ||
            !d_k1.alloc(8 * (size_t) (2 * S)) || !d_cnt.alloc(8) || !d_nnz.alloc(8 * m))
        {
            rc = GM_OUT_OF_MEMORY;
            goto done;
        }
        GEN_HIP(hipMemcpyAsync(d_n.p, hn.data(), 8 * m, hipMemcpyHostToDevice, st));
        GEN_HIP(hipMemcpyAsync(d_sc.p, hsc.data(), 4 * m, hipMemcpyHostToDevice, st));
        GEN_HIP(hipMemcpyAsync(d_seed.p, hseed.data(), 8 * m, hipMemcpyHostToDevice, st));
        GEN_HIP(hipMemcpyAsync(d_wseed.p, hwseed.data(), 8 * m, hipMemcpyHostToDevice, st));
        GEN_HIP(hipMemcpyAsync(d_sp.p, hsp.data(), 8 * (m + 1), hipMemcpyHostToDevice, st));
        GEN_HIP(hipMemcpyAsync(d_rb.p, hrb.data(), 8 * (m + 1), hipMemcpyHostToDevice, st));
        GEN_HIP(hipMemcpyAsync(d_pb.p, hpb.data(), 8 * (m + 1), hipMemcpyHostToDevice, st));
        {
            ChunkArgs ca{m, (const int64_t *) d_n.p, (const int32_t *) d_sc.p,
                         (const uint64_t *) d_seed.p, (const uint64_t *) d_wseed.p,
                         (const int64_t *) d_sp.p, (const int64_t *) d_rb.p,
                         (const int64_t *) d_pb.p, a, ab, abc};
            const int NT = 256, GRID = 8192;
            uint64_t *k0 = (uint64_t *) d_k0.p, *k1 = (uint64_t *) d_k1.p;
            const size_t nk = (size_t) (2 * S);
            size_t sort_bytes = 0, uniq_bytes = 0;
            int64_t cnt = 0;
            uint64_t last = 0;
            if (S > 0)
            {
                hipLaunchKernelGGL(k_gen_keys, dim3(GRID), dim3(NT), 0, st, ca, k0, S);
                GEN_HIP(hipGetLastError());
                GEN_HIP(rocprim::radix_sort_keys(nullptr, sort_bytes, k0, k1, nk, 0, end_bit, st));
                GEN_HIP(rocprim::unique(nullptr, uniq_bytes, k1, k0, (int64_t *) d_cnt.p, nk,
                                        rocprim::equal_to<uint64_t>(), st));
                if (!d_tmp.alloc(sort_bytes > uniq_bytes ? sort_bytes : uniq_bytes))
                {
                    rc = GM_OUT_OF_MEMORY;
                    goto done;
                }
                GEN_HIP(rocprim::radix_sort_keys(d_tmp.p, sort_bytes, k0, k1, nk, 0, end_bit, st));
                GEN_HIP(rocprim::unique(d_tmp.p, uniq_bytes, k1, k0, (int64_t *) d_cnt.p, nk,
                                        rocprim::equal_to<uint64_t>(), st));
                GEN_HIP(hipMemcpyAsync(&cnt, d_cnt.p, 8, hipMemcpyDeviceToHost, st));
                GEN_HIP(hipStreamSynchronize(st));
                if (cnt > 0)
                {
                    GEN_HIP(hipMemcpyAsync(&last, k0 + cnt - 1, 8, hipMemcpyDeviceToHost, st));
                    GEN_HIP(hipStreamSynchronize(st));
                    if (last == ~0ULL) cnt--;
                }
            }
            // with no samples the key array stays empty and every row is empty
            hipLaunchKernelGGL(k_gen_rows, dim3(GRID), dim3(NT), 0, st, ca, k0, cnt,
                               (int64_t *) P + pbase[g0]);
            GEN_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_gen_nnz, dim3((m + NT - 1) / NT), dim3(NT), 0, st, ca, k0, cnt,
                               (int64_t *) d_nnz.p);
            GEN_HIP(hipGetLastError());
            if (cnt > 0)
            {
                hipLaunchKernelGGL(k_gen_vals, dim3(GRID), dim3(NT), 0, st, ca, k0, cnt,
                                   (int64_t *) J + nz_done,
                                   X ? (int32_t *) X + nz_done : nullptr, wlo, range);
                GEN_HIP(hipGetLastError());
            }
            std::vector<int64_t> hnnz(m);
            GEN_HIP(hipMemcpyAsync(hnnz.data(), d_nnz.p, 8 * m, hipMemcpyDeviceToHost, st));
            GEN_HIP(hipStreamSynchronize(st));
            for (int64_t t = 0; t < m; t++) nzbase[g0 + t + 1] = nzbase[g0 + t] + hnnz[t];
            nz_done += cnt;
        }
        g0 = g1;
    }
done:
    if (st)
    {
        (void) hipStreamSynchronize(st);
        (void) hipStreamDestroy(st);
    }
    if (rc != GM_SUCCESS)
    {
        gm_free(P);
        gm_free(J);
        gm_free(X);
        return rc;
    }
    *dP = P;
    *dJ = J;
    if (dX) *dX = X;
    return GM_SUCCESS;
}
