// gm_saxpy3.hip: row-wise Gustavson/hash SpGEMM, C = A*B, C<M> = A*B,
// C<!M> = A*B (the reference's saxpy3 method).
//
// Reference: Source/mxm/GB_AxB_saxpy3.c:119-700 (driver: flop count, tasks,
// symbolic, numeric), GB_AxB_saxpy3_flopcount.c:81-478,
// GB_AxB_saxpy3_symbolic.c + factory/GB_AxB_saxpy3_symbolic_template.c
// (phase1 counts), template/GB_AxB_saxpy3_template.c (numeric), coarse hash
// phase5 (GB_AxB_saxpy3_coarseHash_phase5.c:22-93) for the accumulation order.
//
// Layout: all operands CSR in HBM.  C(i,:) = (+)_k A(i,k) (x) B(k,:).
//
//  flops   one thread per row: f = sum |B(k,:)|, table size t = f (or
//          min(f, |M(i,:)|) for C<M>, and <= ncols).  Rows are binned:
//            single  |A(i,:)| == 1: C(i,:) is B(k,:) filtered by the mask,
//                    already sorted; no hash table at all
//            wave S  t <= 64   : one wavefront, 128-slot LDS table
//            wave L  t <= 512  : one wavefront, 1024-slot LDS table
//            block   t <= 2048 : one 256-thread workgroup, 4096-slot table
//            global  larger    : one workgroup, table of 2^ceil(log2 2t) slots
//                                in HBM (sub-allocated by an atomic bump)
//  symbolic  per row, hash-insert the allowed column indices, count new keys
//            -> Cp[i]; device exclusive scan -> row pointers, nnz(C).
//  numeric   the same inserts plus values; flops are enumerated in the
//            reference's sequential order (k ascending over A(i,:), then j
//            ascending over B(k,:)) and applied in that order (lane order
//            inside a wavefront, wavefront order inside a workgroup), so
//            fp64 results equal the reference's coarse-task sums bit for bit
//            and ANY keeps the last writer, as GB_HX_UPDATE "z = t" does.
//            Rows are then sorted in LDS (rank count / bitonic) and written
//            to C: C is returned with sorted rows.
//  Masks     M(i,j) is tested by binary search in the sorted row M(i,:).
//
// Hash: multiplicative (Fibonacci) hashing of the 31-bit column index,
// linear probing, EMPTY = 0xffffffff.  Column indices are < 2^31-1.

#include "gm_internal.h"
#include "gm_block.h"

namespace gm
{

constexpr int SX_NT = 256;
constexpr uint32_t SX_EMPTY = 0xffffffffu;
enum { BIN_SINGLE = 0, BIN_WS = 1, BIN_WL = 2, BIN_BLK = 3, BIN_GLB = 4, SX_NBINS = 5 };
enum { MODE_NONE = 0, MODE_MASK = 1, MODE_NOTMASK = 2 };

struct SxArgs
{
    int64_t m, n;
    const int64_t *Ap, *Aj;
    const void *Ax;
    int32_t a_iso;
    const int64_t *Bp, *Bj;
    const void *Bx;
    int32_t b_iso;
    const int64_t *Mp, *Mj;
    const void *Mx;
    int32_t msize, m_iso, mask_struct, mode;
    int64_t *Cp;
    int64_t *Ci;
    void *Cx;
    uint32_t *rows;               // SX_NBINS x m row ids
    unsigned long long *cnt;      // SX_NBINS bin counters, then global slot total
    int64_t *gtab_off;            // per global-bin row: offset of its table
    int32_t *gtab_log2;           // per global-bin row: log2 of its table size
    uint32_t *gkeys;
    void *gvals;
    uint64_t *glist;
    int64_t ws_max, wl_max, blk_max;
    // B's entries as 16-byte (column, value bits) pairs, or null: one B row
    // of the gather then spans one contiguous range instead of two.  With
    // rec_lg >= 0 they are B's row records instead (row k at k << rec_lg,
    // its length in the top byte of the first record's column)
    const int64_t *Bjx;
    int32_t rec_lg;
    // row tables: 0 the reference's rule (a table that holds every column of
    // C is indexed by the column itself, Gustavson's dense workspace; a
    // global-bin row whose hash table would be at least ncols/12 slots, or
    // whose flops reach ncols/2, gets such a table), 1 always hash
    // (GxB_AxB_HASH), 2 dense wherever a table holds every column or a
    // global-bin row's dense table is at most 16 times its hash table
    // (GxB_AxB_GUSTAVSON)
    int32_t table;
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class Z>
__device__ __forceinline__ Z readlane(Z v, int src)
{
    if constexpr (sizeof(Z) == 8)
    {
        uint64_t t;
        __builtin_memcpy(&t, &v, 8);
        uint32_t lo = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) t, src);
        uint32_t hi = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) (t >> 32), src);
        uint64_t r = ((uint64_t) hi << 32) | lo;
        Z out;
        __builtin_memcpy(&out, &r, 8);
        return out;
    }
    else if constexpr (sizeof(Z) == 4)
    {
        int t;
        __builtin_memcpy(&t, &v, 4);
        t = __builtin_amdgcn_readlane(t, src);
        Z out;
        __builtin_memcpy(&out, &t, 4);
        return out;
    }
    else
    {
        return (Z) __builtin_amdgcn_readlane((int) v, src);
    }
}

__device__ __forceinline__ uint32_t sx_hash(uint32_t key, int logT)
{
    return (key * 0x9E3779B1u) >> (32 - logT);
}

// Is column j allowed by the mask for row i?  M(i,:) is [pm0, pm1).
__device__ __forceinline__ bool sx_mask_ok(const SxArgs &g, int64_t pm0, int64_t pm1, int64_t j)
{
    if (g.mode == MODE_NONE) return true;
    const int64_t q = lower_bound(g.Mj, pm0, pm1, j);
    bool in = (q < pm1) && (g.Mj[q] == j);
    if (in && !g.mask_struct) in = mask_true(g.Mx, q, g.msize, g.m_iso);
    return (g.mode == MODE_MASK) ? in : !in;
}

// insert key; returns its slot and whether this call created it.  dense: the
// table has a slot for every column, the key's own (no probing happens)
__device__ __forceinline__ int32_t sx_insert(uint32_t *keys, int logT, uint32_t key, bool &created,
                                             bool dense)
{
    const uint32_t mask = (1u << logT) - 1u;
    uint32_t h = dense ? key : sx_hash(key, logT);
    created = false;
    while (true)
    {
        uint32_t k = keys[h];
        if (k == key) return (int32_t) h;
        if (k == SX_EMPTY)
        {
            uint32_t prev = atomicCAS(&keys[h], SX_EMPTY, key);
            if (prev == SX_EMPTY) { created = true; return (int32_t) h; }
            if (prev == key) return (int32_t) h;
        }
        h = (h + 1) & mask;
    }
}

__device__ __forceinline__ int32_t ilog2_floor32(uint32_t x) { return 31 - __clz(x); }

//------------------------------------------------------------------------------
// team-of-TG helpers (TG lanes of one wavefront work on one row)
//------------------------------------------------------------------------------

constexpr int TG = 8;          // lanes per row in the flop-count and fused kernels
constexpr int FMAX = 64;       // largest flop count of a "small" row

template <class T>
__device__ __forceinline__ T team_incl_scan(T v)
{
    const int sub = threadIdx.x & (TG - 1);
    #pragma unroll
    for (int off = 1; off < TG; off <<= 1)
    {
        T u = __shfl_up(v, off, TG);
        if (sub >= off) v += u;
    }
    return v;
}

template <class T>
__device__ __forceinline__ T team_sum(T v)
{
    #pragma unroll
    for (int off = 1; off < TG; off <<= 1) v += __shfl_xor(v, off, TG);
    return v;
}

//------------------------------------------------------------------------------
// max row lengths of A and B (decides whether every row is "small"):
// out[0] / out[1]; and nnz(A) / nnz(B) in out[2] / out[5], so that one copy
// brings all four to the host
//------------------------------------------------------------------------------

__global__ __launch_bounds__(SX_NT) void k_sx_maxlen(const int64_t *__restrict__ Ap, int64_t m,
                                                     const int64_t *__restrict__ Bp, int64_t kb,
                                                     unsigned long long *out)
{
    __shared__ unsigned long long s[2][SX_NT / 64];
    unsigned long long ma = 0, mb = 0;
    const int64_t lim = m > kb ? m : kb;
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        out[2] = (unsigned long long) Ap[m];
        out[5] = (unsigned long long) Bp[kb];
    }
    // four consecutive rows per thread and pass, so that one pass of a
    // modest grid has all of its loads in flight together
    for (int64_t i0 = ((int64_t) blockIdx.x * SX_NT + threadIdx.x) * 4; i0 < lim;
         i0 += (int64_t) gridDim.x * SX_NT * 4)
    {
        #pragma unroll
        for (int u = 0; u < 4; u++)
        {
            const int64_t i = i0 + u;
            if (i < m) { unsigned long long l = (unsigned long long) (Ap[i + 1] - Ap[i]); ma = l > ma ? l : ma; }
            if (i < kb) { unsigned long long l = (unsigned long long) (Bp[i + 1] - Bp[i]); mb = l > mb ? l : mb; }
        }
    }
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        unsigned long long ua = __shfl_xor(ma, off, 64), ub = __shfl_xor(mb, off, 64);
        ma = ua > ma ? ua : ma;
        mb = ub > mb ? ub : mb;
    }
    if ((threadIdx.x & 63) == 0) { s[0][threadIdx.x >> 6] = ma; s[1][threadIdx.x >> 6] = mb; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < SX_NT / 64; w++)
        {
            ma = s[0][w] > ma ? s[0][w] : ma;
            mb = s[1][w] > mb ? s[1][w] : mb;
        }
        // an atomic only where this block raises the maximum seen so far
        // (a stale read only costs an atomic that was not needed)
        const unsigned long long ca = __hip_atomic_load(&out[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long cb = __hip_atomic_load(&out[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ma > ca) atomicMax(&out[0], ma);
        if (mb > cb) atomicMax(&out[1], mb);
    }
}

//------------------------------------------------------------------------------
// flops + binning: TG lanes per row (coalesced reads of A(i,:) and of the
// B row pointers it names).  Rows with f <= fs are "small"; their flops are
// summed into cnt[7] (capacity for the fused kernel) and they are listed in
// BIN_WS (used only when the fused kernel is not).  Larger rows go to the
// single / wave1024 / block4096 / global bins.
//------------------------------------------------------------------------------

__global__ __launch_bounds__(SX_NT) void k_sx_flops(SxArgs g, int64_t fs, int list_small)
{
    __shared__ unsigned int s_bc[SX_NBINS];
    __shared__ unsigned long long s_base[SX_NBINS];
    __shared__ unsigned long long s_red[SX_NT / 64];
    const int sub = threadIdx.x & (TG - 1);
    const int lane = threadIdx.x & 63;
    unsigned long long fsum = 0;          // flops of this thread's small rows
    for (int64_t base = (int64_t) blockIdx.x * (SX_NT / TG); base < g.m;
         base += (int64_t) gridDim.x * (SX_NT / TG))
    {
        const int64_t i = base + threadIdx.x / TG;
        int64_t f = 0, a0 = 0, a1 = 0;
        if (i < g.m)
        {
            a0 = g.Ap[i];
            a1 = g.Ap[i + 1];
            for (int64_t e = a0 + sub; e < a1; e += TG)
            {
                const int64_t k = g.Aj[e];
                f += g.Bp[k + 1] - g.Bp[k];
            }
        }
        f = team_sum<int64_t>(f);
        int bin = -1;
        int64_t t = f;
        if (threadIdx.x < SX_NBINS) s_bc[threadIdx.x] = 0;
        if (i < g.m && sub == 0)
        {
            if (g.mode == MODE_MASK)
            {
                const int64_t mi = g.Mp[i + 1] - g.Mp[i];
                if (mi < t) t = mi;
            }
            if (t > g.n) t = g.n;
            if (f == 0 || t == 0) g.Cp[i] = 0;
            else if (f <= fs) { bin = list_small ? BIN_WS : -1; fsum += (unsigned long long) f; }
            else if (a1 - a0 == 1) bin = BIN_SINGLE;
            else if (t <= g.wl_max) bin = BIN_WL;
            else if (t <= g.blk_max) bin = BIN_BLK;
            else bin = BIN_GLB;
        }
        __syncthreads();
        // block-aggregated appends: LDS counters, one global atomic per bin
        unsigned int local = 0;
        if (bin >= 0) local = atomicAdd(&s_bc[bin], 1u);
        __syncthreads();
        if (threadIdx.x < SX_NBINS && s_bc[threadIdx.x] > 0)
            s_base[threadIdx.x] = atomicAdd(&g.cnt[threadIdx.x], (unsigned long long) s_bc[threadIdx.x]);
        __syncthreads();
        if (bin >= 0)
        {
            const unsigned long long slot = s_base[bin] + local;
            g.rows[(int64_t) bin * g.m + (int64_t) slot] = (uint32_t) i;
            if (bin == BIN_GLB)
            {
                // table of 2^ceil(log2(2t)) slots, load factor <= 1/2, or a
                // dense one of 2^ceil(log2(ncols)) (see SxArgs::table)
                int lg = ilog2_ceil(2 * t);
                const int lgd = ilog2_ceil(g.n);
                const bool dn = (g.table == 0 && (((int64_t) 1 << lg) >= g.n / 12 || t >= g.n / 2)) ||
                                (g.table == 2 && lgd <= lg + 4);
                if (dn && lgd > lg) lg = lgd;
                const unsigned long long sz = 1ull << lg;
                g.gtab_off[slot] = (int64_t) atomicAdd(&g.cnt[SX_NBINS], sz);
                g.gtab_log2[slot] = lg;
            }
        }
        __syncthreads();
    }
    // flops of small rows: one atomic per workgroup
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) fsum += __shfl_xor(fsum, off, 64);
    if (lane == 0) s_red[threadIdx.x >> 6] = fsum;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long tsum = 0;
        for (int w = 0; w < SX_NT / 64; w++) tsum += s_red[w];
        if (tsum) atomicAdd(&g.cnt[7], tsum);
    }
}

// sum of the symbolic counts of the rows in the big bins -> cnt[6]
__global__ __launch_bounds__(SX_NT) void k_sx_sum_big(SxArgs g)
{
    __shared__ unsigned long long s_red[SX_NT / 64];
    unsigned long long sum = 0;
    const int bins[4] = {BIN_SINGLE, BIN_WL, BIN_BLK, BIN_GLB};
    for (int b = 0; b < 4; b++)
    {
        const int64_t count = (int64_t) g.cnt[bins[b]];
        const uint32_t *list = g.rows + (int64_t) bins[b] * g.m;
        for (int64_t q = (int64_t) blockIdx.x * SX_NT + threadIdx.x; q < count;
             q += (int64_t) gridDim.x * SX_NT)
            sum += (unsigned long long) g.Cp[list[q]];
    }
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long tsum = 0;
        for (int w = 0; w < SX_NT / 64; w++) tsum += s_red[w];
        if (tsum) atomicAdd(&g.cnt[6], tsum);
    }
}

//------------------------------------------------------------------------------
// single-entry rows: C(i,:) = A(i,k) (x) B(k,:) filtered by the mask
//------------------------------------------------------------------------------

template <class S, int PHASE>
__global__ __launch_bounds__(SX_NT) void k_sx_single(SxArgs g)
{
    using Z = typename S::Z;
    const int lane = threadIdx.x & 63;
    const int64_t count = (int64_t) g.cnt[BIN_SINGLE];
    const uint32_t *list = g.rows + (int64_t) BIN_SINGLE * g.m;
    for (int64_t q = (int64_t) blockIdx.x * 4 + (threadIdx.x >> 6); q < count;
         q += (int64_t) gridDim.x * 4)
    {
        const int64_t i = list[q];
        const int64_t e = g.Ap[i];
        const int64_t k = g.Aj[e];
        const int64_t b0 = g.Bp[k], b1 = g.Bp[k + 1];
        const int64_t pm0 = g.Mp ? g.Mp[i] : 0, pm1 = g.Mp ? g.Mp[i + 1] : 0;
        if (PHASE == 0 && g.mode == MODE_NONE)
        {
            if (lane == 0) g.Cp[i] = b1 - b0;
            continue;
        }
        const Z av = S::geta((const Z *) g.Ax, e, g.a_iso);
        const int64_t cbase = PHASE ? g.Cp[i] : 0;
        const int64_t cend = PHASE ? g.Cp[i + 1] : 0;
        int64_t run = 0;
        for (int64_t c0 = b0; c0 < b1; c0 += 64)
        {
            const int64_t pb = c0 + lane;
            bool ok = false;
            int64_t j = 0;
            if (pb < b1)
            {
                j = g.Bj[pb];
                ok = sx_mask_ok(g, pm0, pm1, j);
            }
            const unsigned long long bal = __ballot(ok);
            if (PHASE == 1 && ok)
            {
                const int64_t pos = cbase + run + __popcll(bal & ((1ull << lane) - 1ull));
                if (pos < cend)
                {
                    g.Ci[pos] = j;
                    if constexpr (!S::iso_out)
                        ((Z *) g.Cx)[pos] = S::mul(av, S::geta((const Z *) g.Bx, pb, g.b_iso), k);
                }
            }
            run += __popcll(bal);
        }
        if (PHASE == 0 && lane == 0) g.Cp[i] = run;
    }
}

//------------------------------------------------------------------------------
// one row on a team of NT lanes (NT = 64: one wavefront; NT = 256: one
// workgroup).  keys/vals/list are the row's table (LDS or HBM), logT its size.
//------------------------------------------------------------------------------

template <int NT>
__device__ __forceinline__ void team_sync()
{
    if constexpr (NT == 64) wave_sync(); else __syncthreads();
}

// bitonic sort of list[0..P) (P a power of two), ascending
template <int NT>
__device__ void team_bitonic(uint64_t *list, int64_t P, int tid)
{
    for (int64_t k = 2; k <= P; k <<= 1)
    {
        for (int64_t jj = k >> 1; jj > 0; jj >>= 1)
        {
            for (int64_t t = tid; t < P / 2; t += NT)
            {
                const int64_t lo = 2 * t - (t & (jj - 1));
                const int64_t hi = lo + jj;
                const bool asc = (lo & k) == 0;
                const uint64_t a = list[lo], b = list[hi];
                if ((a > b) == asc) { list[lo] = b; list[hi] = a; }
            }
            team_sync<NT>();
        }
    }
}

template <class S, int NT, int PHASE>
__device__ void sx_row(const SxArgs &g, int64_t i, uint32_t *keys, typename S::Z *vals,
                       uint64_t *list, int logT, int64_t *s_pre, int64_t *s_k, int64_t *s_bs,
                       typename S::Z *s_av, int *s_flag, int64_t *s_red)
{
    using Z = typename S::Z;
    constexpr int NWV = NT / 64;
    const int tid = (NT == 64) ? (threadIdx.x & 63) : threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int wv = (NT == 64) ? 0 : (threadIdx.x >> 6);
    const int64_t T = (int64_t) 1 << logT;
    const bool dense = g.table != 1 && T >= g.n;    // Gustavson: index by column

    for (int64_t t = tid; t < T; t += NT)
    {
        keys[t] = SX_EMPTY;
        if (PHASE) vals[t] = S::identity();
    }
    team_sync<NT>();

    const int64_t a0 = g.Ap[i], a1 = g.Ap[i + 1];
    const int64_t pm0 = g.Mp ? g.Mp[i] : 0, pm1 = g.Mp ? g.Mp[i + 1] : 0;
    int64_t nnew = 0;

    for (int64_t ea = a0; ea < a1; ea += NT)
    {
        // stage up to NT entries of A(i,:) with their B-row extents
        const int64_t e = ea + tid;
        int64_t len = 0;
        if (e < a1)
        {
            const int64_t k = g.Aj[e];
            const int64_t bs = g.Bp[k];
            len = g.Bp[k + 1] - bs;
            s_k[tid] = k;
            s_bs[tid] = bs;
            if (PHASE) s_av[tid] = S::geta((const Z *) g.Ax, e, g.a_iso);
        }
        int64_t total;
        int64_t pre;
        if constexpr (NT == 64)
        {
            const int64_t inc = wave_inclusive_scan<int64_t>(len);
            pre = inc - len;
            total = __shfl(inc, 63, 64);
        }
        else
        {
            pre = block_exclusive_scan<int64_t, NT>(len, s_red, &total);
        }
        s_pre[tid] = pre;
        const int64_t nA = (a1 - ea) < NT ? (a1 - ea) : NT;
        team_sync<NT>();

        for (int64_t f0 = 0; f0 < total; f0 += NT)
        {
            const int64_t f = f0 + tid;
            int32_t slot = -1;
            bool created = false;
            Z tv = S::identity();
            if (f < total)
            {
                // owning A entry: last index with s_pre <= f
                int lo = 0, hi = (int) nA - 1;
                while (lo < hi)
                {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= f) lo = mid; else hi = mid - 1;
                }
                const int64_t pb = s_bs[lo] + (f - s_pre[lo]);
                const int64_t j = g.Bj[pb];
                if (sx_mask_ok(g, pm0, pm1, j))
                {
                    if (PHASE)
                        tv = S::mul(s_av[lo], S::geta((const Z *) g.Bx, pb, g.b_iso), s_k[lo]);
                    slot = sx_insert(keys, logT, (uint32_t) j, created, dense);
                }
            }
            if (PHASE == 0)
            {
                nnew += created ? 1 : 0;
            }
            else
            {
                // fast path: every active lane of the team created a fresh key,
                // so all slots are distinct and new
                const bool dup = (slot >= 0) && !created;
                bool any_dup;
                if constexpr (NT == 64) any_dup = __ballot(dup) != 0;
                else any_dup = __syncthreads_or(dup) != 0;
                if (!any_dup)
                {
                    if (slot >= 0) vals[slot] = S::upd(vals[slot], tv);
                }
                else
                {
                    // apply in flop order: wavefront by wavefront, lane by lane
                    for (int w = 0; w < NWV; w++)
                    {
                        if (w == wv)
                        {
                            bool owner = slot >= 0;
                            Z z = owner ? vals[slot] : S::identity();
                            for (int src = 0; src < 64; src++)
                            {
                                const int32_t ss = __builtin_amdgcn_readlane(slot, src);
                                const Z ts = readlane<Z>(tv, src);
                                if (slot >= 0 && ss == slot)
                                {
                                    if (src < lane) owner = false;
                                    else z = S::upd(z, ts);
                                }
                            }
                            if (owner) vals[slot] = z;
                        }
                        team_sync<NT>();
                    }
                }
                team_sync<NT>();
            }
        }
        team_sync<NT>();
    }

    if (PHASE == 0)
    {
        int64_t tot;
        if constexpr (NT == 64)
        {
            tot = wave_inclusive_scan<int64_t>(nnew);
            tot = __shfl(tot, 63, 64);
        }
        else
        {
            tot = block_sum<int64_t, NT>(nnew, s_red);
        }
        if (tid == 0) g.Cp[i] = tot;
        return;
    }

    // numeric: gather the row's keys, sort, write C(i,:)
    const int64_t cbase = g.Cp[i], c = g.Cp[i + 1] - cbase;
    int64_t run = 0;
    for (int64_t t0 = 0; t0 < T; t0 += NT)
    {
        const int64_t t = t0 + tid;
        const uint32_t k = (t < T) ? keys[t] : SX_EMPTY;
        const bool v = (k != SX_EMPTY);
        int64_t pos, cnt;
        if constexpr (NT == 64)
        {
            const unsigned long long bal = __ballot(v);
            pos = __popcll(bal & ((1ull << lane) - 1ull));
            cnt = __popcll(bal);
        }
        else
        {
            pos = block_exclusive_scan<int64_t, NT>(v ? 1 : 0, s_red, &cnt);
        }
        if (v && run + pos < c) list[run + pos] = ((uint64_t) k << 32) | (uint64_t) t;
        run += cnt;
    }
    team_sync<NT>();
    if (c <= 64 && NT == 64)
    {
        // rank by counting: keys in a row are distinct
        if (lane < c)
        {
            const uint64_t my = list[lane];
            int r = 0;
            for (int t = 0; t < c; t++) r += (list[t] < my) ? 1 : 0;
            g.Ci[cbase + r] = (int64_t) (my >> 32);
            if constexpr (!S::iso_out) ((Z *) g.Cx)[cbase + r] = vals[(uint32_t) my & (uint32_t) (T - 1)];
        }
        team_sync<NT>();
        return;
    }
    int64_t P = 1;
    while (P < c) P <<= 1;
    for (int64_t t = c + tid; t < P; t += NT) list[t] = ~0ull;
    team_sync<NT>();
    team_bitonic<NT>(list, P, tid);
    for (int64_t t = tid; t < c; t += NT)
    {
        const uint64_t w = list[t];
        g.Ci[cbase + t] = (int64_t) (w >> 32);
        if constexpr (!S::iso_out) ((Z *) g.Cx)[cbase + t] = vals[(uint32_t) w & (uint32_t) (T - 1)];
    }
    team_sync<NT>();
}

// one wavefront per row, LDS table of 2^LOGT slots; 4 rows per workgroup
template <class S, int LOGT, int PHASE>
__global__ __launch_bounds__(SX_NT) void k_sx_wave(SxArgs g, int bin)
{
    using Z = typename S::Z;
    constexpr int T = 1 << LOGT;
    constexpr int NWV = SX_NT / 64;
    __shared__ uint32_t s_keys[NWV][T];
    __shared__ Z s_vals[NWV][PHASE ? T : 1];
    __shared__ uint64_t s_list[NWV][PHASE ? T / 2 : 1];
    __shared__ int64_t s_pre[NWV][64], s_k[NWV][64], s_bs[NWV][64];
    __shared__ Z s_av[NWV][64];
    const int wv = threadIdx.x >> 6;
    const int64_t count = (int64_t) g.cnt[bin];
    const uint32_t *list = g.rows + (int64_t) bin * g.m;
    for (int64_t q = (int64_t) blockIdx.x * NWV + wv; q < count; q += (int64_t) gridDim.x * NWV)
    {
        sx_row<S, 64, PHASE>(g, (int64_t) list[q], s_keys[wv], s_vals[wv], s_list[wv], LOGT,
                             s_pre[wv], s_k[wv], s_bs[wv], s_av[wv], nullptr, nullptr);
    }
}

// one workgroup per row, LDS table of 2^LOGT slots
template <class S, int LOGT, int PHASE>
__global__ __launch_bounds__(SX_NT) void k_sx_block(SxArgs g)
{
    using Z = typename S::Z;
    constexpr int T = 1 << LOGT;
    __shared__ uint32_t s_keys[T];
    __shared__ Z s_vals[PHASE ? T : 1];
    __shared__ uint64_t s_list[PHASE ? T / 2 : 1];
    __shared__ int64_t s_pre[SX_NT], s_k[SX_NT], s_bs[SX_NT];
    __shared__ Z s_av[SX_NT];
    __shared__ int64_t s_red[SX_NT / 64 + 1];
    const int64_t count = (int64_t) g.cnt[BIN_BLK];
    const uint32_t *list = g.rows + (int64_t) BIN_BLK * g.m;
    for (int64_t q = blockIdx.x; q < count; q += gridDim.x)
    {
        sx_row<S, SX_NT, PHASE>(g, (int64_t) list[q], s_keys, s_vals, s_list, LOGT,
                                s_pre, s_k, s_bs, s_av, nullptr, s_red);
    }
}

// one workgroup per row, table in HBM
template <class S, int PHASE>
__global__ __launch_bounds__(SX_NT) void k_sx_global(SxArgs g)
{
    using Z = typename S::Z;
    __shared__ int64_t s_pre[SX_NT], s_k[SX_NT], s_bs[SX_NT];
    __shared__ Z s_av[SX_NT];
    __shared__ int64_t s_red[SX_NT / 64 + 1];
    const int64_t count = (int64_t) g.cnt[BIN_GLB];
    const uint32_t *list = g.rows + (int64_t) BIN_GLB * g.m;
    for (int64_t q = blockIdx.x; q < count; q += gridDim.x)
    {
        const int64_t off = g.gtab_off[q];
        const int lg = g.gtab_log2[q];
        sx_row<S, SX_NT, PHASE>(g, (int64_t) list[q], g.gkeys + off, ((Z *) g.gvals) + off,
                                g.glist + off, lg, s_pre, s_k, s_bs, s_av, nullptr, s_red);
    }
}

__global__ void k_sx_zero(unsigned long long *x, int n)
{
    if ((int) threadIdx.x < n) x[threadIdx.x] = 0;
}

__global__ void k_sx_iso_one(void *x, int size)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        uint8_t *b = (uint8_t *) x;
        for (int t = 0; t < size; t++) b[t] = 0;
        b[0] = 1;
    }
}

//------------------------------------------------------------------------------
// single-pass fused kernel for small rows (f <= fs <= 64)
//------------------------------------------------------------------------------
//
// A persistent grid takes tiles of FUSED_NT/TG = 64 consecutive rows in ticket
// order.  One team of TG = 8 lanes per row:
//   expand   the row's flops, in the reference's order (A(i,:) in chunks of
//            TG entries, B(k,:) in order), into LDS as sort keys
//            (j << 6 | flop index) plus values; masked-out flops stay at the
//            all-ones key
//   sort     bitonic sort of the keys (the flop index makes them unique and
//            keeps equal columns in flop order)
//   compress runs of one column are reduced in flop order, so fp64 sums are
//            the reference's sequential ascending-k sums and ANY keeps the
//            last writer
// Rows with f > fs ("big") were counted beforehand by the hash kernels; the
// tile only adds their counts to the scan.  The tile's output offset comes
// from a decoupled look-back over the previous tiles' published counts
// (status word: 2-bit flag | 62-bit count); then C(tile rows) is written with
// contiguous, coalesced stores, together with Cp.  Ticket order means a tile
// only waits on tiles already running; a bounded spin turns a lost
// predecessor into an error flag instead of a hang.

// B packed as (column, value bits) pairs for the fused kernel's gathers: for
// 8-byte values that the semiring reads
template <class S>
constexpr bool SX_PACK_OK = sizeof(typename S::Z) == 8 && !S::pattern;

__global__ __launch_bounds__(256) void k_sx_pack(const int64_t *__restrict__ Bj,
                                                 const int64_t *__restrict__ Bx, int64_t nnz,
                                                 longlong2 *__restrict__ out)
{
    for (int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t) gridDim.x * blockDim.x)
        out[e] = make_longlong2(Bj[e], Bx[e]);
}

// Row records of B (gm_csr_aux): slot s of row k at (k << lg) + s holds
// (column, value bits) of the row's s-th entry, or (0, 0) past its end; slot
// 0's column carries the row length in its top byte (columns < 2^56)
constexpr long long SX_REC_COL = (1ll << 56) - 1;
constexpr int SX_REC_MAX_LG = 4;          // rows of at most 16 entries

__global__ __launch_bounds__(256) void k_sx_rec_build(const int64_t *__restrict__ Bp,
                                                      const int64_t *__restrict__ Bj,
                                                      const int64_t *__restrict__ Bx, int64_t nrows,
                                                      int lg, longlong2 *__restrict__ out)
{
    const int64_t total = nrows << lg;
    const int64_t smask = ((int64_t) 1 << lg) - 1;
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t) gridDim.x * blockDim.x)
    {
        const int64_t k = t >> lg, sl = t & smask;
        const int64_t p0 = Bp[k], len = Bp[k + 1] - p0;
        long long c = 0, v = 0;
        if (sl < len) { c = Bj[p0 + sl]; v = Bx[p0 + sl]; }
        if (sl == 0) c |= (long long) ((unsigned long long) len << 56);
        out[t] = make_longlong2(c, v);
    }
}

// threads of a fused-kernel workgroup: 64-row tiles halve the look-backs of
// 32-row ones (ER 2^20: 0.861 against 0.899 ms; 16-row tiles 1.32 ms)
constexpr int FUSED_NT = 512;
constexpr int SX_FUSED_WAVES = 4;
constexpr int FUSED_ROWS = FUSED_NT / TG;  // rows of a tile
constexpr unsigned long long LB_AGG = 1ull << 62;
constexpr unsigned long long LB_PRE = 2ull << 62;
constexpr unsigned long long LB_MASK = (1ull << 62) - 1;
constexpr int LB_SPIN_LIMIT = 1 << 22;
constexpr int LB_WIN = 1;            // look-back window: LB_WIN x 64 predecessors (one wavefront's loads per step measured fastest)

// K: the sort key type, (j << 6 | flop).  32-bit keys (columns < 2^26) halve
// the key LDS and the sort's shuffles and registers.
// The bound's second argument is the wavefronts per SIMD asked for; with
// 32-bit keys SX_FUSED_WAVES, so that a workgroup waiting in its look-back
// is covered by the others on the CU; 64-bit keys double the key LDS and ask
// for three.
// PK: 0 B read from Bp/Bj/Bx, 1 from the packed pairs, 2 from the row records
template <class S, class K, int NT, int PK>
__global__ __launch_bounds__(NT, sizeof(K) == 4 ? SX_FUSED_WAVES : 3)
void k_sx_fused(SxArgs g, int64_t fs,
                                                         unsigned long long *status, int64_t ntiles,
                                                         int *err)
{
    using Z = typename S::Z;
    constexpr int FUSED_ROWS = NT / TG;
    static_assert(FUSED_ROWS <= 64, "the tile's row offsets are scanned by one wavefront");
    // per-team rows padded by one element so the 8 teams of a wavefront
    // start on different LDS banks
    __shared__ K s_key[FUSED_ROWS][FMAX + 1];         // sort keys, then output columns (int32)
    __shared__ Z s_val[FUSED_ROWS][FMAX + 1];         // flop values, then output values
    __shared__ int64_t s_cnt[FUSED_ROWS], s_off[FUSED_ROWS + 1];
    __shared__ int s_big[FUSED_ROWS];
    __shared__ int64_t s_tile, s_prefix;

    const int team = threadIdx.x / TG, sub = threadIdx.x & (TG - 1);
    K *key = s_key[team];
    Z *val = s_val[team];
    int32_t *outj = (int32_t *) s_key[team];
    unsigned long long *ticket = status + ntiles;

    while (true)
    {
        if (threadIdx.x == 0) s_tile = (int64_t) atomicAdd(ticket, 1ull);
        __syncthreads();
        const int64_t tile = s_tile;
        if (tile >= ntiles) break;
        const int64_t i = tile * FUSED_ROWS + team;

        #pragma unroll
        for (int u = 0; u < FMAX / TG; u++) key[sub * (FMAX / TG) + u] = (K) ~(K) 0;
        int64_t f = 0, cnt = 0;
        int big = 0, nallow = 0;
        if (i < g.m)
        {
            const int64_t a0 = g.Ap[i], a1 = g.Ap[i + 1];
            const int64_t pm0 = g.Mp ? g.Mp[i] : 0, pm1 = g.Mp ? g.Mp[i + 1] : 0;
            wave_sync();
            for (int64_t e0 = a0; e0 < a1; e0 += TG)
            {
                const int64_t e = e0 + sub;
                int64_t k = 0, bs = 0, len = 0;
                Z av = S::identity();
                if (e < a1)
                {
                    k = g.Aj[e];
                    if constexpr (PK == 2)
                    {
                        // the row's run of records; its length rides in the
                        // first record, whose line the team reads next anyway
                        bs = k << g.rec_lg;
                        len = (int64_t) ((unsigned long long) g.Bjx[2 * bs] >> 56);
                    }
                    else
                    {
                        bs = g.Bp[k];
                        len = g.Bp[k + 1] - bs;
                    }
                    av = S::geta((const Z *) g.Ax, e, g.a_iso);
                }
                const int64_t inc = team_incl_scan<int64_t>(len);
                const int64_t tot = __shfl(inc, TG - 1, TG);
                if (f + tot > fs) { big = 1; break; }
                // the chunk's flop offsets in every lane's registers; flop q
                // belongs to the last A entry whose offset is <= q (entries
                // past the row's end have offset tot, never <= q)
                const int32_t pre = (int32_t) (inc - len);
                int32_t pres[TG];
                #pragma unroll
                for (int m = 0; m < TG; m++) pres[m] = __shfl(pre, m, TG);
                // the chunk's flops in three passes, so that the B gathers of
                // all FMAX/TG flops of a lane are in flight together: owners
                // and B positions (a lane past the chunk's flops takes its
                // last flop's, a valid position), the gathers, then the keys
                if (tot > 0)
                {
                    constexpr int NIT = FMAX / TG;
                    int los[NIT];
                    int64_t pbs[NIT];
                    #pragma unroll
                    for (int it = 0; it < NIT; it++)
                    {
                        const int64_t q = (int64_t) it * TG + sub;
                        const int32_t qc = (int32_t) (q < tot ? q : tot - 1);
                        int lo = -1;
                        #pragma unroll
                        for (int m = 0; m < TG; m++) lo += (pres[m] <= qc) ? 1 : 0;
                        // shuffled by every lane of the team (a lane outside a
                        // branch would return zeros)
                        los[it] = lo;
                        pbs[it] = __shfl(bs, lo, TG) + (qc - __shfl(pre, lo, TG));
                    }
                    int64_t js[NIT];
                    uint64_t xb[NIT];
                    #pragma unroll
                    for (int it = 0; it < NIT; it++)
                    {
                        if constexpr (PK != 0)
                        {
                            const longlong2 jx = *(const longlong2 *) (g.Bjx + 2 * pbs[it]);
                            js[it] = (PK == 2) ? (jx.x & SX_REC_COL) : jx.x;
                            xb[it] = (uint64_t) jx.y;
                        }
                        else
                        {
                            js[it] = g.Bj[pbs[it]];
                            xb[it] = 0;
                        }
                    }
                    #pragma unroll
                    for (int it = 0; it < NIT; it++)
                    {
                        const int64_t q = (int64_t) it * TG + sub;
                        const Z avl = shfl_any<Z>(av, los[it], TG);
                        const int64_t kl = (S::mult == GM_MULT_K) ? __shfl(k, los[it], TG) : 0;
                        if (q < tot && sx_mask_ok(g, pm0, pm1, js[it]))
                        {
                            const int64_t fq = f + q;
                            key[fq] = (K) (((K) js[it] << 6) | (K) fq);
                            if constexpr (!S::iso_out)
                            {
                                Z bv;
                                if constexpr (PK != 0) bv = __builtin_bit_cast(Z, xb[it]);
                                else bv = S::geta((const Z *) g.Bx, pbs[it], g.b_iso);
                                val[fq] = S::mul(avl, bv, kl);
                            }
                            nallow++;
                        }
                    }
                }
                f += tot;
                wave_sync();
            }
            if (big) cnt = g.Cp[i];
        }
        nallow = team_sum<int>(nallow);
        if (!big && nallow > 0)
        {
            // bitonic sort of the 64 keys: lane sub holds positions
            // sub*8 .. sub*8+7 in registers; distances < 8 are register
            // compare-exchanges, distances >= 8 are width-8 shuffles
            constexpr int PER_L = FMAX / TG;
            K kr[PER_L];
            #pragma unroll
            for (int r = 0; r < PER_L; r++) kr[r] = key[sub * PER_L + r];
            #pragma unroll
            for (int kk = 2; kk <= FMAX; kk <<= 1)
            {
                #pragma unroll
                for (int jj = kk >> 1; jj > 0; jj >>= 1)
                {
                    if (jj >= PER_L)
                    {
                        const int lmask = jj / PER_L;
                        const bool lower = (sub & lmask) == 0;
                        #pragma unroll
                        for (int r = 0; r < PER_L; r++)
                        {
                            const int idx = sub * PER_L + r;
                            const bool asc = (idx & kk) == 0;
                            K o;
                            if constexpr (sizeof(K) == 4)
                                o = (K) __shfl_xor((unsigned int) kr[r], lmask, TG);
                            else
                                o = (K) __shfl_xor((unsigned long long) kr[r], lmask, TG);
                            const K mn = kr[r] < o ? kr[r] : o, mx = kr[r] < o ? o : kr[r];
                            kr[r] = (lower == asc) ? mn : mx;
                        }
                    }
                    else
                    {
                        #pragma unroll
                        for (int r = 0; r < PER_L; r++)
                        {
                            const int partner = r ^ jj;
                            if (partner > r)
                            {
                                const int idx = sub * PER_L + r;
                                const bool asc = (idx & kk) == 0;
                                const K a = kr[r], b = kr[partner];
                                if ((a > b) == asc) { kr[r] = b; kr[partner] = a; }
                            }
                        }
                    }
                }
            }
            #pragma unroll
            for (int r = 0; r < PER_L; r++) key[sub * PER_L + r] = kr[r];
            wave_sync();
            // compress: lane owns sorted positions sub*8 .. sub*8+7, whose
            // keys are still in kr[]; a run's head sums its values in flop
            // order from registers, and from LDS only past the lane's end
            constexpr int PER = FMAX / TG;
            const K kprev = (K) shfl_any<K>(kr[PER - 1], sub > 0 ? sub - 1 : 0, TG);
            Z vv[PER];
            #pragma unroll
            for (int u = 0; u < PER; u++)
            {
                vv[u] = S::identity();
                if constexpr (!S::iso_out)
                    if (sub * PER + u < nallow) vv[u] = val[kr[u] & 63];
            }
            int32_t hj[PER];
            Z hz[PER];
            int nh = 0;
            #pragma unroll
            for (int u = 0; u < PER; u++)
            {
                const int p = sub * PER + u;
                hj[u] = -1;
                if (p < nallow)
                {
                    const int64_t jp = (int64_t) (kr[u] >> 6);
                    const K kb = (u == 0) ? kprev : kr[u > 0 ? u - 1 : 0];
                    if (p == 0 || (int64_t) (kb >> 6) != jp)
                    {
                        Z z = vv[u];
                        bool open = true;
                        #pragma unroll
                        for (int u2 = u + 1; u2 < PER; u2++)
                        {
                            if (open && sub * PER + u2 < nallow && (int64_t) (kr[u2] >> 6) == jp)
                            {
                                if constexpr (!S::iso_out) z = S::upd(z, vv[u2]);
                            }
                            else open = false;
                        }
                        if (open)
                        {
                            for (int p2 = sub * PER + PER; p2 < nallow; p2++)
                            {
                                const K k2 = key[p2];
                                if ((int64_t) (k2 >> 6) != jp) break;
                                if constexpr (!S::iso_out) z = S::upd(z, val[k2 & 63]);
                            }
                        }
                        hj[u] = (int32_t) jp;
                        hz[u] = z;
                        nh++;
                    }
                }
            }
            const int inc = team_incl_scan<int>(nh);
            int r = inc - nh;
            cnt = __shfl(inc, TG - 1, TG);
            wave_sync();
            #pragma unroll
            for (int u = 0; u < PER; u++)
            {
                if (hj[u] >= 0)
                {
                    outj[r] = hj[u];
                    if constexpr (!S::iso_out) val[r] = hz[u];
                    r++;
                }
            }
        }
        if (sub == 0)
        {
            s_cnt[team] = cnt;
            s_big[team] = big;
        }
        __syncthreads();
        if (threadIdx.x < 64)
        {
            // wavefront 0: tile offsets, then a look-back over up to 64
            // predecessors per step (lane l reads tile - 1 - l)
            const int lane = threadIdx.x;
            // row offsets of the tile: inclusive scan of the row counts
            // across the wavefront's lanes
            const int64_t c = lane < FUSED_ROWS ? s_cnt[lane] : 0;
            int64_t incl = c;
            #pragma unroll
            for (int off = 1; off < FUSED_ROWS; off <<= 1)
            {
                const int64_t o = __shfl_up(incl, off, 64);
                if (lane >= off) incl += o;
            }
            const int64_t run = __shfl(incl, FUSED_ROWS - 1, 64);
            if (lane == 0)
                __hip_atomic_store(&status[tile], (tile == 0 ? LB_PRE : LB_AGG) | (unsigned long long) run,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (lane < FUSED_ROWS) s_off[lane] = incl - c;
            if (lane == 0) s_off[FUSED_ROWS] = run;
            int64_t prefix = 0;
            if (tile > 0)
            {
                // window of LB_WIN*64 predecessors, position u*64 + lane =
                // distance - 1; all loads are independent (relaxed: the
                // counts travel inside the status words themselves)
                int64_t end = tile - 1;
                int spins = 0;
                while (true)
                {
                    unsigned long long w[LB_WIN];
                    #pragma unroll
                    for (int u = 0; u < LB_WIN; u++)
                    {
                        const int64_t idx = end - (int64_t) u * 64 - lane;
                        w[u] = (idx >= 0) ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT)
                                          : LB_PRE;
                    }
                    int firstp = LB_WIN * 64;
                    bool zero_before = false;
                    #pragma unroll
                    for (int u = 0; u < LB_WIN; u++)
                    {
                        const unsigned long long flag = w[u] & ~LB_MASK;
                        const unsigned long long bp = __ballot(flag == LB_PRE);
                        const unsigned long long b0 = __ballot(flag == 0);
                        if (firstp == LB_WIN * 64)
                        {
                            const int fp = bp ? (__ffsll((long long) bp) - 1) : 64;
                            const unsigned long long upto =
                                (fp >= 63) ? ~0ull : ((2ull << fp) - 1ull);
                            if (b0 & upto) zero_before = true;
                            if (fp < 64) firstp = u * 64 + fp;
                        }
                    }
                    if (zero_before)
                    {
                        // a predecessor has not published yet
                        if (++spins > LB_SPIN_LIMIT)
                        {
                            if (lane == 0) atomicExch(err, 1);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    int64_t v = 0;
                    #pragma unroll
                    for (int u = 0; u < LB_WIN; u++)
                        if (u * 64 + lane <= firstp) v += (int64_t) (w[u] & LB_MASK);
                    #pragma unroll
                    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                    prefix += v;
                    if (firstp < LB_WIN * 64) break;
                    end -= LB_WIN * 64;
                }
                if (lane == 0)
                    __hip_atomic_store(&status[tile], LB_PRE | (unsigned long long) (prefix + run),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (lane == 0) s_prefix = prefix;
        }
        __syncthreads();
        const int64_t prefix = s_prefix;
        const int64_t total = s_off[FUSED_ROWS];
        if (sub == 0 && i < g.m) g.Cp[i] = prefix + s_off[team];
        if (tile == ntiles - 1 && threadIdx.x == 0) g.Cp[g.m] = prefix + total;
        // each team copies its own small row: the row's entries are
        // contiguous in C, its lanes store consecutive ones
        if (!big && i < g.m)
        {
            const int64_t r0 = prefix + s_off[team];
            for (int pos = sub; pos < (int) cnt; pos += TG)
            {
                g.Ci[r0 + pos] = (int64_t) outj[pos];
                if constexpr (!S::iso_out) ((Z *) g.Cx)[r0 + pos] = val[pos];
            }
        }
        __syncthreads();
    }
}

__global__ void k_sx_fill_u64(unsigned long long *x, int64_t n)
{
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t) gridDim.x * blockDim.x)
        x[t] = 0;
}

static inline unsigned sx_grid(int64_t items, int64_t per_block, int64_t cap)
{
    int64_t g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned) g;
}

// pattern-only instance used by the symbolic phase of every semiring
using SxSym = SR<GM_ADD_ANY, GM_MULT_PAIR, uint8_t>;

template <class S, int PHASE>
static int sx_launch_phase(SxArgs &g, const int64_t *cnt, bool skip_ws)
{
    hipStream_t st = stream();
    const int64_t cap = (int64_t) cu_count() * 32;
    const char *names[2][5] = {
        {"sx_sym_single", "sx_sym_wave128", "sx_sym_wave1024", "sx_sym_block4096", "sx_sym_global"},
        {"sx_num_single", "sx_num_wave128", "sx_num_wave1024", "sx_num_block4096", "sx_num_global"}};
    if (cnt[BIN_SINGLE] > 0)
    {
        ProfScope ps(names[PHASE][0]);
        k_sx_single<S, PHASE><<<sx_grid(cnt[BIN_SINGLE], 4, cap), SX_NT, 0, st>>>(g);
        GM_HIP(hipGetLastError(), "k_sx_single");
    }
    if (cnt[BIN_WS] > 0 && !skip_ws)
    {
        ProfScope ps(names[PHASE][1]);
        k_sx_wave<S, 7, PHASE><<<sx_grid(cnt[BIN_WS], 4, cap), SX_NT, 0, st>>>(g, BIN_WS);
        GM_HIP(hipGetLastError(), "k_sx_wave128");
    }
    if (cnt[BIN_WL] > 0)
    {
        ProfScope ps(names[PHASE][2]);
        k_sx_wave<S, 10, PHASE><<<sx_grid(cnt[BIN_WL], 4, cap), SX_NT, 0, st>>>(g, BIN_WL);
        GM_HIP(hipGetLastError(), "k_sx_wave1024");
    }
    if (cnt[BIN_BLK] > 0)
    {
        ProfScope ps(names[PHASE][3]);
        k_sx_block<S, 12, PHASE><<<sx_grid(cnt[BIN_BLK], 1, cap), SX_NT, 0, st>>>(g);
        GM_HIP(hipGetLastError(), "k_sx_block4096");
    }
    if (cnt[BIN_GLB] > 0)
    {
        ProfScope ps(names[PHASE][4]);
        k_sx_global<S, PHASE><<<sx_grid(cnt[BIN_GLB], 1, cap), SX_NT, 0, st>>>(g);
        GM_HIP(hipGetLastError(), "k_sx_global");
    }
    return GM_SUCCESS;
}

template <class S>
static int saxpy3_run(gm_csr_out *C, const gm_csr *M, bool mask_comp, bool mask_struct,
                      const gm_csr *A, const gm_csr *B, const gm_saxpy3_params *prm)
{
    using Z = typename S::Z;
    hipStream_t st = stream();
    const int64_t m = A->nrows, n = B->ncols;
    const size_t zs = sizeof(Z);
    const size_t entry_bytes = sizeof(int64_t) + (S::iso_out ? 0 : zs);

    DevBuf Cp, rows, cnt, goff, glog;
    if (!Cp.alloc(sizeof(int64_t) * (m + 1)) || !rows.alloc(sizeof(uint32_t) * SX_NBINS * (m > 0 ? m : 1)) ||
        !cnt.alloc(sizeof(unsigned long long) * 16) || !goff.alloc(sizeof(int64_t) * (m > 0 ? m : 1)) ||
        !glog.alloc(sizeof(int32_t) * (m > 0 ? m : 1)))
        return set_error(GM_OUT_OF_MEMORY, "saxpy3: workspace");

    SxArgs g;
    g.m = m; g.n = n;
    g.Ap = A->p; g.Aj = A->j; g.Ax = A->x; g.a_iso = A->iso;
    g.Bp = B->p; g.Bj = B->j; g.Bx = B->x; g.b_iso = B->iso;
    g.Mp = M ? M->p : nullptr; g.Mj = M ? M->j : nullptr;
    g.Mx = (M && !mask_struct) ? M->x : nullptr;
    g.msize = M ? (int32_t) type_size(M->type) : 0;
    g.m_iso = M ? M->iso : 0;
    g.mask_struct = (mask_struct || !M || !M->x) ? 1 : 0;
    g.mode = !M ? MODE_NONE : (mask_comp ? MODE_NOTMASK : MODE_MASK);
    g.Cp = Cp.as<int64_t>(); g.Ci = nullptr; g.Cx = nullptr;
    g.rows = rows.as<uint32_t>(); g.cnt = cnt.as<unsigned long long>();
    g.gtab_off = goff.as<int64_t>(); g.gtab_log2 = glog.as<int32_t>();
    g.gkeys = nullptr; g.gvals = nullptr; g.glist = nullptr;
    g.Bjx = nullptr;
    g.rec_lg = -1;
    g.table = prm ? (int32_t) prm->table : 0;
    if (g.table < 0 || g.table > 2) return set_error(GM_INVALID_VALUE, "saxpy3: table %d", (int) g.table);
    const gm_saxpy3_params tu = saxpy3_tuned();
    g.ws_max = pick_param(prm ? prm->wave_small_max : 0, tu.wave_small_max, 64);
    g.wl_max = pick_param(prm ? prm->wave_large_max : 0, tu.wave_large_max, 512);
    g.blk_max = pick_param(prm ? prm->block_max : 0, tu.block_max, 2048);
    if (g.ws_max > FMAX) g.ws_max = FMAX;
    if (g.wl_max > 512) g.wl_max = 512;
    if (g.blk_max > 2048) g.blk_max = 2048;
    const int64_t fs = g.ws_max;        // flops bound of a "small" row
    const int mode = (prm && prm->mode) ? (int) prm->mode : (int) tu.mode;
    // the fused kernel holds C's small rows at their flop bound until the
    // scan; it is used while that bound stays within a third of HBM
    const double budget = (double) device_hbm_bytes() / 3.0;

    int64_t hc[16] = {0};
    unsigned long long *hp = (unsigned long long *) pinned_scratch(16 * sizeof(unsigned long long));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    if (m == 0)
    {
        k_sx_fill_u64<<<1, 64, 0, st>>>((unsigned long long *) g.Cp, 1);
        GM_HIP(hipGetLastError(), "k_sx_fill_u64");
        C->nrows = m; C->ncols = n; C->nvals = 0;
        if (!Cp.ptr) return GM_OUT_OF_MEMORY;
        DevBuf Ci, Cx;
        if (!Ci.alloc(8) || !Cx.alloc(zs)) return set_error(GM_OUT_OF_MEMORY, "saxpy3: output");
        GM_HIP(hipStreamSynchronize(st), "saxpy3 sync");
        C->p = (int64_t *) Cp.release(); C->j = (int64_t *) Ci.release(); C->x = Cx.release();
        C->iso = S::iso_out ? 1 : 0;
        return GM_SUCCESS;
    }

    // 1. max row lengths of A and B, and nnz(A), for the all-small shortcut.
    // When A has far fewer rows than B (a few frontier rows against a large
    // graph) the flop pass below is as cheap as this one, so it is skipped
    // together with its synchronisation.
    int64_t maxA = 0, maxB = 0, nnzA = 0, nnzB = -1;
    const bool probe_small = m * 8 >= B->nrows;
    k_sx_fill_u64<<<1, 64, 0, st>>>(g.cnt, 16);
    GM_HIP(hipGetLastError(), "k_sx_fill_u64");
    if (probe_small)
    {
        {
            ProfScope ps("sx_maxlen");
            const int64_t lim = m > B->nrows ? m : B->nrows;
            k_sx_maxlen<<<sx_grid((lim + 3) / 4, SX_NT, (int64_t) cu_count() * 8), SX_NT, 0, st>>>(
                A->p, m, B->p, B->nrows, g.cnt + 8);
            GM_HIP(hipGetLastError(), "k_sx_maxlen");
        }
        // counters 8..13: max |A(i,:)|, max |B(k,:)|, nnz(A), -, -, nnz(B)
        GM_HIP(hipMemcpyAsync(hp + 8, g.cnt + 8, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
               "maxlen d2h");
        GM_HIP(hipStreamSynchronize(st), "maxlen sync");
        maxA = (int64_t) hp[8];
        maxB = (int64_t) hp[9];
        nnzA = (int64_t) hp[10];
        nnzB = (int64_t) hp[13];
    }

    bool fused = (mode != 2);
    bool all_small = fused && probe_small && (double) maxA * (double) maxB <= (double) fs &&
                     (double) nnzA * (double) maxB * (double) entry_bytes <= budget;
    int64_t cap_small = all_small ? nnzA * maxB : 0;

    // 2. flop count and binning (skipped when every row is small)
    if (!all_small)
    {
        {
            ProfScope ps("sx_flops");
            k_sx_flops<<<sx_grid(m, SX_NT / TG, (int64_t) cu_count() * 16), SX_NT, 0, st>>>(
                g, fs, fused ? 0 : 1);
            GM_HIP(hipGetLastError(), "k_sx_flops");
        }
        GM_HIP(hipMemcpyAsync(hp, g.cnt, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
               "bin counts");
        GM_HIP(hipStreamSynchronize(st), "bin counts sync");
        for (int t = 0; t < 8; t++) hc[t] = (int64_t) hp[t];
        cap_small = hc[7];
        if (fused && (double) cap_small * (double) entry_bytes > budget)
        {
            // too large to hold at the flop bound: redo the binning with the
            // small rows listed for the symbolic + numeric path
            fused = false;
            k_sx_fill_u64<<<1, 64, 0, st>>>(g.cnt, 16);
            k_sx_flops<<<sx_grid(m, SX_NT / TG, (int64_t) cu_count() * 16), SX_NT, 0, st>>>(g, fs, 1);
            GM_HIP(hipGetLastError(), "k_sx_flops");
            GM_HIP(hipMemcpyAsync(hp, g.cnt, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
                   "bin counts");
            GM_HIP(hipStreamSynchronize(st), "bin counts sync");
            for (int t = 0; t < 8; t++) hc[t] = (int64_t) hp[t];
        }
    }

    DevBuf gkeys, gvals, glist;
    const int64_t gslots = hc[SX_NBINS];
    if (gslots > 0)
    {
        if (!gkeys.alloc(sizeof(uint32_t) * gslots) || !gvals.alloc(zs * gslots) ||
            !glist.alloc(sizeof(uint64_t) * gslots))
            return set_error(GM_OUT_OF_MEMORY, "saxpy3: global hash tables (%lld slots)",
                             (long long) gslots);
        g.gkeys = gkeys.as<uint32_t>(); g.gvals = gvals.ptr; g.glist = glist.as<uint64_t>();
    }
    const bool have_big = hc[BIN_SINGLE] + hc[BIN_WL] + hc[BIN_BLK] + hc[BIN_GLB] > 0;
    int64_t nvals = 0;
    DevBuf Ci, Cx;
    if (fused)
    {
        // 3. symbolic phase of the big rows only
        int64_t big_total = 0;
        if (have_big)
        {
            GM_TRY((sx_launch_phase<SxSym, 0>(g, hc, true)));
            const int64_t nbig = hc[BIN_SINGLE] + hc[BIN_WL] + hc[BIN_BLK] + hc[BIN_GLB];
            k_sx_sum_big<<<sx_grid(nbig, SX_NT, 512), SX_NT, 0, st>>>(g);
            GM_HIP(hipGetLastError(), "k_sx_sum_big");
            GM_HIP(hipMemcpyAsync(hp + 6, g.cnt + 6, sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, st), "big total");
            GM_HIP(hipStreamSynchronize(st), "big total sync");
            big_total = (int64_t) hp[6];
        }
        const int64_t cap = cap_small + big_total;
        if (!Ci.alloc(sizeof(int64_t) * (cap > 0 ? cap : 1)) ||
            !Cx.alloc(S::iso_out ? zs : zs * (cap > 0 ? cap : 1)))
            return set_error(GM_OUT_OF_MEMORY, "saxpy3: output (%lld entries)", (long long) cap);
        g.Ci = Ci.as<int64_t>();
        g.Cx = Cx.ptr;
        // 4. fused small-row kernel with the look-back scan over all rows
        const int64_t ntiles = (m + FUSED_ROWS - 1) / FUSED_ROWS;
        DevBuf status, errf;
        if (!status.alloc(sizeof(unsigned long long) * (ntiles + 1)) || !errf.alloc(sizeof(int)))
            return set_error(GM_OUT_OF_MEMORY, "saxpy3: scan status");
        k_sx_fill_u64<<<sx_grid(ntiles + 1, 256, 1024), 256, 0, st>>>(
            status.as<unsigned long long>(), ntiles + 1);
        k_sx_fill_u64<<<1, 64, 0, st>>>((unsigned long long *) g.cnt + 12, 1);
        GM_HIP(hipGetLastError(), "k_sx_fill_u64");
        // B's rows: from its row records when the operand carries them (one
        // aligned run per gathered row, no row pointers); else as 16-byte
        // (column, value) pairs when its rows are gathered at least about once
        // each on average (the pack reads and writes B once)
        DevBuf bjx;
        int pk = 0;
        if constexpr (SX_PACK_OK<S>)
        {
            const gm_csr_aux *ax = B->aux;
            if (!B->iso && B->x && ax && ax->rec && ax->rec_lg >= 0 && ax->nrows == B->nrows)
            {
                g.Bjx = (const int64_t *) ax->rec;
                g.rec_lg = ax->rec_lg;
                pk = 2;
            }
            else if (!B->iso && B->x && nnzB > 0 && 2 * nnzA >= B->nrows &&
                     bjx.alloc(sizeof(longlong2) * nnzB))
            {
                ProfScope ps("sx_pack");
                k_sx_pack<<<sx_grid(nnzB, 256, (int64_t) cu_count() * 16), 256, 0, st>>>(
                    B->j, (const int64_t *) B->x, nnzB, bjx.as<longlong2>());
                GM_HIP(hipGetLastError(), "k_sx_pack");
                g.Bjx = bjx.as<int64_t>();
                pk = 1;
            }
        }
        {
            ProfScope ps("sx_fused");
            const int64_t gcap = (int64_t) cu_count() * 8 * 256 / FUSED_NT;
            const int64_t grid = ntiles < gcap ? ntiles : gcap;
            unsigned long long *stp = status.as<unsigned long long>();
            int *errp = (int *) (g.cnt + 12);
            const bool k32 = B->ncols <= ((int64_t) 1 << 26);
#define GM_SX_FUSED(K, PKV) \
    k_sx_fused<S, K, FUSED_NT, PKV><<<(unsigned) grid, FUSED_NT, 0, st>>>(g, fs, stp, ntiles, errp)
            if constexpr (SX_PACK_OK<S>)
            {
                if (pk == 2 && k32) GM_SX_FUSED(uint32_t, 2);
                else if (pk == 2) GM_SX_FUSED(uint64_t, 2);
                else if (pk == 1 && k32) GM_SX_FUSED(uint32_t, 1);
                else if (pk == 1) GM_SX_FUSED(uint64_t, 1);
            }
            if (pk == 0 && k32) GM_SX_FUSED(uint32_t, 0);
            else if (pk == 0) GM_SX_FUSED(uint64_t, 0);
#undef GM_SX_FUSED
            GM_HIP(hipGetLastError(), "k_sx_fused");
        }
        GM_HIP(hipMemcpyAsync(hp + 11, g.Cp + m, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz(C)");
        GM_HIP(hipMemcpyAsync(hp + 12, g.cnt + 12, sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
               "fused err");
        GM_HIP(hipStreamSynchronize(st), "fused sync");
        if (hp[12] != 0) return set_error(GM_PANIC, "saxpy3: look-back scan did not complete");
        nvals = (int64_t) hp[11];
        // 5. numeric phase of the big rows
        if (have_big && nvals > 0) GM_TRY((sx_launch_phase<S, 1>(g, hc, true)));
    }
    else
    {
        // symbolic + scan + numeric for every row
        GM_TRY((sx_launch_phase<SxSym, 0>(g, hc, false)));
        GM_TRY(scan_exclusive_i64(g.Cp, g.Cp, m));
        GM_HIP(hipMemcpyAsync(hp + 11, g.Cp + m, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz(C)");
        GM_HIP(hipStreamSynchronize(st), "nnz(C) sync");
        nvals = (int64_t) hp[11];
        if (!Ci.alloc(sizeof(int64_t) * (nvals > 0 ? nvals : 1)) ||
            !Cx.alloc(S::iso_out ? zs : zs * (nvals > 0 ? nvals : 1)))
            return set_error(GM_OUT_OF_MEMORY, "saxpy3: output (%lld entries)", (long long) nvals);
        g.Ci = Ci.as<int64_t>();
        g.Cx = Cx.ptr;
        if (nvals > 0) GM_TRY((sx_launch_phase<S, 1>(g, hc, false)));
    }
    if (S::iso_out)
    {
        k_sx_iso_one<<<1, 1, 0, st>>>(Cx.ptr, (int) zs);
        GM_HIP(hipGetLastError(), "k_sx_iso_one");
    }
    GM_HIP(hipStreamSynchronize(st), "saxpy3 sync");
    C->nrows = m;
    C->ncols = n;
    C->nvals = nvals;
    C->p = (int64_t *) Cp.release();
    C->j = (int64_t *) Ci.release();
    C->x = Cx.release();
    C->iso = S::iso_out ? 1 : 0;
    return GM_SUCCESS;
}

} // namespace gm

using namespace gm;

//------------------------------------------------------------------------------
// total multiplies of C=A*B: the sum over A's entries (i,k) of |B(k,:)|.
// Partial sums per workgroup, then one workgroup adds them (no atomics).
//------------------------------------------------------------------------------

constexpr int FLOPS_NT = 256;
constexpr int FLOPS_MAXG = 2048;

__global__ __launch_bounds__(FLOPS_NT) void k_flopcount(const int64_t *__restrict__ Aj, int64_t nnz,
                                                        const int64_t *__restrict__ Bp,
                                                        unsigned long long *part)
{
    __shared__ unsigned long long s[FLOPS_NT / 64];
    unsigned long long f = 0;
    for (int64_t e = (int64_t) blockIdx.x * FLOPS_NT + threadIdx.x; e < nnz;
         e += (int64_t) gridDim.x * FLOPS_NT)
    {
        const int64_t k = Aj[e];
        f += (unsigned long long) (Bp[k + 1] - Bp[k]);
    }
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = f;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < FLOPS_NT / 64; w++) f += s[w];
        part[blockIdx.x] = f;
    }
}

__global__ __launch_bounds__(FLOPS_NT) void k_flopcount_sum(unsigned long long *part, int n)
{
    __shared__ unsigned long long s[FLOPS_NT / 64];
    unsigned long long f = 0;
    for (int t = threadIdx.x; t < n; t += FLOPS_NT) f += part[t];
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = f;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < FLOPS_NT / 64; w++) f += s[w];
        part[FLOPS_MAXG] = f;
    }
}

extern "C" int gm_flopcount(int64_t *flops, const gm_csr *A, const gm_csr *B)
{
    if (!flops || !A || !B) return set_error(GM_INVALID_VALUE, "gm_flopcount: NULL argument");
    if (A->ncols != B->nrows)
        return set_error(GM_INVALID_VALUE, "gm_flopcount: A is %lld x %lld, B is %lld x %lld",
                         (long long) A->nrows, (long long) A->ncols, (long long) B->nrows,
                         (long long) B->ncols);
    *flops = 0;
    if (A->nrows == 0 || B->nrows == 0) return GM_SUCCESS;
    if (!A->p || !B->p) return set_error(GM_INVALID_VALUE, "gm_flopcount: NULL row pointers");
    GM_TRY(ensure_init());
    hipStream_t st = stream();
    unsigned long long *hp = (unsigned long long *) pinned_scratch(16 * sizeof(unsigned long long));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(hp, A->p + A->nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz(A)");
    GM_HIP(hipStreamSynchronize(st), "nnz(A) sync");
    const int64_t nnz = (int64_t) hp[0];
    if (nnz <= 0) return GM_SUCCESS;
    if (!A->j) return set_error(GM_INVALID_VALUE, "gm_flopcount: NULL column indices");
    DevBuf part;
    if (!part.alloc(sizeof(unsigned long long) * (FLOPS_MAXG + 1)))
        return set_error(GM_OUT_OF_MEMORY, "gm_flopcount: workspace");
    const unsigned grid = sx_grid(nnz, FLOPS_NT * 4, FLOPS_MAXG);
    k_flopcount<<<grid, FLOPS_NT, 0, st>>>(A->j, nnz, B->p, part.as<unsigned long long>());
    GM_HIP(hipGetLastError(), "k_flopcount");
    k_flopcount_sum<<<1, FLOPS_NT, 0, st>>>(part.as<unsigned long long>(), (int) grid);
    GM_HIP(hipGetLastError(), "k_flopcount_sum");
    GM_HIP(hipMemcpyAsync(hp, part.as<unsigned long long>() + FLOPS_MAXG, sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, st), "flops d2h");
    GM_HIP(hipStreamSynchronize(st), "flops sync");
    *flops = (int64_t) hp[0];
    return GM_SUCCESS;
}

extern "C" bool gm_saxpy3_branch(const gm_csr *M, bool mask_comp, bool mask_struct,
                                 const gm_csr *A, const gm_csr *B, const gm_semiring *sr)
{
    (void) mask_comp;
    (void) mask_struct;
    if (!A || !B || !sr) return false;
    if (!semiring_supported(sr)) return false;
    if (!operand_ok(sr, A) || !operand_ok(sr, B)) return false;
    if (A->ncols != B->nrows) return false;
    if (M && (M->nrows != A->nrows || M->ncols != B->ncols)) return false;
    // column indices are hashed as 31-bit keys; row ids are 32-bit
    if (B->ncols >= (int64_t) 0x7fffffff || A->nrows >= ((int64_t) 1 << 32)) return false;
    return true;
}

extern "C" int gm_saxpy3(gm_csr_out *C, const gm_csr *M, bool mask_comp, bool mask_struct,
                         const gm_csr *A, const gm_csr *B, const gm_semiring *sr,
                         const gm_saxpy3_params *params)
{
    if (!C || !A || !B || !sr) return set_error(GM_INVALID_VALUE, "gm_saxpy3: NULL argument");
    if (!gm_saxpy3_branch(M, mask_comp, mask_struct, A, B, sr))
        return set_error(GM_NO_VALUE, "gm_saxpy3: case not handled by the GPU kernels");
    GM_TRY(ensure_init());
#define GM_SX_CASE(ADD, MULT, ZC, Z)                                                      \
    if (sr->add_op == ADD && sr->mult_op == MULT && sr->ztype == ZC)                      \
    {                                                                                     \
        int rc = saxpy3_run<SR<ADD, MULT, Z>>(C, M, mask_comp, mask_struct, A, B, params); \
        if (rc == GM_SUCCESS) C->type = ZC;                                               \
        return rc;                                                                        \
    }
    GM_SEMIRING_LIST(GM_SX_CASE)
#undef GM_SX_CASE
    return set_error(GM_NO_VALUE, "gm_saxpy3: semiring not instantiated");
}

//------------------------------------------------------------------------------
// derived forms of an operand (gm_csr_aux)
//------------------------------------------------------------------------------

extern "C" int gm_csr_aux_build(gm_csr_aux **aux, const gm_csr *X)
{
    if (!aux || !X) return set_error(GM_INVALID_VALUE, "gm_csr_aux_build: NULL argument");
    *aux = nullptr;
    // row records need 8-byte values, one per entry, and columns below 2^56
    if (!X->p || X->nrows <= 0 || !X->x || X->iso || type_size(X->type) != 8 ||
        X->ncols >= ((int64_t) 1 << 56))
        return GM_SUCCESS;
    GM_TRY(ensure_init());
    hipStream_t st = stream();
    DevBuf cnt;
    if (!cnt.alloc(8 * sizeof(unsigned long long))) return set_error(GM_OUT_OF_MEMORY, "aux: workspace");
    unsigned long long *hp = (unsigned long long *) pinned_scratch(8 * sizeof(unsigned long long));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    k_sx_fill_u64<<<1, 64, 0, st>>>(cnt.as<unsigned long long>(), 8);
    k_sx_maxlen<<<sx_grid(X->nrows, SX_NT * 4, (int64_t) cu_count() * 4), SX_NT, 0, st>>>(
        X->p, X->nrows, X->p, 0, cnt.as<unsigned long long>());
    GM_HIP(hipGetLastError(), "aux: k_sx_maxlen");
    GM_HIP(hipMemcpyAsync(hp, cnt.ptr, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
           "aux: lengths");
    GM_HIP(hipStreamSynchronize(st), "aux: sync");
    const int64_t maxlen = (int64_t) hp[0], nnz = (int64_t) hp[2];
    int lg = 0;
    while (((int64_t) 1 << lg) < maxlen) lg++;
    if (nnz <= 0 || lg > SX_REC_MAX_LG || (X->nrows << lg) > nnz + nnz / 2 + 1024) return GM_SUCCESS;
    gm_csr_aux *a = (gm_csr_aux *) calloc(1, sizeof(gm_csr_aux));
    if (!a) return set_error(GM_OUT_OF_MEMORY, "aux: header");
    a->nrows = X->nrows;
    a->rec_lg = lg;
    a->rec = gm_malloc(sizeof(longlong2) * (size_t) (X->nrows << lg));
    if (!a->rec)
    {
        free(a);
        return set_error(GM_OUT_OF_MEMORY, "aux: row records");
    }
    k_sx_rec_build<<<sx_grid(X->nrows << lg, 256, (int64_t) cu_count() * 16), 256, 0, st>>>(
        X->p, X->j, (const int64_t *) X->x, X->nrows, lg, (longlong2 *) a->rec);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess)
    {
        gm_free(a->rec);
        free(a);
        return check_hip(e, "aux: k_sx_rec_build");
    }
    *aux = a;
    return GM_SUCCESS;
}

extern "C" void gm_csr_aux_free(gm_csr_aux *aux)
{
    if (!aux) return;
    gm_free(aux->rec);
    free(aux);
}
