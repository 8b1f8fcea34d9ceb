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

// insert key; returns its slot and whether this call created it
__device__ __forceinline__ int32_t sx_insert(uint32_t *keys, int logT, uint32_t key, bool &created)
{
    const uint32_t mask = (1u << logT) - 1u;
    uint32_t h = sx_hash(key, logT);
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
// flops + binning
//------------------------------------------------------------------------------

__global__ __launch_bounds__(SX_NT) void k_sx_flops(SxArgs g)
{
    for (int64_t base = (int64_t) blockIdx.x * SX_NT; base < g.m; base += (int64_t) gridDim.x * SX_NT)
    {
        const int64_t i = base + threadIdx.x;
        int bin = -1;
        int64_t t = 0;
        if (i < g.m)
        {
            const int64_t a0 = g.Ap[i], a1 = g.Ap[i + 1];
            int64_t f = 0;
            for (int64_t e = a0; e < a1; e++)
            {
                const int64_t k = g.Aj[e];
                f += g.Bp[k + 1] - g.Bp[k];
            }
            t = f;
            if (g.mode == MODE_MASK)
            {
                const int64_t mi = g.Mp[i + 1] - g.Mp[i];
                if (mi < t) t = mi;
            }
            if (t > g.n) t = g.n;
            if (f == 0 || t == 0) { g.Cp[i] = 0; }
            else if (a1 - a0 == 1) bin = BIN_SINGLE;
            else if (t <= g.ws_max) bin = BIN_WS;
            else if (t <= g.wl_max) bin = BIN_WL;
            else if (t <= g.blk_max) bin = BIN_BLK;
            else bin = BIN_GLB;
        }
        #pragma unroll
        for (int b = 0; b < SX_NBINS; b++)
        {
            const bool take = (bin == b);
            const unsigned long long mask = __ballot(take);
            if (mask == 0) continue;
            const int lane = threadIdx.x & 63;
            const int leader = __ffsll((long long) mask) - 1;
            unsigned long long slot = 0;
            if (lane == leader) slot = atomicAdd(&g.cnt[b], (unsigned long long) __popcll(mask));
            slot = __shfl(slot, leader, 64) + (unsigned long long) __popcll(mask & ((1ull << lane) - 1ull));
            if (take)
            {
                g.rows[(int64_t) b * g.m + (int64_t) slot] = (uint32_t) i;
                if (b == BIN_GLB)
                {
                    // table of 2^ceil(log2(2t)) slots, load factor <= 1/2
                    const int lg = ilog2_ceil(2 * t);
                    const unsigned long long sz = 1ull << lg;
                    g.gtab_off[slot] = (int64_t) atomicAdd(&g.cnt[SX_NBINS], sz);
                    g.gtab_log2[slot] = lg;
                }
            }
        }
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
                    slot = sx_insert(keys, logT, (uint32_t) j, created);
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
static int sx_launch_phase(SxArgs &g, const int64_t *cnt)
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
    if (cnt[BIN_WS] > 0)
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

    DevBuf Cp, rows, cnt, goff, glog;
    if (!Cp.alloc(sizeof(int64_t) * (m + 1)) || !rows.alloc(sizeof(uint32_t) * SX_NBINS * (m > 0 ? m : 1)) ||
        !cnt.alloc(sizeof(unsigned long long) * 8) || !goff.alloc(sizeof(int64_t) * (m > 0 ? m : 1)) ||
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
    g.ws_max = (prm && prm->wave_small_max > 0) ? prm->wave_small_max : 64;
    g.wl_max = (prm && prm->wave_large_max > 0) ? prm->wave_large_max : 512;
    g.blk_max = (prm && prm->block_max > 0) ? prm->block_max : 2048;
    if (g.ws_max > 64) g.ws_max = 64;
    if (g.wl_max > 512) g.wl_max = 512;
    if (g.blk_max > 2048) g.blk_max = 2048;

    int64_t hc[8] = {0};
    if (m > 0)
    {
        k_sx_zero<<<1, 64, 0, st>>>(g.cnt, 8);
        {
            ProfScope ps("sx_flops");
            k_sx_flops<<<sx_grid(m, SX_NT, (int64_t) cu_count() * 16), SX_NT, 0, st>>>(g);
            GM_HIP(hipGetLastError(), "k_sx_flops");
        }
        unsigned long long *hp = (unsigned long long *) pinned_scratch(8 * sizeof(unsigned long long));
        if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
        GM_HIP(hipMemcpyAsync(hp, g.cnt, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
               "bin counts");
        GM_HIP(hipStreamSynchronize(st), "bin counts sync");
        for (int t = 0; t < 8; t++) hc[t] = (int64_t) hp[t];
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
    // symbolic
    if (m > 0) GM_TRY((sx_launch_phase<SxSym, 0>(g, hc)));
    GM_TRY(scan_exclusive_i64(g.Cp, g.Cp, m));
    int64_t nvals = 0;
    {
        int64_t *hp = (int64_t *) pinned_scratch(sizeof(int64_t));
        GM_HIP(hipMemcpyAsync(hp, g.Cp + m, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz(C)");
        GM_HIP(hipStreamSynchronize(st), "nnz(C) sync");
        nvals = *hp;
    }
    DevBuf Ci, Cx;
    if (!Ci.alloc(sizeof(int64_t) * (nvals > 0 ? nvals : 1)) ||
        !Cx.alloc(S::iso_out ? zs : zs * (nvals > 0 ? nvals : 1)))
        return set_error(GM_OUT_OF_MEMORY, "saxpy3: output (%lld entries)", (long long) nvals);
    g.Ci = Ci.as<int64_t>();
    g.Cx = Cx.ptr;
    if (nvals > 0) GM_TRY((sx_launch_phase<S, 1>(g, hc)));
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
