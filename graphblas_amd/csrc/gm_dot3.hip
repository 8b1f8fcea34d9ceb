// gm_dot3.hip: masked dot-product SpGEMM, C<M> = A*B' (the reference's dot3).
//
// Reference: Source/mxm/GB_AxB_dot3.c:37-357 (driver), template/
// GB_AxB_dot3_template.c:13-169 (per mask entry, zombies) and
// template/GB_AxB_dot_cij.c:516-652 (sparse x sparse dot product).
//
// All operands are CSR.  For each mask entry (i,j): C(i,j) = A(i,:) . B(j,:)
// under the semiring.  Work is organised per mask entry (never per row, so a
// hub row of M does not serialise):
//
//  bitmaps the longest rows of A and B (within a memory budget) get a bit
//          array over their column range in HBM, plus a per-word rank for
//          valued semirings.  A dot product whose longer row has a bitmap
//          costs min(a,b) independent bit probes instead of a merge.
//  phase1  one thread per entry: row lookup, mask test, sizes a=|A(i,:)|,
//          b=|B(j,:)|, work = min(a,b) with a bitmap, else
//          min(a+b, s*ceil(log2(l+1))).  Small work is done in place by that
//          thread; larger entries are appended (wavefront-aggregated atomics)
//          to one of three bins.
//  team16 / wave / block: 16, 64 or 256 lanes per entry.  A balanced pair is
//          split by merge-path diagonals; an unbalanced pair (s*log l < a+b)
//          is split over the short list, each lane binary-searching the long
//          one.  Lanes reduce in lane order (= k order), so fp64 sums are
//          deterministic and ANY keeps the first match in k order, like the
//          reference's sequential merge.
//  compact one pass per 2048-entry tile: block scan of the live flags, scatter
//          of the surviving (j, value) pairs, and the new row pointers of all
//          rows that start in the tile.  The result has no zombies.

#include <cstring>
#include <type_traits>
#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>
#include "gm_internal.h"
#include "gm_block.h"

namespace gm
{

constexpr int D3_NT = 256;            // threads per workgroup everywhere here
constexpr int D3_P1_PER_THREAD = 4;   // entries per thread per phase1 tile
constexpr int D3_P1_TILE = D3_NT * D3_P1_PER_THREAD;
constexpr int D3_CT_IPT = 8;          // compaction items per thread
constexpr int D3_CT_TILE = D3_NT * D3_CT_IPT;

// Probe indexes of the long rows of one operand (rows of length >= T).  A row
// with column range [lo, hi] gets
//   a bitmap  when its words are at most BM_DENSITY times its length
//             (bm_eligible; on RMAT 22, 32 builds the indexes 0.6 ms faster
//             than 64 at the same probe times, and 16 probes 2.8 ms slower):
//             ((hi-lo)>>6)+1 words, and rank[w] (valued semirings) counts
//             the row's entries before word w;
//   a hash    otherwise: 2^ceil(log2(2 len)) slots of (column << 32 | position),
//             linear probing, EMPTY = all ones.
// off[r] is the row's first 8-byte unit in the shared pool words[].
constexpr int64_t BM_DENSITY = 32;
constexpr uint64_t HX_EMPTY = ~0ull;

struct RowBits
{
    int64_t T;                    // INT64_MAX: no bitmaps
    const int64_t *off;
    const uint64_t *words;
    const uint32_t *rank;
};

__host__ __device__ __forceinline__ bool bm_eligible(int64_t T, int64_t len, int64_t lo, int64_t hi)
{
    return len >= T && ((hi - lo) >> 6) + 1 <= BM_DENSITY * len;
}

__device__ __forceinline__ int hx_log2(int64_t len) { return ilog2_ceil(2 * len); }

// 8-byte units of a long row's probe index
__device__ __forceinline__ int64_t px_units(int64_t len, int64_t lo, int64_t hi)
{
    return bm_eligible(0, len, lo, hi) ? ((hi - lo) >> 6) + 1 : ((int64_t) 1 << hx_log2(len));
}

__device__ __forceinline__ uint32_t hx_hash(uint32_t key, int lg)
{
    return (key * 0x9E3779B1u) >> (32 - lg);
}

// One row's summary for pair setup, read as one aligned 32-byte record:
// start, length, first and last column, and its probe index (kind -1: none,
// 0: bitmap, else log2 of the hash size) at unit offset off of the pool.
// Built when the columns fit 31 bits.
struct __align__(32) RowDesc
{
    int64_t p0;
    int64_t off;
    int32_t len, lo, hi, kind;
};

struct Dot3Args
{
    const int64_t *Mp, *Mj;
    const void *Mx;
    int32_t msize, m_iso, mask_struct;
    int64_t m_nrows, mnz;
    const int64_t *Ap, *Aj;
    const void *Ax;
    int32_t a_iso;
    const int64_t *Bp, *Bj;
    const void *Bx;
    int32_t b_iso;
    void *W;                      // one Z per mask entry
    uint8_t *live;                // 1 if C(i,j) exists
    uint64_t *bins;               // 3 bins x mnz, packed (i << 32) | p
    unsigned long long *bin_cnt;  // 3 counters
    int64_t thread_max, team16_max, wave_max;
    RowBits bA, bB;               // bitmaps of the long rows of A and B
    int32_t cols32;               // column indices fit 32 bits (LDS staging allowed)
    uint64_t *bkeys;              // NULL, or 3 x mnz grouping keys of the binned entries
    const uint8_t *rowmode;       // NULL, or 1 per mask row handled whole by k_dot3_rows
    int64_t key_nA;               // probe-2 keys are key_nA + j
    const RowDesc *dA, *dB;       // NULL, or one summary per row of A / B
};

// lane-aggregated append: returns the slot for lanes with pred set.
// Every lane of the wavefront must call it.
__device__ __forceinline__ unsigned long long wave_append(bool pred, unsigned long long *counter)
{
    const unsigned long long mask = __ballot(pred);
    if (mask == 0) return 0;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long) mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long) __popcll(mask));
    base = __shfl(base, leader, 64);
    const unsigned long long below = mask & ((1ull << lane) - 1ull);
    return base + (unsigned long long) __popcll(below);
}

// One dot product's operands: row ranges, and how to intersect them.
//   probe 0: merge / binary search (no bitmap on the longer row)
//   probe 1: the elements of B(j,:) are probed in A(i,:)'s bitmap
//   probe 2: the elements of A(i,:) are probed in B(j,:)'s bitmap
struct Pair
{
    int64_t pa, pa_end, pb, pb_end;
    int64_t off, lo, hi;          // probe index of the longer row (merges: lo and hi are
                                  // the last columns of A's and B's rows)
    int64_t slo, shi;             // first / last column of the probing (short) row
    int probe;
    int hlg;                      // 0: bitmap, else log2 of the hash size
    bool empty;                   // no possible common index
};

__device__ __forceinline__ int64_t d3_work_merge(int64_t a, int64_t b)
{
    const int64_t s = a < b ? a : b, l = a < b ? b : a;
    const int64_t srch = s * (int64_t) (ilog2_ceil(l + 1) + 1);
    return srch < a + b ? srch : a + b;
}

// fill the pair for A(i,:) . B(j,:); returns the estimated work
__device__ __forceinline__ int64_t d3_pair(const Dot3Args &g, int64_t i, int64_t j, Pair &P)
{
    if (g.dA)
    {
        const RowDesc da = g.dA[i], db = g.dB[j];
        P.pa = da.p0; P.pa_end = da.p0 + da.len;
        P.pb = db.p0; P.pb_end = db.p0 + db.len;
        P.probe = 0;
        P.empty = true;
        const int64_t a = da.len, b = db.len;
        if (a == 0 || b == 0) return 0;
        if (da.hi < db.lo || db.hi < da.lo) return 0;
        P.empty = false;
        if (a >= b && da.kind >= 0)
        {
            P.probe = 1; P.off = da.off; P.lo = da.lo; P.hi = da.hi; P.hlg = da.kind;
            P.slo = db.lo; P.shi = db.hi;
            return b;
        }
        if (db.kind >= 0)
        {
            P.probe = 2; P.off = db.off; P.lo = db.lo; P.hi = db.hi; P.hlg = db.kind;
            P.slo = da.lo; P.shi = da.hi;
            return a;
        }
        if (da.kind >= 0)
        {
            P.probe = 1; P.off = da.off; P.lo = da.lo; P.hi = da.hi; P.hlg = da.kind;
            P.slo = db.lo; P.shi = db.hi;
            return b;
        }
        P.lo = da.hi; P.hi = db.hi;    // merges: the rows' last columns
        return d3_work_merge(a, b);
    }
    P.pa = g.Ap[i]; P.pa_end = g.Ap[i + 1];
    P.pb = g.Bp[j]; P.pb_end = g.Bp[j + 1];
    P.probe = 0;
    P.empty = true;
    const int64_t a = P.pa_end - P.pa, b = P.pb_end - P.pb;
    if (a == 0 || b == 0) return 0;
    const int64_t alo = g.Aj[P.pa], ahi = g.Aj[P.pa_end - 1];
    const int64_t blo = g.Bj[P.pb], bhi = g.Bj[P.pb_end - 1];
    if (ahi < blo || bhi < alo) return 0;
    P.empty = false;
    // probe the shorter row into the longer one's index when it has one
    if (a >= b && a >= g.bA.T)
    {
        P.probe = 1; P.off = g.bA.off[i]; P.lo = alo; P.hi = ahi;
        P.hlg = bm_eligible(0, a, alo, ahi) ? 0 : hx_log2(a);
        P.slo = blo; P.shi = bhi;
        return b;
    }
    if (b >= g.bB.T)
    {
        P.probe = 2; P.off = g.bB.off[j]; P.lo = blo; P.hi = bhi;
        P.hlg = bm_eligible(0, b, blo, bhi) ? 0 : hx_log2(b);
        P.slo = alo; P.shi = ahi;
        return a;
    }
    if (a >= g.bA.T)
    {
        P.probe = 1; P.off = g.bA.off[i]; P.lo = alo; P.hi = ahi;
        P.hlg = bm_eligible(0, a, alo, ahi) ? 0 : hx_log2(a);
        P.slo = blo; P.shi = bhi;
        return b;
    }
    P.lo = ahi; P.hi = bhi;
    return d3_work_merge(a, b);
}

// position of x in the probed row (row start p0), or -1
template <bool NEED_POS>
__device__ __forceinline__ int64_t bm_find(const RowBits &bm, const Pair &P, int64_t p0, int64_t x)
{
    if (x < P.lo || x > P.hi) return -1;
    if (P.hlg)
    {
        const uint64_t *T = bm.words + P.off;
        const uint32_t mask = (1u << P.hlg) - 1u;
        uint32_t h = hx_hash((uint32_t) x, P.hlg);
        while (true)
        {
            const uint64_t e = T[h];
            if (e == HX_EMPTY) return -1;
            if ((int64_t) (e >> 32) == x) return p0 + (int64_t) (e & 0xffffffffull);
            h = (h + 1) & mask;
        }
    }
    const uint64_t d = (uint64_t) (x - P.lo);
    const uint64_t w = bm.words[P.off + (int64_t) (d >> 6)];
    const uint64_t bit = 1ull << (d & 63);
    if (!(w & bit)) return -1;
    if constexpr (!NEED_POS) return p0;
    else return p0 + (int64_t) bm.rank[P.off + (int64_t) (d >> 6)] + __popcll(w & (bit - 1));
}

// Two halves of bm_find, so that several probes' first reads can be issued
// before any is used: the index unit x lands on (a bitmap word or its home
// hash slot; unit 0 of the row's index when x is outside the row's range,
// a safe read whose value is ignored), then the position from that unit
// (hash: further slots on a collision).
__device__ __forceinline__ uint64_t bm_first(const RowBits &bm, const Pair &P, int64_t x)
{
    const bool inr = x >= P.lo && x <= P.hi;
    int64_t u = 0;
    if (inr) u = P.hlg ? (int64_t) hx_hash((uint32_t) x, P.hlg) : (int64_t) ((uint64_t) (x - P.lo) >> 6);
    return bm.words[P.off + u];
}

template <bool NEED_POS>
__device__ __forceinline__ int64_t bm_resolve(const RowBits &bm, const Pair &P, int64_t p0, int64_t x,
                                              uint64_t w)
{
    if (x < P.lo || x > P.hi) return -1;
    if (P.hlg)
    {
        const uint64_t *T = bm.words + P.off;
        const uint32_t mask = (1u << P.hlg) - 1u;
        uint32_t h = hx_hash((uint32_t) x, P.hlg);
        while (true)
        {
            if (w == HX_EMPTY) return -1;
            if ((int64_t) (w >> 32) == x) return p0 + (int64_t) (w & 0xffffffffull);
            h = (h + 1) & mask;
            w = T[h];
        }
    }
    const uint64_t d = (uint64_t) (x - P.lo);
    const uint64_t bit = 1ull << (d & 63);
    if (!(w & bit)) return -1;
    if constexpr (!NEED_POS) return p0;
    else return p0 + (int64_t) bm.rank[P.off + (int64_t) (d >> 6)] + __popcll(w & (bit - 1));
}

// probe elements q0, q0+step, ... < q1 of the short row in the long row's
// index; the iterations are independent, so several probes are in flight
// STOP: the lane leaves after the first step that passes the long row's
// range (the short row is sorted), which spares the team path a second
// search; steps of four elements keep their loads and probes issued
// together (team16 bin at RMAT 22: 25.9 ms with an exit per element, 23.1
// ms in steps of four).  Without STOP the loop has no data-dependent exit.
constexpr int PR_UN = 4;                      // elements per lane per probe step
constexpr int MERGE_BLK = 4;                  // block width of the serial merge
template <class S, bool STOP>
__device__ __forceinline__ void d3_probe_range(const Dot3Args &g, const Pair &P, int64_t q0, int64_t q1,
                                               int64_t step, bool &e, typename S::Z &z)
{
    using Z = typename S::Z;
    const Z *Ax = (const Z *) g.Ax;
    const Z *Bx = (const Z *) g.Bx;
    const bool shortA = (P.probe == 2);
    const int64_t *Sj = shortA ? g.Aj : g.Bj;
    const RowBits &bm = shortA ? g.bB : g.bA;
    const int64_t lp0 = shortA ? P.pb : P.pa;
    if constexpr (STOP)
    {
        // PR_UN elements per step: their loads, then their probes, issue
        // together; the lane stops after a step that passed the range
        for (int64_t q = q0; q < q1; q += PR_UN * step)
        {
            // past the end: the last element read again (no branch around
            // the loads), then a sentinel above every column
            int64_t xs[PR_UN];
            #pragma unroll
            for (int u = 0; u < PR_UN; u++)
            {
                const int64_t qq = q + u * step;
                xs[u] = Sj[qq < q1 ? qq : q1 - 1];
            }
            #pragma unroll
            for (int u = 0; u < PR_UN; u++)
                if (q + u * step >= q1) xs[u] = INT64_MAX;
            // the probes' first index reads, issued together
            uint64_t ws[PR_UN];
            #pragma unroll
            for (int u = 0; u < PR_UN; u++) ws[u] = bm_first(bm, P, xs[u]);
            bool done = false;
            #pragma unroll
            for (int u = 0; u < PR_UN; u++)
            {
                const int64_t x = xs[u];
                if (x > P.hi) { done = true; break; }     // also the end of the row
                const int64_t pos = bm_resolve<!S::pattern>(bm, P, lp0, x, ws[u]);
                if (pos >= 0)
                {
                    const int64_t qq = q + u * step;
                    const int64_t qa = shortA ? qq : pos, qb = shortA ? pos : qq;
                    Z t = S::mul(S::geta(Ax, qa, g.a_iso), S::geta(Bx, qb, g.b_iso), x);
                    if (!e) { z = t; e = true; }
                    else if constexpr (S::any_min_k) { if (t < z) z = t; }
                    else z = S::upd(z, t);
                    if constexpr (S::is_any && !S::any_min_k) { done = true; break; }
                    if (S::terminal(z) && !S::any_min_k) { done = true; break; }
                }
            }
            if (done) break;
        }
        return;
    }
    #pragma unroll 4
    for (int64_t q = q0; q < q1; q += step)
    {
        const int64_t x = Sj[q];
        const int64_t pos = bm_find<!S::pattern>(bm, P, lp0, x);
        if (pos >= 0)
        {
            const int64_t qa = shortA ? q : pos, qb = shortA ? pos : q;
            Z t = S::mul(S::geta(Ax, qa, g.a_iso), S::geta(Bx, qb, g.b_iso), x);
            if (!e) { z = t; e = true; }
            else if constexpr (S::any_min_k) { if (t < z) z = t; }
            else z = S::upd(z, t);
            if constexpr (S::is_any && !S::any_min_k) break;
            if (S::terminal(z) && !S::any_min_k) break;
        }
    }
}

// trim the short row of a probe pair to the overlap of both rows' ranges
__device__ __forceinline__ void d3_trim_short(const Dot3Args &g, const Pair &P, int64_t &q0, int64_t &q1)
{
    const bool shortA = (P.probe == 2);
    const int64_t *Sj = shortA ? g.Aj : g.Bj;
    int64_t s0 = shortA ? P.pa : P.pb, s1 = shortA ? P.pa_end : P.pb_end;
    const int64_t slo = P.slo, shi = P.shi;
    const int64_t olo = slo > P.lo ? slo : P.lo, ohi = shi < P.hi ? shi : P.hi;
    if (olo > slo) s0 = lower_bound(Sj, s0, s1, olo);
    if (ohi < shi) s1 = lower_bound(Sj, s0, s1, ohi + 1);
    q0 = s0;
    q1 = s1;
}

// First index in [lo, hi) with X[idx] >= v (hi if none), found by the G lanes
// of a team together (G <= 64, one wavefront): each round the lanes sample G
// evenly spaced positions and a ballot narrows the range G-fold, so the
// search costs log_G(hi-lo) rounds of independent loads instead of
// log_2(hi-lo) dependent ones.  Every lane of the team must call it.
template <int G, class T = int64_t>
__device__ __forceinline__ int64_t team_lower_bound(const T *X, int64_t lo, int64_t hi, int64_t v,
                                                    int lane)
{
    static_assert(G <= 64, "one wavefront");
    const int base = (int) (threadIdx.x & 63) & ~(G - 1);
    const unsigned long long tmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    while (hi > lo)
    {
        const int64_t len = hi - lo;
        const int64_t stride = (len + G - 1) / G;
        const int64_t pos = lo + (int64_t) lane * stride;
        const bool less = pos < hi && (int64_t) X[pos] < v;
        const int cnt = __popcll((__ballot(less) >> base) & tmask);
        if (stride == 1) return lo + cnt;
        if (cnt == 0) return lo;
        const int64_t nlo = lo + (int64_t) (cnt - 1) * stride + 1;
        const int64_t nhi = lo + (int64_t) cnt * stride;
        lo = nlo;
        hi = nhi < hi ? nhi : hi;
    }
    return lo;
}

// d3_trim_short with the two searches done by the team
template <int G>
__device__ __forceinline__ void d3_trim_short_team(const Dot3Args &g, const Pair &P, int lane,
                                                   int64_t &q0, int64_t &q1)
{
    const bool shortA = (P.probe == 2);
    const int64_t *Sj = shortA ? g.Aj : g.Bj;
    int64_t s0 = shortA ? P.pa : P.pb, s1 = shortA ? P.pa_end : P.pb_end;
    const int64_t slo = P.slo;
    const int64_t olo = slo > P.lo ? slo : P.lo;
    // only the start is searched: the probe loop stops at the first element
    // past the long row's range
    if (olo > slo) s0 = team_lower_bound<G>(Sj, s0, s1, olo, lane);
    q0 = s0;
    q1 = s1;
}

// Single-thread dot product: bitmap probes, or a linear merge / binary trim
// of the longer list when it is much longer (the reference's rule at
// GB_AxB_dot_cij.c:538-616, here chosen by estimated cost).
template <class S>
__device__ bool dot_serial(const Dot3Args &g, const Pair &P, typename S::Z &out)
{
    using Z = typename S::Z;
    bool e = false;
    Z z = S::identity();
    if (P.empty) return false;
    if (P.probe)
    {
        if (P.probe == 2) d3_probe_range<S, true>(g, P, P.pa, P.pa_end, 1, e, z);
        else d3_probe_range<S, true>(g, P, P.pb, P.pb_end, 1, e, z);
        if constexpr (S::plus_pair_wide) e = (z != 0);
        out = z;
        return e;
    }
    const int64_t *Aj = g.Aj, *Bj = g.Bj;
    const Z *Ax = (const Z *) g.Ax, *Bx = (const Z *) g.Bx;
    int64_t pa = P.pa, pb = P.pb;
    const int64_t pa_end = P.pa_end, pb_end = P.pb_end;
    const int64_t a = pa_end - pa, b = pb_end - pb;
    const bool trimA = a > 8 * b, trimB = b > 8 * a;
    if (!trimA && !trimB)
    {
        // blocks of MERGE_BLK from each row, their loads issued together; all
        // pairs of a block compared in registers, then the block with the
        // smaller last column (or both) moves on.  Matches come in
        // ascending column order, as in the element-wise merge.
        const int64_t ahi = P.lo, bhi = P.hi;
        constexpr int MB = MERGE_BLK;
        while (pa < pa_end && pb < pb_end)
        {
            int64_t xa[MB], xb[MB];
            #pragma unroll
            for (int u = 0; u < MB; u++)
            {
                xa[u] = Aj[pa + u < pa_end ? pa + u : pa_end - 1];
                xb[u] = Bj[pb + u < pb_end ? pb + u : pb_end - 1];
            }
            // past a row's end: sentinels that match nothing
            #pragma unroll
            for (int u = 0; u < MB; u++)
            {
                if (pa + u >= pa_end) xa[u] = INT64_MAX;
                if (pb + u >= pb_end) xb[u] = INT64_MAX - 1;
            }
            if (xa[0] > bhi || xb[0] > ahi) break;
            bool stop = false;
            #pragma unroll
            for (int u = 0; u < MB; u++)
            {
                #pragma unroll
                for (int v = 0; v < MB; v++)
                {
                    if (xa[u] == xb[v] && !stop)
                    {
                        Z t = S::mul(S::geta(Ax, pa + u, g.a_iso), S::geta(Bx, pb + v, g.b_iso), xa[u]);
                        if (!e) { z = t; e = true; } else z = S::upd(z, t);
                        if constexpr (S::is_any) stop = true;
                        if (S::terminal(z)) stop = true;
                    }
                }
            }
            if (stop) break;
            const int64_t am = xa[MB - 1], bm = xb[MB - 1];
            if (am <= bm) pa += MB;
            if (bm <= am) pb += MB;
        }
        if constexpr (S::plus_pair_wide) e = (z != 0);
        out = z;
        return e;
    }
    int64_t ia = Aj[pa], ib = Bj[pb];
    while (true)
    {
        if (ia < ib)
        {
            if (trimA) pa = lower_bound(Aj, pa + 1, pa_end, ib); else pa++;
            if (pa >= pa_end) break;
            ia = Aj[pa];
        }
        else if (ib < ia)
        {
            if (trimB) pb = lower_bound(Bj, pb + 1, pb_end, ia); else pb++;
            if (pb >= pb_end) break;
            ib = Bj[pb];
        }
        else
        {
            Z t = S::mul(S::geta(Ax, pa, g.a_iso), S::geta(Bx, pb, g.b_iso), ia);
            if (!e) { z = t; e = true; } else z = S::upd(z, t);
            if constexpr (S::is_any) break;
            if (S::terminal(z)) break;
            pa++;
            pb++;
            if (pa >= pa_end || pb >= pb_end) break;
            ia = Aj[pa];
            ib = Bj[pb];
        }
    }
    if constexpr (S::plus_pair_wide) e = (z != 0);
    out = z;
    return e;
}

// One lane's share of a team dot product.  lane in [0,G).  Returns the
// lane's partial (e, z); partials combine in lane order (= k order).
template <class S, int G>
__device__ void dot_lane(const Dot3Args &g, const Pair &P, int lane, bool &e, typename S::Z &z)
{
    using Z = typename S::Z;
    e = false;
    z = S::identity();
    if (P.empty) return;
    const int64_t *Aj = g.Aj, *Bj = g.Bj;
    const Z *Ax = (const Z *) g.Ax, *Bx = (const Z *) g.Bx;
    const int64_t pa = P.pa, pa_end = P.pa_end, pb = P.pb, pb_end = P.pb_end;
    const int64_t a = pa_end - pa, b = pb_end - pb;
    if (P.probe)
    {
        // the short row from its first element in the long row's range
        // (teams: lanes stop past the range; a workgroup trims both ends),
        // interleaved over the lanes (coalesced reads), each element probed
        // in the long row's index
        int64_t q0, q1;
        if constexpr (G <= 64) d3_trim_short_team<G>(g, P, lane, q0, q1);
        else d3_trim_short(g, P, q0, q1);
        d3_probe_range<S, (G <= 64)>(g, P, q0 + lane, q1, G, e, z);
        return;
    }
    const int64_t s = a < b ? a : b, l = a < b ? b : a;
    const bool search = s * (int64_t) (ilog2_ceil(l + 1) + 1) < a + b;
    if (search)
    {
        // split the short list into G contiguous chunks; binary search each
        // element in the long list from the previous hit onward
        const bool shortA = a <= b;
        const int64_t *Sj = shortA ? Aj : Bj;
        const int64_t *Lj = shortA ? Bj : Aj;
        const int64_t s0 = shortA ? pa : pb;
        int64_t lcur = shortA ? pb : pa;
        const int64_t lend = shortA ? pb_end : pa_end;
        const int64_t chunk = (s + G - 1) / G;
        const int64_t q0 = s0 + lane * chunk;
        const int64_t q1 = (s0 + s) < (q0 + chunk) ? (s0 + s) : (q0 + chunk);
        for (int64_t q = q0; q < q1; q++)
        {
            const int64_t x = Sj[q];
            lcur = lower_bound(Lj, lcur, lend, x);
            if (lcur >= lend) break;
            if (Lj[lcur] == x)
            {
                const int64_t qa = shortA ? q : lcur, qb = shortA ? lcur : q;
                Z t = S::mul(S::geta(Ax, qa, g.a_iso), S::geta(Bx, qb, g.b_iso), x);
                if (!e) { z = t; e = true; } else z = S::upd(z, t);
                if constexpr (S::is_any) break;
                if (S::terminal(z)) break;
                lcur++;
            }
        }
    }
    else
    {
        // merge path: diagonals d0..d1 of the (a+b)-long merge, A first on
        // ties; a match is taken when B[y] is consumed right after an equal
        // A[x-1], so each common index is seen by exactly one lane
        const int64_t tot = a + b;
        const int64_t d0 = (tot * lane) / G, d1 = (tot * (lane + 1)) / G;
        int64_t lo = d0 > b ? d0 - b : 0, hi = d0 < a ? d0 : a;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi) >> 1;
            if (Aj[pa + mid] <= Bj[pb + d0 - mid - 1]) lo = mid + 1; else hi = mid;
        }
        int64_t x = lo, y = d0 - lo;
        for (int64_t d = d0; d < d1; d++)
        {
            if (x < a && (y >= b || Aj[pa + x] <= Bj[pb + y]))
            {
                x++;
            }
            else
            {
                const int64_t kb = Bj[pb + y];
                if (x > 0 && Aj[pa + x - 1] == kb)
                {
                    Z t = S::mul(S::geta(Ax, pa + x - 1, g.a_iso), S::geta(Bx, pb + y, g.b_iso), kb);
                    if (!e) { z = t; e = true; } else z = S::upd(z, t);
                    if constexpr (S::is_any) break;
                    if (S::terminal(z)) break;
                }
                y++;
            }
        }
    }
}

// LDS words per team for staged pairs (16 KB per workgroup, so LDS does not
// cap the occupancy): both index lists (as 32-bit columns)
// of a merge/search pair with a+b <= cap are copied to LDS by the team with
// coalesced loads, then intersected there, so the merge walk and the binary
// searches pay LDS latency instead of a chain of dependent global loads.
template <int G> struct StageCap { static constexpr int value = G == 16 ? 256 : 1024; };

// make the team's LDS stores visible to the team's later loads (the team is
// part of one wavefront, so no workgroup barrier is needed)
__device__ __forceinline__ void team_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Same partition of the work over lanes as dot_lane's merge/search cases, on
// LDS copies sa[0..a) and sb[0..b) of the two lists.
template <class S, int G>
__device__ void dot_lane_staged(const Dot3Args &g, const Pair &P, int lane, uint32_t *sl, bool &e,
                                typename S::Z &z)
{
    using Z = typename S::Z;
    e = false;
    z = S::identity();
    const Z *Ax = (const Z *) g.Ax, *Bx = (const Z *) g.Bx;
    const int64_t pa = P.pa, pb = P.pb;
    const int a = (int) (P.pa_end - pa), b = (int) (P.pb_end - pb);
    for (int t = lane; t < a; t += G) sl[t] = (uint32_t) g.Aj[pa + t];
    for (int t = lane; t < b; t += G) sl[a + t] = (uint32_t) g.Bj[pb + t];
    team_lds_sync();
    const uint32_t *sa = sl, *sb = sl + a;
    const int s = a < b ? a : b, l = a < b ? b : a;
    const bool search = (int64_t) s * (int64_t) (ilog2_ceil(l + 1) + 1) < (int64_t) (a + b);
    if (search)
    {
        const bool shortA = a <= b;
        const uint32_t *Sj = shortA ? sa : sb;
        const uint32_t *Lj = shortA ? sb : sa;
        int lcur = 0;
        const int chunk = (s + G - 1) / G;
        const int q0 = lane * chunk;
        const int q1 = s < q0 + chunk ? s : q0 + chunk;
        for (int q = q0; q < q1; q++)
        {
            const uint32_t x = Sj[q];
            int lo = lcur, hi = l;
            while (lo < hi)
            {
                const int mid = (lo + hi) >> 1;
                if (Lj[mid] < x) lo = mid + 1; else hi = mid;
            }
            lcur = lo;
            if (lcur >= l) break;
            if (Lj[lcur] == x)
            {
                const int64_t qa = shortA ? pa + q : pa + lcur, qb = shortA ? pb + lcur : pb + q;
                Z t = S::mul(S::geta(Ax, qa, g.a_iso), S::geta(Bx, qb, g.b_iso), (int64_t) x);
                if (!e) { z = t; e = true; } else z = S::upd(z, t);
                if constexpr (S::is_any) break;
                if (S::terminal(z)) break;
                lcur++;
            }
        }
    }
    else
    {
        const int tot = a + b;
        const int d0 = (tot * lane) / G, d1 = (tot * (lane + 1)) / G;
        int lo = d0 > b ? d0 - b : 0, hi = d0 < a ? d0 : a;
        while (lo < hi)
        {
            const int mid = (lo + hi) >> 1;
            if (sa[mid] <= sb[d0 - mid - 1]) lo = mid + 1; else hi = mid;
        }
        int x = lo, y = d0 - lo;
        for (int d = d0; d < d1; d++)
        {
            if (x < a && (y >= b || sa[x] <= sb[y]))
            {
                x++;
            }
            else
            {
                const uint32_t kb = sb[y];
                if (x > 0 && sa[x - 1] == kb)
                {
                    Z t = S::mul(S::geta(Ax, pa + x - 1, g.a_iso), S::geta(Bx, pb + y, g.b_iso),
                                 (int64_t) kb);
                    if (!e) { z = t; e = true; } else z = S::upd(z, t);
                    if constexpr (S::is_any) break;
                    if (S::terminal(z)) break;
                }
                y++;
            }
        }
    }
    // the next pair of this team overwrites the lists
    team_lds_sync();
}

// final existence rule and store of one entry
template <class S>
__device__ __forceinline__ void d3_store(const Dot3Args &g, int64_t p, bool e, typename S::Z z)
{
    if constexpr (S::plus_pair_wide) e = (z != 0);
    g.live[p] = e ? 1 : 0;
    if (e && !S::iso_out) ((typename S::Z *) g.W)[p] = z;
}

//------------------------------------------------------------------------------
// phase1: classify every mask entry, finish the small ones
//------------------------------------------------------------------------------

template <class S>
__global__ __launch_bounds__(D3_NT) void k_dot3_phase1(Dot3Args g)
{
    using Z = typename S::Z;
    __shared__ int64_t s_rows[2];
    // the tile's row pointers Mp[r0 .. r1+1], when the window is short
    __shared__ int64_t s_mp[D3_P1_TILE + 2];
    // the tile's binned entries: bin (-1: none), record, grouping key; the
    // wavefronts' ballots per bin, their offsets in the tile, the bases
    __shared__ int8_t s_bin[D3_P1_TILE];
    __shared__ uint64_t s_rec[D3_P1_TILE], s_bkey[D3_P1_TILE];
    __shared__ unsigned long long s_mb[D3_P1_PER_THREAD][D3_NT / 64][3];
    __shared__ unsigned long long s_pre[D3_P1_PER_THREAD][D3_NT / 64][3];
    __shared__ unsigned long long s_base[3];
    const int64_t ntiles = (g.mnz + D3_P1_TILE - 1) / D3_P1_TILE;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
    {
        const int64_t t0 = tile * D3_P1_TILE;
        const int64_t t1 = (t0 + D3_P1_TILE < g.mnz) ? t0 + D3_P1_TILE : g.mnz;
        if (threadIdx.x < 2)
        {
            // row containing entry t0 (thread 0) / t1-1 (thread 1)
            const int64_t pp = threadIdx.x == 0 ? t0 : t1 - 1;
            int64_t lo = 0, hi = g.m_nrows - 1;
            while (lo < hi)
            {
                const int64_t mid = (lo + hi + 1) >> 1;
                if (g.Mp[mid] <= pp) lo = mid; else hi = mid - 1;
            }
            s_rows[threadIdx.x] = lo;
        }
        __syncthreads();
        const int64_t r0 = s_rows[0], r1 = s_rows[1];
        const bool lds_rows = (r1 - r0 + 2) <= (int64_t) (D3_P1_TILE + 2);
        if (lds_rows)
            for (int64_t t = threadIdx.x; t <= r1 - r0 + 1; t += D3_NT) s_mp[t] = g.Mp[r0 + t];
        __syncthreads();
        #pragma unroll 1
        for (int u = 0; u < D3_P1_PER_THREAD; u++)
        {
            const int64_t p = t0 + (int64_t) u * D3_NT + threadIdx.x;
            const bool valid = p < t1;
            int bin = -1;
            int64_t i = 0;
            uint64_t key = 0;
            if (valid)
            {
                int64_t lo = r0, hi = r1;
                if (lds_rows)
                {
                    while (lo < hi)
                    {
                        const int64_t mid = (lo + hi + 1) >> 1;
                        if (s_mp[mid - r0] <= p) lo = mid; else hi = mid - 1;
                    }
                }
                else
                {
                    while (lo < hi)
                    {
                        const int64_t mid = (lo + hi + 1) >> 1;
                        if (g.Mp[mid] <= p) lo = mid; else hi = mid - 1;
                    }
                }
                i = lo;
                const int64_t j = g.Mj[p];
                bool e = false;
                Z z = S::identity();
                if (g.rowmode && g.rowmode[i]) bin = 3;          // the row kernel's entry
                else if (g.mask_struct || mask_true(g.Mx, p, g.msize, g.m_iso))
                {
                    Pair P;
                    const int64_t w = d3_pair(g, i, j, P);
                    if (!P.empty)
                    {
                        if (w <= g.thread_max) e = dot_serial<S>(g, P, z);
                        else bin = (w <= g.team16_max) ? 0 : (w <= g.wave_max) ? 1 : 2;
                        // entries probing the same long row's index sort together
                        key = (P.probe == 2) ? (uint64_t) (g.key_nA + j) : (uint64_t) i;
                    }
                }
                if (bin == 3) bin = -2;
                else if (bin < 0)
                {
                    g.live[p] = e ? 1 : 0;
                    if (e && !S::iso_out) ((Z *) g.W)[p] = z;
                }
            }
            // the entry's bin and record, kept in LDS for the tile's appends;
            // lane 0 keeps the wavefront's ballot of each bin
            const int tslot = u * D3_NT + (int) threadIdx.x;
            s_bin[tslot] = (int8_t) (bin >= 0 ? bin : -1);
            if (bin >= 0)
            {
                s_rec[tslot] = ((uint64_t) i << 32) | (uint64_t) p;
                s_bkey[tslot] = key;
            }
            #pragma unroll
            for (int b = 0; b < 3; b++)
            {
                const unsigned long long m = __ballot(bin == b);
                if ((threadIdx.x & 63) == 0) s_mb[u][threadIdx.x >> 6][b] = m;
            }
        }
        // one global atomic per bin per tile: every wavefront appending to
        // three shared counters serialised phase1 on them (RMAT 22: 17 of
        // its 19 ms)
        __syncthreads();
        if (threadIdx.x < 3)
        {
            const int b = (int) threadIdx.x;
            unsigned long long run = 0;
            for (int u = 0; u < D3_P1_PER_THREAD; u++)
                for (int w = 0; w < D3_NT / 64; w++)
                {
                    s_pre[u][w][b] = run;
                    run += (unsigned long long) __popcll(s_mb[u][w][b]);
                }
            s_base[b] = run ? atomicAdd(&g.bin_cnt[b], run) : 0ull;
        }
        __syncthreads();
        #pragma unroll
        for (int u = 0; u < D3_P1_PER_THREAD; u++)
        {
            const int tslot = u * D3_NT + (int) threadIdx.x;
            const int b = s_bin[tslot];
            if (b < 0) continue;
            const int wv = (int) (threadIdx.x >> 6), wl = (int) (threadIdx.x & 63);
            const unsigned long long below = s_mb[u][wv][b] & ((1ull << wl) - 1ull);
            const int64_t slot = (int64_t) (s_base[b] + s_pre[u][wv][b] + (unsigned long long) __popcll(below));
            g.bins[(int64_t) b * g.mnz + slot] = s_rec[tslot];
            if (g.bkeys) g.bkeys[(int64_t) b * g.mnz + slot] = s_bkey[tslot];
        }
        __syncthreads();
    }
}

//------------------------------------------------------------------------------
// team kernels: G lanes per deferred entry (G = 16 or 64)
//------------------------------------------------------------------------------

// one deferred entry w = (i << 32 | p), computed by a team of G lanes (sl:
// the team's LDS staging area for short merges, or NULL for none)
template <class S, int G>
__device__ __forceinline__ void team_task(const Dot3Args &g, uint64_t w, int lane, uint32_t *sl)
{
    using Z = typename S::Z;
    constexpr int CAP = StageCap<G>::value;
    const int64_t i = (int64_t) (w >> 32), p = (int64_t) (w & 0xffffffffull);
    const int64_t j = g.Mj[p];
    bool e;
    Z z;
    Pair P;
    (void) d3_pair(g, i, j, P);
    if (sl && !P.empty && P.probe == 0 && g.cols32 &&
        (P.pa_end - P.pa) + (P.pb_end - P.pb) <= CAP)
        dot_lane_staged<S, G>(g, P, lane, sl, e, z);
    else
        dot_lane<S, G>(g, P, lane, e, z);
    team_reduce<S, G>(e, z);
    if (lane == 0) d3_store<S>(g, p, e, z);
}

// the bin's entries in their appended order, strided over the grid
template <class S, int G>
__global__ __launch_bounds__(D3_NT) void k_dot3_team(Dot3Args g, int bin)
{
    const int64_t count = (int64_t) g.bin_cnt[bin];
    const uint64_t *list = g.bins + (int64_t) bin * g.mnz;
    constexpr int TPB = D3_NT / G;
    __shared__ uint32_t s_lists[TPB][StageCap<G>::value];
    const int lane = threadIdx.x & (G - 1);
    uint32_t *sl = s_lists[threadIdx.x / G];
    for (int64_t q = (int64_t) blockIdx.x * TPB + threadIdx.x / G; q < count;
         q += (int64_t) gridDim.x * TPB)
        team_task<S, G>(g, list[q], lane, sl);
}

// The bin's entries sorted by the row whose probe index they read.  The list
// is cut into 8 regions, one per XCD (workgroups are dispatched to the 8 XCDs
// round-robin, so workgroup b runs on XCD b mod 8): an XCD's workgroups take
// chunks of its own region from a counter, so the entries that probe one
// long row run on one XCD and find that row's index in its L2.  A workgroup
// whose region is used up moves on to the next regions.
template <class S, int G>
__global__ __launch_bounds__(D3_NT) void k_dot3_team_grp(Dot3Args g, const uint64_t *list,
                                                         int64_t count, unsigned long long *ctr)
{
    constexpr int TPB = D3_NT / G;
    constexpr int CH = TPB * 4;
    __shared__ uint32_t s_lists[TPB][StageCap<G>::value];
    __shared__ int64_t s_start;
    const int lane = threadIdx.x & (G - 1);
    const int team = threadIdx.x / G;
    uint32_t *sl = s_lists[team];
    const int x0 = (int) (blockIdx.x & 7);
    for (int k = 0; k < 8; k++)
    {
        const int r = (x0 + k) & 7;
        const int64_t lo = count * r / 8, hi = count * (r + 1) / 8;
        while (true)
        {
            if (threadIdx.x == 0) s_start = lo + (int64_t) atomicAdd(&ctr[r], 1ull) * CH;
            __syncthreads();
            const int64_t start = s_start;
            __syncthreads();
            if (start >= hi) break;
            const int64_t end = start + CH < hi ? start + CH : hi;
            for (int64_t q = start + team; q < end; q += TPB) team_task<S, G>(g, list[q], lane, sl);
        }
    }
}

// one deferred entry computed by a whole 256-thread workgroup (the heaviest bin)
template <class S>
__device__ __forceinline__ void block_task(const Dot3Args &g, uint64_t w, typename S::Z *s_z, int *s_e)
{
    using Z = typename S::Z;
    const int wid = threadIdx.x >> 6;
    const int64_t i = (int64_t) (w >> 32), p = (int64_t) (w & 0xffffffffull);
    const int64_t j = g.Mj[p];
    bool e;
    Z z;
    Pair P;
    (void) d3_pair(g, i, j, P);
    dot_lane<S, D3_NT>(g, P, threadIdx.x, e, z);
    team_reduce<S, 64>(e, z);
    if ((threadIdx.x & 63) == 0) { s_z[wid] = z; s_e[wid] = e ? 1 : 0; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        bool ee = s_e[0] != 0;
        Z zz = s_z[0];
        for (int t = 1; t < D3_NT / 64; t++) combine<S>(ee, zz, s_e[t] != 0, s_z[t]);
        d3_store<S>(g, p, ee, zz);
    }
    __syncthreads();
}

template <class S>
__global__ __launch_bounds__(D3_NT) void k_dot3_block(Dot3Args g)
{
    using Z = typename S::Z;
    __shared__ Z s_z[D3_NT / 64];
    __shared__ int s_e[D3_NT / 64];
    const int64_t count = (int64_t) g.bin_cnt[2];
    const uint64_t *list = g.bins + 2 * g.mnz;
    for (int64_t q = blockIdx.x; q < count; q += gridDim.x) block_task<S>(g, list[q], s_z, s_e);
}

// the heaviest bin sorted by probed row, over the same 8 XCD regions as the
// grouped team kernel (one entry per chunk)
template <class S>
__global__ __launch_bounds__(D3_NT) void k_dot3_block_grp(Dot3Args g, const uint64_t *list,
                                                          int64_t count, unsigned long long *ctr)
{
    using Z = typename S::Z;
    __shared__ Z s_z[D3_NT / 64];
    __shared__ int s_e[D3_NT / 64];
    __shared__ int64_t s_q;
    const int x0 = (int) (blockIdx.x & 7);
    for (int k = 0; k < 8; k++)
    {
        const int r = (x0 + k) & 7;
        const int64_t lo = count * r / 8, hi = count * (r + 1) / 8;
        while (true)
        {
            if (threadIdx.x == 0) s_q = lo + (int64_t) atomicAdd(&ctr[r], 1ull);
            __syncthreads();
            const int64_t q = s_q;
            __syncthreads();
            if (q >= hi) break;
            block_task<S>(g, list[q], s_z, s_e);
        }
    }
}

//------------------------------------------------------------------------------
// LDS probing of the heavy bins (pattern semirings)
//------------------------------------------------------------------------------

// In a bin sorted by probed row, consecutive entries read one long row.
// The sorted bin is cut into fixed slices of LC_SLICE entries; in a slice, a
// run of at least LC_GMIN entries with the same probed long row builds that
// row's bitmap in LDS, LC_BITS columns at a time, from its sorted column
// list (one streamed pass over the row, whatever index it has in HBM), and
// each entry's short row (one wavefront per entry, 64 elements per load)
// is tested against the chunk in LDS instead of with one random L2 access
// per element.  An entry keeps its position in its short row and its
// partial result in LDS from chunk to chunk; elements are visited in
// ascending column order, so partials combine in k order.  Other entries
// of the slice take the global path of the wavefront kernel.  Slices are
// handed out per XCD region, as in the grouped team kernels, so a long
// row's bitmap chunks are read from one L2.
constexpr int64_t LC_POS_MAX = (int64_t) 1 << 28;   // chunk offsets kept (1 GB)
constexpr int LC_WAVE_G = 16;                 // lanes per entry of the wavefront bin
constexpr bool LC_NARROW = true;              // 32-bit short-row columns when they fit
constexpr bool LC_TEAM16 = false;             // the 16-lane bin on the chunk kernel too
constexpr int LC_LG = 20;
constexpr int LC_BITS = 1 << LC_LG;           // columns per chunk: 128 KB of LDS
constexpr int LC_WORDS = LC_BITS / 64;
constexpr int LC_SLICE = 512;
constexpr int LC_GMIN = 8;
constexpr int LC_UNROLL = 16;
constexpr int LC_NT = 1024;                   // 16 wavefronts share a chunk (one workgroup per CU)

template <class S, int G, bool N32>
__global__ __launch_bounds__(LC_NT) void k_dot3_lds(Dot3Args g, const uint64_t *list,
                                                    const uint64_t *keys, int64_t count,
                                                    unsigned long long *ctr, const uint32_t *posA,
                                                    const uint32_t *posB, int ncp, const uint32_t *A32,
                                                    const uint32_t *B32)
{
    using Z = typename S::Z;
    static_assert(S::pattern, "LDS probing covers the pattern semirings");
    static_assert(G <= 64 && (64 % G) == 0, "teams inside one wavefront");
    constexpr int NW = LC_NT / G;            // teams of G lanes, one entry each
    constexpr int UN = N32 ? LC_UNROLL : LC_UNROLL / 2;   // windows per step (64-bit: registers)
    const int tbase = (int) (threadIdx.x & 63) & ~(G - 1);
    const unsigned long long tmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    __shared__ uint64_t s_bits[LC_WORDS];
    __shared__ uint64_t s_key[LC_SLICE];
    __shared__ int64_t s_srow[LC_SLICE], s_rp0[LC_SLICE];   // the short row and its start
    __shared__ uint32_t s_qs[LC_SLICE], s_qe[LC_SLICE];       // its part in the current chunk
    __shared__ Z s_z[LC_SLICE];
    __shared__ uint32_t s_p[LC_SLICE];
    __shared__ uint8_t s_st[LC_SLICE];        // bit 0: result exists, 1: LDS entry
    __shared__ int64_t s_slice;
    __shared__ int s_ge;
    const int lane = threadIdx.x & (G - 1), wv = threadIdx.x / G;
    const int64_t nslices = (count + LC_SLICE - 1) / LC_SLICE;
    const int x0 = (int) (blockIdx.x & 7);
    for (int k = 0; k < 8; k++)
    {
        const int r = (x0 + k) & 7;
        const int64_t slo = nslices * r / 8, shi = nslices * (r + 1) / 8;
        while (true)
        {
            if (threadIdx.x == 0) s_slice = slo + (int64_t) atomicAdd(&ctr[r], 1ull);
            __syncthreads();
            const int64_t sl = s_slice;
            __syncthreads();
            if (sl >= shi) break;
            const int64_t q0 = sl * LC_SLICE;
            const int n = (int) ((count - q0) < LC_SLICE ? (count - q0) : LC_SLICE);
            for (int t = threadIdx.x; t < LC_SLICE; t += LC_NT) s_key[t] = t < n ? keys[q0 + t] : ~0ull;
            __syncthreads();
            int gs = 0;
            while (gs < n)
            {
                // the run of entries with the slice's key at gs (keys are sorted)
                if (threadIdx.x == 0) s_ge = n;
                __syncthreads();
                const uint64_t key = s_key[gs];
                for (int t = gs + 1 + threadIdx.x; t < n; t += LC_NT)
                    if (s_key[t] != key) atomicMin(&s_ge, t);
                __syncthreads();
                const int ge = s_ge;
                const bool rowA = key < (uint64_t) g.key_nA;
                const int64_t rr = rowA ? (int64_t) key : (int64_t) key - g.key_nA;
                const RowDesc d = rowA ? g.dA[rr] : g.dB[rr];
                const bool use_lds = (ge - gs >= LC_GMIN) && d.len > 0;
                for (int t = gs + threadIdx.x; t < ge; t += LC_NT)
                {
                    const uint64_t w = list[q0 + t];
                    const int64_t i = (int64_t) (w >> 32), p = (int64_t) (w & 0xffffffffull);
                    uint8_t st = 0;
                    if (use_lds)
                    {
                        Pair P;
                        (void) d3_pair(g, i, g.Mj[p], P);
                        if (!P.empty && P.probe == (rowA ? 1 : 2))
                        {
                            st = 2;
                            s_srow[t] = rowA ? g.Mj[p] : i;
                            s_rp0[t] = rowA ? P.pb : P.pa;
                            s_z[t] = S::identity();
                        }
                    }
                    s_p[t] = (uint32_t) p;
                    s_st[t] = st;
                }
                __syncthreads();
                // entries on the global path: one wavefront each
                for (int t = gs + wv; t < ge; t += NW)
                {
                    if (s_st[t] != 0) continue;
                    const uint64_t w = list[q0 + t];
                    const int64_t i = (int64_t) (w >> 32), p = (int64_t) (w & 0xffffffffull);
                    Pair P;
                    (void) d3_pair(g, i, g.Mj[p], P);
                    bool e;
                    Z z;
                    dot_lane<S, G>(g, P, lane, e, z);
                    team_reduce<S, G>(e, z);
                    if (lane == 0) d3_store<S>(g, p, e, z);
                }
                if (use_lds)
                {
                    // the long row's bitmap one aligned chunk of LC_BITS
                    // columns at a time; every row's position at each chunk
                    // boundary was found beforehand (pos arrays), so an
                    // entry's visit reads exactly its short row's elements
                    // inside the chunk and the long row's part is known
                    const int64_t *Sj = rowA ? g.Bj : g.Aj;
                    const uint32_t *Sj32 = rowA ? B32 : A32;   // N32: the same columns, 32-bit
                    const int64_t *Lj = rowA ? g.Aj : g.Bj;
                    const uint32_t *posS = rowA ? posB : posA;
                    const uint32_t *posL = rowA ? posA : posB;
                    const int64_t c_lo = (int64_t) d.lo >> LC_LG, c_hi = (int64_t) d.hi >> LC_LG;
                    for (int64_t c = c_lo; c <= c_hi; c++)
                    {
                        const int64_t cbeg = c << LC_LG;
                        const int64_t clo = (int64_t) d.lo > cbeg ? (int64_t) d.lo : cbeg;
                        const int64_t chi = (int64_t) d.hi < cbeg + LC_BITS - 1 ? (int64_t) d.hi
                                                                             : cbeg + LC_BITS - 1;
                        const int wlo = (int) ((clo - cbeg) >> 6), whi = (int) ((chi - cbeg) >> 6) + 1;
                        __syncthreads();              // the previous chunk's probes are done
                        for (int w = wlo + (int) threadIdx.x; w < whi; w += LC_NT) s_bits[w] = 0;
                        // every entry's short-row part in this chunk, loaded
                        // together while the bitmap is built
                        for (int t = gs + (int) threadIdx.x; t < ge; t += LC_NT)
                            if (s_st[t] & 2)
                            {
                                const uint32_t *ps = posS + s_srow[t] * ncp + c;
                                s_qs[t] = ps[0];
                                s_qe[t] = ps[1];
                            }
                        __syncthreads();
                        const int64_t l0 = d.p0 + posL[rr * ncp + c], l1 = d.p0 + posL[rr * ncp + c + 1];
                        for (int64_t e = l0 + threadIdx.x; e < l1; e += LC_NT)
                        {
                            const int64_t dd = Lj[e] - cbeg;
                            atomicOr((unsigned long long *) &s_bits[dd >> 6], 1ull << (dd & 63));
                        }
                        __syncthreads();
                        for (int t = gs + wv; t < ge; t += NW)
                        {
                            const uint8_t st = s_st[t];
                            if (!(st & 2)) continue;
                            const uint32_t rs = s_qs[t], re = s_qe[t];
                            if (rs >= re) continue;
                            const int64_t rp0 = s_rp0[t];
                            int64_t qs = rp0 + rs;
                            const int64_t qe = rp0 + re;
                            // the long row's first chunk: skip the short row's
                            // elements below the long row's first column
                            if (c == c_lo && clo > cbeg)
                            {
                                if constexpr (N32) qs = team_lower_bound<G>(Sj32, qs, qe, clo, lane);
                                else qs = team_lower_bound<G>(Sj, qs, qe, clo, lane);
                            }
                            bool e = false;
                            Z z = S::identity();
                            uint32_t cnt = 0;         // plus_pair: matches of this lane
                            // UN windows of G elements per step, their
                            // loads issued together; elements outside the long
                            // row's range inside the chunk are skipped
                            for (int64_t q = qs; q < qe; q += UN * G)
                            {
                                // lanes past the end read the last element
                                // again (same line): the loads carry no branch
                                // and all issue before the first is used
                                using CT = std::conditional_t<N32, uint32_t, int64_t>;
                                CT raw[UN];
                                #pragma unroll
                                for (int u = 0; u < UN; u++)
                                {
                                    const int64_t qq = q + u * G + lane;
                                    const int64_t qc = qq < qe ? qq : qe - 1;
                                    if constexpr (N32) raw[u] = Sj32[qc];
                                    else raw[u] = Sj[qc];
                                }
                                if constexpr (N32)
                                {
                                    // 32-bit columns: one unsigned compare for the
                                    // range, 32-bit words of the bitmap
                                    const uint32_t c0 = (uint32_t) clo, span = (uint32_t) (chi - clo);
                                    const uint32_t boff = (uint32_t) (clo - cbeg);
                                    const uint32_t *bits32 = (const uint32_t *) s_bits;
                                    #pragma unroll
                                    for (int u = 0; u < UN; u++)
                                    {
                                        const uint32_t d = raw[u] - c0;
                                        if (q + u * G + lane >= qe || d > span) continue;
                                        const uint32_t bi = d + boff;
                                        const uint32_t hit = (bits32[bi >> 5] >> (bi & 31)) & 1u;
                                        if constexpr (S::add == GM_ADD_PLUS && S::mult == GM_MULT_PAIR) cnt += hit;
                                        else if (hit)
                                        {
                                            const Z tz = S::mul(Z(1), Z(1), (int64_t) raw[u]);
                                            if (!e) { z = tz; e = true; }
                                            else if constexpr (S::any_min_k) { if (tz < z) z = tz; }
                                            else z = S::upd(z, tz);
                                        }
                                    }
                                    const bool past = q + (UN - 1) * G + lane < qe &&
                                                      raw[UN - 1] > (uint32_t) chi;
                                    if ((__ballot(past) >> tbase) & tmask) break;
                                }
                                else
                                {
                                    int64_t xs[UN];
                                    #pragma unroll
                                    for (int u = 0; u < UN; u++)
                                        xs[u] = (q + u * G + lane < qe) ? (int64_t) raw[u] : -1;
                                    #pragma unroll
                                    for (int u = 0; u < UN; u++)
                                    {
                                        const int64_t x = xs[u];
                                        if (x < clo || x > chi) continue;
                                        const int64_t dd = x - cbeg;
                                        if ((s_bits[dd >> 6] >> (dd & 63)) & 1ull)
                                        {
                                            const Z tz = S::mul(Z(1), Z(1), x);
                                            if (!e) { z = tz; e = true; }
                                            else if constexpr (S::any_min_k) { if (tz < z) z = tz; }
                                            else z = S::upd(z, tz);
                                        }
                                    }
                                    // sorted: once the step's last window passes
                                    // the long row's last column, nothing after
                                    // it matches
                                    if ((__ballot(xs[UN - 1] > chi) >> tbase) & tmask) break;
                                }
                            }
                            if constexpr (S::add == GM_ADD_PLUS && S::mult == GM_MULT_PAIR)
                                if (cnt) { e = true; z = (Z) cnt; }
                            team_reduce<S, G>(e, z);
                            if (lane == 0)
                            {
                                bool ee = (st & 1) != 0;
                                Z zz = s_z[t];
                                combine<S>(ee, zz, e, z);
                                s_z[t] = zz;
                                s_st[t] = (uint8_t) ((st & ~1) | (ee ? 1 : 0));
                            }
                        }
                    }
                    __syncthreads();
                    for (int t = gs + threadIdx.x; t < ge; t += LC_NT)
                        if (s_st[t] & 2) d3_store<S>(g, (int64_t) s_p[t], (s_st[t] & 1) != 0, s_z[t]);
                }
                __syncthreads();
                gs = ge;
            }
        }
    }
}

// Single-chunk form for operands whose rows all span at most L1_BITS
// columns (small graphs, e.g. a batch of them): a run of at least L1_GMIN
// entries of a sorted bin that probe the same long row builds that row's
// whole bitmap in LDS once (plus, for valued semirings, the row position of
// the first entry of every 64-column word), and teams of G lanes probe their
// short rows against it: lane l takes elements l, l+G, ... of the 32-bit
// column copy (eight loads in flight) and the partials combine in lane
// order, as in the global probe.  Other entries run the team kernel's own
// path, without LDS staging: the kernel keeps its LDS small and its
// registers capped at eight wavefronts per SIMD, because it is bound by the
// latency of its entries' loads (4096-graph batch: 16-lane bin 387 -> 291
// ms, wavefront bin 220 -> 164 ms).
constexpr int L1_BITS = 1 << 16;
constexpr int L1_WORDS = L1_BITS / 64;
constexpr int L1_SLICE = 256;
constexpr int L1_GMIN = 4;

template <class S, int G, bool N32>
__global__ __launch_bounds__(D3_NT) __attribute__((amdgpu_waves_per_eu(8))) void k_dot3_lds1(Dot3Args g, const uint64_t *list,
                                                     const uint64_t *keys, int64_t count,
                                                     unsigned long long *ctr, const uint32_t *A32,
                                                     const uint32_t *B32)
{
    using Z = typename S::Z;
    constexpr int TPB = D3_NT / G;
    constexpr bool RANK = !S::pattern;
    __shared__ uint64_t s_bits[L1_WORDS];
    __shared__ uint32_t s_rank[RANK ? L1_WORDS : 1];
    __shared__ uint64_t s_key[L1_SLICE];
    // every slice entry's short row (start, length) and probe direction
    // (0: none, 1: into A's row, 2: into B's row), set up together at the
    // slice's start so that a team's entry starts with its row's loads
    __shared__ int64_t s_sp0[L1_SLICE];
    __shared__ uint32_t s_slen[L1_SLICE];
    __shared__ int8_t s_pdir[L1_SLICE];
    __shared__ int64_t s_slice;
    __shared__ int s_ge;
    const int lane = threadIdx.x & (G - 1), team = threadIdx.x / G;
    const Z *Ax = (const Z *) g.Ax;
    const Z *Bx = (const Z *) g.Bx;
    const int64_t nslices = (count + L1_SLICE - 1) / L1_SLICE;
    const int x0 = (int) (blockIdx.x & 7);
    for (int k = 0; k < 8; k++)
    {
        const int r = (x0 + k) & 7;
        const int64_t slo = nslices * r / 8, shi = nslices * (r + 1) / 8;
        while (true)
        {
            if (threadIdx.x == 0) s_slice = slo + (int64_t) atomicAdd(&ctr[r], 1ull);
            __syncthreads();
            const int64_t sl_id = s_slice;
            __syncthreads();
            if (sl_id >= shi) break;
            const int64_t q0 = sl_id * L1_SLICE;
            const int n = (int) ((count - q0) < L1_SLICE ? (count - q0) : L1_SLICE);
            for (int t = threadIdx.x; t < L1_SLICE; t += D3_NT)
            {
                s_key[t] = t < n ? keys[q0 + t] : ~0ull;
                int8_t dir = 0;
                if (t < n)
                {
                    const uint64_t w = list[q0 + t];
                    const int64_t i = (int64_t) (w >> 32), p = (int64_t) (w & 0xffffffffull);
                    Pair P;
                    (void) d3_pair(g, i, g.Mj[p], P);
                    if (!P.empty && P.probe != 0)
                    {
                        dir = (int8_t) P.probe;
                        const bool shortA = (P.probe == 2);
                        s_sp0[t] = shortA ? P.pa : P.pb;
                        s_slen[t] = (uint32_t) (shortA ? P.pa_end - P.pa : P.pb_end - P.pb);
                    }
                }
                s_pdir[t] = dir;
            }
            __syncthreads();
            int gs = 0;
            while (gs < n)
            {
                if (threadIdx.x == 0) s_ge = n;
                __syncthreads();
                const uint64_t key = s_key[gs];
                for (int t = gs + 1 + threadIdx.x; t < n; t += D3_NT)
                    if (s_key[t] != key) atomicMin(&s_ge, t);
                __syncthreads();
                const int ge = s_ge;
                const bool rowA = key < (uint64_t) g.key_nA;
                const int64_t rr = rowA ? (int64_t) key : (int64_t) key - g.key_nA;
                const RowDesc d = rowA ? g.dA[rr] : g.dB[rr];
                const int64_t nwords = (((int64_t) d.hi - (int64_t) d.lo) >> 6) + 1;
                const bool use_lds = (ge - gs >= L1_GMIN) && d.len > 0 && nwords <= L1_WORDS;
                if (use_lds)
                {
                    const int64_t *Lj = rowA ? g.Aj : g.Bj;
                    for (int t = threadIdx.x; t < nwords; t += D3_NT) s_bits[t] = 0;
                    __syncthreads();
                    for (int64_t e = d.p0 + threadIdx.x; e < d.p0 + d.len; e += D3_NT)
                    {
                        const int64_t dd = Lj[e] - d.lo;
                        const int64_t wd = dd >> 6;
                        atomicOr((unsigned long long *) &s_bits[wd], 1ull << (dd & 63));
                        if constexpr (RANK)
                        {
                            // the first entry of its word records its position
                            if (e == d.p0 || ((Lj[e - 1] - d.lo) >> 6) != wd)
                                s_rank[wd] = (uint32_t) (e - d.p0);
                        }
                    }
                    __syncthreads();
                }
                for (int t = gs + team; t < ge; t += TPB)
                {
                    if (!(use_lds && s_pdir[t] == (rowA ? 1 : 2)))
                    {
                        team_task<S, G>(g, list[q0 + t], lane, nullptr);
                        continue;
                    }
                    const int64_t p = (int64_t) (list[q0 + t] & 0xffffffffull);
                    // the short row against the LDS bitmap of the long one
                    const bool shortA = !rowA;
                    const int64_t *Sj = shortA ? g.Aj : g.Bj;
                    const int64_t s0 = s_sp0[t], s1 = s0 + (int64_t) s_slen[t];
                    bool e = false;
                    Z z = S::identity();
                    uint32_t cnt = 0;                     // plus_pair: this lane's matches
                    // steps of UN elements per lane, loads issued together (past
                    // the end: the last element again, no branch around the
                    // loads); the lane stops after a step past the long row's range
                    constexpr int UN = N32 ? 8 : 4;
                    const uint32_t *Sj32 = shortA ? A32 : B32;
                    for (int64_t q0s = s0 + lane; q0s < s1; q0s += UN * G)
                    {
                        int64_t xs[UN];
                        #pragma unroll
                        for (int u = 0; u < UN; u++)
                        {
                            const int64_t qq = q0s + u * G;
                            const int64_t qc = qq < s1 ? qq : s1 - 1;
                            if constexpr (N32) xs[u] = (int64_t) Sj32[qc];
                            else xs[u] = Sj[qc];
                        }
                        #pragma unroll
                        for (int u = 0; u < UN; u++)
                            if (q0s + u * G >= s1) xs[u] = INT64_MAX;
                        #pragma unroll
                        for (int u = 0; u < UN; u++)
                        {
                            const int64_t x = xs[u];
                            if (x < (int64_t) d.lo || x > (int64_t) d.hi) continue;
                            const uint32_t dd = (uint32_t) (x - d.lo);
                            if constexpr (S::pattern)
                            {
                                const uint32_t hit = (((const uint32_t *) s_bits)[dd >> 5] >> (dd & 31)) & 1u;
                                if constexpr (S::add == GM_ADD_PLUS && S::mult == GM_MULT_PAIR)
                                {
                                    cnt += hit;
                                }
                                else if (hit)
                                {
                                    const Z tz = S::mul(Z(1), Z(1), x);
                                    if (!e) { z = tz; e = true; }
                                    else if constexpr (S::any_min_k) { if (tz < z) z = tz; }
                                    else z = S::upd(z, tz);
                                }
                            }
                            else
                            {
                                const uint64_t word = s_bits[dd >> 6];
                                const uint64_t bit = 1ull << (dd & 63);
                                if (!(word & bit)) continue;
                                const int64_t q = q0s + u * G;
                                const int64_t pos = d.p0 + (int64_t) s_rank[dd >> 6] + __popcll(word & (bit - 1));
                                const int64_t qa = shortA ? q : pos, qb = shortA ? pos : q;
                                const Z tz = S::mul(S::geta(Ax, qa, g.a_iso), S::geta(Bx, qb, g.b_iso), x);
                                if (!e) { z = tz; e = true; }
                                else z = S::upd(z, tz);
                            }
                        }
                        if (xs[UN - 1] > (int64_t) d.hi) break;
                    }
                    if constexpr (S::add == GM_ADD_PLUS && S::mult == GM_MULT_PAIR)
                        if (cnt) { e = true; z = (Z) cnt; }
                    team_reduce<S, G>(e, z);
                    if (lane == 0) d3_store<S>(g, p, e, z);
                }
                __syncthreads();
                gs = ge;
            }
        }
    }
}

//------------------------------------------------------------------------------
// 16-lane bin against LDS hash tables of the long rows (pattern semirings)
//------------------------------------------------------------------------------

// The 16-lane bin sorted by probed long row, for operands whose rows span
// more than L1_BITS columns (so a whole-row LDS bitmap does not fit).  The
// bin is cut into slices of LH_SLICE entries.  In a slice, the runs of
// entries with one key are found by a workgroup scan; a run whose long row
// has a probe index, at most LH_MAX entries, and enough entries to pay for
// a table (LH_GMIN, and 32 entries per LH_PAY entries of the long row) loads
// that row's columns into an LDS hash table once (2^ceil(log2 2l) slots of
// 32-bit columns, linear probing), and the teams probe their short rows
// there: LDS reads in place of one random L2 / HBM access per element.  The
// slice's other entries take the global path of the team kernel first (the
// table's LDS doubles as their staging area then).  Slices are handed out
// per XCD region as in the grouped team kernels.  Seven workgroups share a
// CU (LDS allows seven); registers are capped to match (72 VGPRs, a few
// spilled: 15.1 against 16.4 ms at five waves per SIMD, RMAT 22).
constexpr int LH_SLICE = 256;
constexpr int LH_MAX = 2048;                   // long row entries at most
constexpr int LH_SLOTS = 2 * LH_MAX;           // 16 KB of LDS
constexpr int LH_GMIN = 4;
constexpr int LH_PAY = 32;
constexpr uint32_t LH_EMPTY = 0xffffffffu;
static_assert(LH_SLOTS * sizeof(uint32_t) >= (D3_NT / 16) * StageCap<16>::value * sizeof(uint32_t),
              "the table's LDS holds the global path's staging lists");

// One lane's share of a short row [s0, s1) probed against an LDS hash table
// of the long row's 32-bit columns (2^lg slots, linear probing): UN elements
// per lane per step, their loads issued together (past the end: the last
// element again); the lane stops after a step past the long row's last column.
template <class S, int G, bool N32, int UN>
__device__ __forceinline__ void lh_probe(const int64_t *Sj, const uint32_t *Sj32, int64_t s0, int64_t s1,
                                         int lane, int64_t hi, const uint32_t *tab, int lg,
                                         uint32_t hmask, bool &ex, typename S::Z &z, uint32_t &cnt)
{
    using Z = typename S::Z;
    for (int64_t qs = s0 + lane; qs < s1; qs += UN * G)
    {
        int64_t xs[UN];
        #pragma unroll
        for (int u = 0; u < UN; u++)
        {
            const int64_t qq = qs + u * G;
            const int64_t qc = qq < s1 ? qq : s1 - 1;
            if constexpr (N32) xs[u] = (int64_t) Sj32[qc];
            else xs[u] = Sj[qc];
        }
        #pragma unroll
        for (int u = 0; u < UN; u++)
            if (qs + u * G >= s1) xs[u] = INT64_MAX;
        #pragma unroll
        for (int u = 0; u < UN; u++)
        {
            const int64_t x = xs[u];
            if (x > hi) break;
            const uint32_t x32 = (uint32_t) x;
            uint32_t h = hx_hash(x32, lg);
            bool hit = false;
            while (true)
            {
                const uint32_t v = tab[h];
                if (v == x32) { hit = true; break; }
                if (v == LH_EMPTY) break;
                h = (h + 1) & hmask;
            }
            if constexpr (S::add == GM_ADD_PLUS && S::mult == GM_MULT_PAIR)
            {
                cnt += hit ? 1u : 0u;
            }
            else
            {
                if (!hit) continue;
                const Z tz = S::mul(Z(1), Z(1), x);
                if (!ex) { z = tz; ex = true; }
                else if constexpr (S::any_min_k) { if (tz < z) z = tz; }
                else z = S::upd(z, tz);
            }
        }
        if (xs[UN - 1] > hi) break;
    }
}

template <class S, int G, bool N32>
__global__ __launch_bounds__(D3_NT) __attribute__((amdgpu_waves_per_eu(7))) void k_dot3_team_lh(Dot3Args g, const uint64_t *list,
                                                        const uint64_t *keys, int64_t count,
                                                        unsigned long long *ctr, const uint32_t *A32,
                                                        const uint32_t *B32)
{
    using Z = typename S::Z;
    static_assert(S::pattern, "LDS hash probing covers the pattern semirings");
    constexpr int TPB = D3_NT / G;
    static_assert(LH_SLICE == D3_NT, "one slice entry per thread");
    __shared__ uint32_t s_tab[LH_SLOTS];        // hash table, or the teams' staging lists
    __shared__ uint64_t s_key[LH_SLICE];
    __shared__ int s_rs[LH_SLICE + 1];          // start of run r in the slice
    __shared__ uint8_t s_re[LH_SLICE];          // run r takes the LDS table
    __shared__ int s_rid[LH_SLICE];             // run of entry t
    __shared__ int s_scan[D3_NT / 64 + 1];
    __shared__ int64_t s_slice;
    const int lane = threadIdx.x & (G - 1), team = threadIdx.x / G;
    const int64_t nslices = (count + LH_SLICE - 1) / LH_SLICE;
    const int x0 = (int) (blockIdx.x & 7);
    for (int k = 0; k < 8; k++)
    {
        const int r = (x0 + k) & 7;
        const int64_t slo = nslices * r / 8, shi = nslices * (r + 1) / 8;
        while (true)
        {
            if (threadIdx.x == 0) s_slice = slo + (int64_t) atomicAdd(&ctr[r], 1ull);
            __syncthreads();
            const int64_t sl_id = s_slice;
            __syncthreads();
            if (sl_id >= shi) break;
            const int64_t q0 = sl_id * LH_SLICE;
            const int n = (int) ((count - q0) < LH_SLICE ? (count - q0) : LH_SLICE);
            // runs of equal keys: flag the first entry of each, scan for run ids
            const int t = threadIdx.x;
            const uint64_t kt = t < n ? keys[q0 + t] : ~0ull;
            s_key[t] = kt;
            __syncthreads();
            const int first = (t < n && (t == 0 || s_key[t - 1] != kt)) ? 1 : 0;
            int nruns;
            const int rid = block_exclusive_scan<int, D3_NT>(first, s_scan, &nruns) + first - 1;
            if (first)
            {
                s_rs[rid] = t;
                s_rid[t] = rid;
            }
            if (t == 0) s_rs[nruns] = n;
            __syncthreads();
            if (t < n && !first) s_rid[t] = rid;
            // run eligibility, decided by its first entry
            if (first)
            {
                const int rlen = s_rs[rid + 1] - t;
                const bool rowA = kt < (uint64_t) g.key_nA;
                const int64_t rr = rowA ? (int64_t) kt : (int64_t) kt - g.key_nA;
                const RowDesc d = rowA ? g.dA[rr] : g.dB[rr];
                s_re[rid] = (rlen >= LH_GMIN && d.kind >= 0 && d.len > 0 && d.len <= LH_MAX &&
                             (int64_t) rlen * LH_PAY >= (int64_t) d.len) ? 1 : 0;
            }
            __syncthreads();
            // entries of the other runs: the team kernel's own path
            for (int e = team; e < n; e += TPB)
                if (!s_re[s_rid[e]]) team_task<S, G>(g, list[q0 + e], lane, s_tab + team * StageCap<G>::value);
            for (int rr_ = 0; rr_ < nruns; rr_++)
            {
                if (!s_re[rr_]) continue;
                const int gs = s_rs[rr_], ge = s_rs[rr_ + 1];
                const uint64_t key = s_key[gs];
                const bool rowA = key < (uint64_t) g.key_nA;
                const int64_t rr = rowA ? (int64_t) key : (int64_t) key - g.key_nA;
                const RowDesc d = rowA ? g.dA[rr] : g.dB[rr];
                const int lg = ilog2_ceil(2 * (int64_t) d.len);
                const uint32_t hmask = (1u << lg) - 1u;
                __syncthreads();                  // previous users of s_tab are done
                for (int w = threadIdx.x; w <= (int) hmask; w += D3_NT) s_tab[w] = LH_EMPTY;
                __syncthreads();
                {
                    const int64_t *Lj = rowA ? g.Aj : g.Bj;
                    for (int e = threadIdx.x; e < d.len; e += D3_NT)
                    {
                        const uint32_t x = (uint32_t) Lj[d.p0 + e];
                        uint32_t h = hx_hash(x, lg);
                        while (atomicCAS(&s_tab[h], LH_EMPTY, x) != LH_EMPTY) h = (h + 1) & hmask;
                    }
                }
                __syncthreads();
                for (int e = gs + team; e < ge; e += TPB)
                {
                    const uint64_t w = list[q0 + e];
                    const int64_t i = (int64_t) (w >> 32), p = (int64_t) (w & 0xffffffffull);
                    Pair P;
                    (void) d3_pair(g, i, g.Mj[p], P);
                    bool ex = false;
                    Z z = S::identity();
                    if (!P.empty && P.probe == (rowA ? 1 : 2))
                    {
                        // the short row from its first element in the long
                        // row's range, probed in the LDS table (lh_probe)
                        int64_t s0, s1;
                        d3_trim_short_team<G>(g, P, lane, s0, s1);
                        const int64_t *Sj = rowA ? g.Bj : g.Aj;
                        const uint32_t *Sj32 = rowA ? B32 : A32;   // N32: the same columns, 32-bit
                        uint32_t cnt = 0;                          // plus_pair: this lane's matches
                        lh_probe<S, G, N32, N32 ? 8 : 4>(Sj, Sj32, s0, s1, lane, d.hi, s_tab, lg, hmask, ex, z, cnt);
                        if constexpr (S::add == GM_ADD_PLUS && S::mult == GM_MULT_PAIR)
                            if (cnt) { ex = true; z = (Z) cnt; }
                    }
                    else if (!P.empty)
                    {
                        // an entry of this run that probes the other way
                        // (its short row is the run's row): global path
                        dot_lane<S, G>(g, P, lane, ex, z);
                    }
                    team_reduce<S, G>(ex, z);
                    if (lane == 0) d3_store<S>(g, p, ex, z);
                }
            }
            __syncthreads();
        }
    }
}

//------------------------------------------------------------------------------
// row mode: whole mask rows against an LDS bitmap of A(i,:)
//------------------------------------------------------------------------------

// A mask row i with enough entries, whose A(i,:) spans fewer than RM_BITS
// columns, is done by one workgroup: A(i,:) becomes a bitmap over its column
// range in LDS (plus a rank per 32-bit word for valued semirings), and every
// entry (i,j) streams B(j,:), trimmed to that range, through it with
// coalesced loads -- LDS lookups instead of random probes into a global
// index.  An entry whose B(j,:) is long and indexed while A(i,:) is short
// probes A(i,:) into B(j,:)'s index instead, as the binned kernels do.
constexpr int RM_BITS = 1 << 17;              // bitmap span: 16 KB of LDS
constexpr int RM_WORDS = RM_BITS / 32;
constexpr int RM_TEAM = 16;                   // lanes per mask entry

__global__ __launch_bounds__(D3_NT) void k_d3_rowmode(const int64_t *__restrict__ Mp,
    const int64_t *__restrict__ Ap, const int64_t *__restrict__ Aj, int64_t n, int64_t min_mrow,
    int64_t min_arow, uint8_t *__restrict__ flag, uint32_t *__restrict__ rows,
    unsigned long long *__restrict__ cnt)
{
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    const int64_t n_pad = ((n + 63) / 64) * 64;          // whole wavefronts for the appends
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n_pad; i += stride)
    {
        bool ok = false;
        if (i < n)
        {
            const int64_t a0 = Ap[i], a1 = Ap[i + 1];
            ok = (Mp[i + 1] - Mp[i]) >= min_mrow && (a1 - a0) >= min_arow && a1 > a0 &&
                 (Aj[a1 - 1] - Aj[a0]) < RM_BITS;
            flag[i] = ok ? 1 : 0;
        }
        const unsigned long long slot = wave_append(ok, cnt);
        if (ok) rows[slot] = (uint32_t) i;
    }
}

template <class S>
__global__ __launch_bounds__(D3_NT) void k_dot3_rows(Dot3Args g, const uint32_t *rows,
                                                     const unsigned long long *nrows_ptr)
{
    using Z = typename S::Z;
    __shared__ uint32_t s_bm[RM_WORDS];
    __shared__ uint32_t s_rank[S::pattern ? 1 : RM_WORDS];
    __shared__ int64_t s_scan[D3_NT / 64 + 1];
    const int64_t nrows = (int64_t) *nrows_ptr;
    const int team = threadIdx.x / RM_TEAM, lane = threadIdx.x & (RM_TEAM - 1);
    constexpr int NTEAM = D3_NT / RM_TEAM;
    const Z *Ax = (const Z *) g.Ax, *Bx = (const Z *) g.Bx;
    for (int64_t r = blockIdx.x; r < nrows; r += gridDim.x)
    {
        const int64_t i = rows[r];
        const int64_t a0 = g.Ap[i], a1 = g.Ap[i + 1];
        const int64_t alo = g.Aj[a0], ahi = g.Aj[a1 - 1];
        const int nw = (int) ((ahi - alo) >> 5) + 1;
        for (int w = threadIdx.x; w < nw; w += D3_NT) s_bm[w] = 0;
        __syncthreads();
        for (int64_t q = a0 + threadIdx.x; q < a1; q += D3_NT)
        {
            const int64_t d = g.Aj[q] - alo;
            atomicOr(&s_bm[d >> 5], 1u << (d & 31));
        }
        __syncthreads();
        if constexpr (!S::pattern)
        {
            // rank[w] = entries of A(i,:) in words before w
            const int chunk = (nw + D3_NT - 1) / D3_NT;
            const int w0 = threadIdx.x * chunk, w1 = (w0 + chunk < nw) ? w0 + chunk : nw;
            int64_t c = 0;
            for (int w = w0; w < w1; w++) c += __popc(s_bm[w]);
            int64_t tot;
            int64_t pre = block_exclusive_scan<int64_t, D3_NT>(c, s_scan, &tot);
            for (int w = w0; w < w1; w++)
            {
                s_rank[w] = (uint32_t) pre;
                pre += __popc(s_bm[w]);
            }
            __syncthreads();
        }
        const int64_t m0 = g.Mp[i], m1 = g.Mp[i + 1];
        // every team runs the same number of rounds (team-wide shuffles inside)
        for (int64_t pb0 = m0; pb0 < m1; pb0 += NTEAM)
        {
            const int64_t p = pb0 + team;
            bool e = false;
            Z z = S::identity();
            if (p < m1)
            {
                const int64_t j = g.Mj[p];
                if (g.mask_struct || mask_true(g.Mx, p, g.msize, g.m_iso))
                {
                    Pair P;
                    (void) d3_pair(g, i, j, P);
                    if (!P.empty && P.probe == 2)
                    {
                        dot_lane<S, RM_TEAM>(g, P, lane, e, z);
                    }
                    else if (!P.empty)
                    {
                        const int64_t q0 = team_lower_bound<RM_TEAM>(g.Bj, P.pb, P.pb_end, alo, lane);
                        const int64_t q1 = team_lower_bound<RM_TEAM>(g.Bj, q0, P.pb_end, ahi + 1, lane);
                        for (int64_t q = q0 + lane; q < q1; q += RM_TEAM)
                        {
                            const int64_t x = g.Bj[q];
                            const int64_t d = x - alo;
                            const uint32_t word = s_bm[d >> 5];
                            const uint32_t bit = 1u << (d & 31);
                            if (!(word & bit)) continue;
                            int64_t qa = a0;
                            if constexpr (!S::pattern) qa = a0 + (int64_t) s_rank[d >> 5] + __popc(word & (bit - 1u));
                            Z t = S::mul(S::geta(Ax, qa, g.a_iso), S::geta(Bx, q, g.b_iso), x);
                            if (!e) { z = t; e = true; }
                            else if constexpr (S::any_min_k) { if (t < z) z = t; }
                            else z = S::upd(z, t);
                            if constexpr (S::is_any && !S::any_min_k) break;
                            if (S::terminal(z) && !S::any_min_k) break;
                        }
                    }
                }
            }
            team_reduce<S, RM_TEAM>(e, z);
            if (lane == 0 && p < m1) d3_store<S>(g, p, e, z);
        }
        __syncthreads();
    }
}

//------------------------------------------------------------------------------
// compaction: drop zombies, build the new row pointers
//------------------------------------------------------------------------------

__global__ __launch_bounds__(D3_NT) void k_dot3_tile_count(const uint8_t *__restrict__ live,
                                                           int64_t mnz, int64_t *__restrict__ cnt)
{
    __shared__ int64_t lds[D3_NT / 64 + 1];
    const int64_t base = (int64_t) blockIdx.x * D3_CT_TILE + (int64_t) threadIdx.x * D3_CT_IPT;
    int64_t c = 0;
    #pragma unroll
    for (int t = 0; t < D3_CT_IPT; t++)
    {
        const int64_t p = base + t;
        if (p < mnz) c += live[p];
    }
    const int64_t tot = block_sum<int64_t, D3_NT>(c, lds);
    if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

template <class Z, bool ISO>
__global__ __launch_bounds__(D3_NT) void k_dot3_compact(const int64_t *__restrict__ Mp,
    const int64_t *__restrict__ Mj, int64_t m_nrows, int64_t mnz,
    const uint8_t *__restrict__ live, const Z *__restrict__ W,
    const int64_t *__restrict__ tile_off, int64_t *__restrict__ Cp,
    int64_t *__restrict__ Ci, Z *__restrict__ Cx)
{
    constexpr int NW = D3_NT / 64;
    static_assert(D3_CT_IPT * NW <= 64, "one wavefront scans the per-step counts");
    __shared__ int32_t s_pre[D3_CT_TILE + 1];
    __shared__ int32_t s_cnt[D3_CT_IPT * NW];
    __shared__ int64_t s_row0;
    const int64_t t0 = (int64_t) blockIdx.x * D3_CT_TILE;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0)
    {
        // first row whose start lies at or after t0
        int64_t lo = 0, hi = m_nrows + 1;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi) >> 1;
            if (Mp[mid] < t0) lo = mid + 1; else hi = mid;
        }
        s_row0 = lo;
    }
    // step t covers entries t0 + t*D3_NT + [0, D3_NT): a wavefront's lanes
    // read and write consecutive entries; the live ones of a step are placed
    // by a ballot, the steps and wavefronts in entry order by one scan
    unsigned long long bal[D3_CT_IPT];
    #pragma unroll
    for (int t = 0; t < D3_CT_IPT; t++)
    {
        const int64_t p = t0 + (int64_t) t * D3_NT + threadIdx.x;
        bal[t] = __ballot((p < mnz) && live[p] != 0);
        if (lane == 0) s_cnt[t * NW + wid] = __popcll(bal[t]);
    }
    __syncthreads();
    if (wid == 0)
    {
        const int v = lane < D3_CT_IPT * NW ? s_cnt[lane] : 0;
        int incl = v;
        #pragma unroll
        for (int d = 1; d < 64; d <<= 1)
        {
            const int o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane < D3_CT_IPT * NW) s_cnt[lane] = incl - v;
        if (lane == 63) s_pre[D3_CT_TILE] = incl;
    }
    __syncthreads();
    const int64_t off = tile_off[blockIdx.x];
    const unsigned long long below = (1ull << lane) - 1ull;
    #pragma unroll
    for (int t = 0; t < D3_CT_IPT; t++)
    {
        const int64_t p = t0 + (int64_t) t * D3_NT + threadIdx.x;
        const int pos = s_cnt[t * NW + wid] + __popcll(bal[t] & below);
        s_pre[t * D3_NT + threadIdx.x] = pos;
        if ((bal[t] >> lane) & 1ull)
        {
            Ci[off + pos] = Mj[p];
            if constexpr (!ISO) Cx[off + pos] = W[p];
        }
    }
    __syncthreads();
    // rows starting inside this tile; the last tile also takes rows that
    // start at mnz (trailing empty rows and the final pointer)
    const bool last = (t0 + D3_CT_TILE >= mnz);
    const int64_t t_end = last ? mnz + 1 : t0 + D3_CT_TILE;
    for (int64_t r = s_row0 + threadIdx.x; r <= m_nrows; r += D3_NT)
    {
        const int64_t mp = Mp[r];
        if (mp >= t_end) break;
        Cp[r] = off + s_pre[mp - t0];
    }
}

__global__ void k_fill_i64(int64_t *x, int64_t n, int64_t v)
{
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t) gridDim.x * blockDim.x)
        x[t] = v;
}

__global__ void k_set_iso_one(void *x, int size)
{
    // iso value of ANY_PAIR: true / 1
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        uint8_t *b = (uint8_t *) x;
        for (int t = 0; t < size; t++) b[t] = 0;
        b[0] = 1;
    }
}

__global__ void k_zero_u64(unsigned long long *x, int n)
{
    if (threadIdx.x < n) x[threadIdx.x] = 0;
}

//------------------------------------------------------------------------------
// bitmaps of the long rows
//------------------------------------------------------------------------------

constexpr int BM_HIST = 64;

// words needed by eligible rows, bucketed by floor(log2(row length))
__global__ __launch_bounds__(D3_NT) void k_d3_bm_hist(const int64_t *__restrict__ Xp,
                                                      const int64_t *__restrict__ Xj, int64_t nrows,
                                                      int64_t minlen, unsigned long long *hist)
{
    // per-workgroup histogram in LDS, then one atomic per nonzero bucket
    __shared__ unsigned long long s_h[64];
    if (threadIdx.x < 64) s_h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t r = (int64_t) blockIdx.x * D3_NT + threadIdx.x; r < nrows;
         r += (int64_t) gridDim.x * D3_NT)
    {
        const int64_t p0 = Xp[r], len = Xp[r + 1] - p0;
        if (len < minlen) continue;
        const int64_t lo = Xj[p0], hi = Xj[p0 + len - 1];
        const int b = 63 - __clzll((unsigned long long) len);
        atomicAdd(&s_h[b], (unsigned long long) px_units(len, lo, hi));
    }
    __syncthreads();
    if (threadIdx.x < 64 && s_h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], s_h[threadIdx.x]);
}

// word offsets of the eligible rows (atomic bump) and the list of them
__global__ __launch_bounds__(D3_NT) void k_d3_bm_assign(const int64_t *__restrict__ Xp,
                                                        const int64_t *__restrict__ Xj, int64_t nrows,
                                                        int64_t T, int64_t *off,
                                                        unsigned long long *wctr,
                                                        unsigned long long *lctr, uint32_t *list)
{
    // one reservation per wavefront for its rows' units and list slots;
    // every lane of a wavefront runs the same iterations
    const int lane = threadIdx.x & 63;
    for (int64_t r0 = (int64_t) blockIdx.x * D3_NT; r0 < nrows; r0 += (int64_t) gridDim.x * D3_NT)
    {
        const int64_t r = r0 + threadIdx.x;
        bool take = false;
        unsigned long long u = 0;
        if (r < nrows)
        {
            const int64_t p0 = Xp[r], len = Xp[r + 1] - p0;
            if (len >= T)
            {
                take = true;
                u = (unsigned long long) px_units(len, Xj[p0], Xj[p0 + len - 1]);
            }
        }
        unsigned long long inc = u;
        #pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const unsigned long long t = __shfl_up(inc, o, 64);
            if (lane >= o) inc += t;
        }
        const unsigned long long total = __shfl(inc, 63, 64);
        const unsigned long long mask = __ballot(take);
        unsigned long long wbase = 0, lbase = 0;
        if (lane == 0 && mask != 0)
        {
            wbase = atomicAdd(wctr, total);
            lbase = atomicAdd(lctr, (unsigned long long) __popcll(mask));
        }
        wbase = __shfl(wbase, 0, 64);
        lbase = __shfl(lbase, 0, 64);
        if (take)
        {
            off[r] = (int64_t) (wbase + inc - u);
            list[lbase + (unsigned long long) __popcll(mask & ((1ull << lane) - 1ull))] = (uint32_t) r;
        }
    }
}

// One workgroup per listed row: clear its words, then OR in its elements.
// Elements are sorted, so the lanes holding one word are contiguous: a
// segmented OR across the wavefront leaves one atomic per word per chunk.
// rank[w] (when requested) = number of row entries before word w.
__global__ __launch_bounds__(D3_NT) void k_d3_bm_build(const int64_t *__restrict__ Xp,
                                                       const int64_t *__restrict__ Xj,
                                                       const int64_t *__restrict__ off,
                                                       const uint32_t *__restrict__ list,
                                                       const unsigned long long *lctr,
                                                       unsigned long long *words, uint32_t *rank)
{
    const int lane = threadIdx.x & 63;
    const int64_t count = (int64_t) *lctr;
    for (int64_t q = blockIdx.x; q < count; q += gridDim.x)
    {
        const int64_t r = list[q];
        const int64_t p0 = Xp[r], p1 = Xp[r + 1];
        const int64_t lo = Xj[p0], hi = Xj[p1 - 1];
        const int64_t o = off[r];
        if (!bm_eligible(0, p1 - p0, lo, hi))
        {
            // hash index: (column << 32 | position)
            const int lg = hx_log2(p1 - p0);
            const int64_t ns = (int64_t) 1 << lg;
            const uint32_t mask = (1u << lg) - 1u;
            for (int64_t t = threadIdx.x; t < ns; t += D3_NT) words[o + t] = HX_EMPTY;
            __syncthreads();
            for (int64_t p = p0 + threadIdx.x; p < p1; p += D3_NT)
            {
                const uint64_t key = (uint64_t) Xj[p];
                const unsigned long long ent = (key << 32) | (uint64_t) (p - p0);
                uint32_t h = hx_hash((uint32_t) key, lg);
                while (atomicCAS(&words[o + h], HX_EMPTY, ent) != HX_EMPTY) h = (h + 1) & mask;
            }
            __syncthreads();
            continue;
        }
        const int64_t nw = ((hi - lo) >> 6) + 1;
        for (int64_t w = threadIdx.x; w < nw; w += D3_NT) words[o + w] = 0;
        __syncthreads();
        for (int64_t base = p0; base < p1; base += D3_NT)
        {
            const int64_t p = base + threadIdx.x;
            const bool valid = p < p1;
            int64_t w = -1;
            unsigned long long bits = 0;
            int64_t wprev = -1;
            if (valid)
            {
                const int64_t d = Xj[p] - lo;
                w = d >> 6;
                bits = 1ull << (d & 63);
                if (p > p0) wprev = (Xj[p - 1] - lo) >> 6;
            }
            #pragma unroll
            for (int s = 1; s < 64; s <<= 1)
            {
                const unsigned long long ob = __shfl_down(bits, s, 64);
                const int64_t ow = __shfl_down(w, s, 64);
                if (lane + s < 64 && ow == w) bits |= ob;
            }
            const bool head = valid && (w != wprev);
            if (valid && (head || lane == 0)) atomicOr(&words[o + w], bits);
            if (head && rank)
            {
                const uint32_t pos = (uint32_t) (p - p0);
                for (int64_t gw = wprev + 1; gw <= w; gw++) rank[o + gw] = pos;
            }
        }
        __syncthreads();
    }
}

// 32-bit copy of a column index array (columns below 2^32)
__global__ __launch_bounds__(D3_NT) void k_d3_narrow(const int64_t *__restrict__ X, int64_t n,
                                                     uint32_t *__restrict__ out)
{
    for (int64_t t = (int64_t) blockIdx.x * D3_NT + threadIdx.x; t < n; t += (int64_t) gridDim.x * D3_NT)
        out[t] = (uint32_t) X[t];
}

// Each row's offsets at the LDS chunk boundaries: pos[r*ncp + c] is the
// number of the row's entries whose column is below c*LC_BITS (so entry 0
// is 0 and entry ncp-1 is the row length).  One thread per (row, boundary),
// a binary search inside the row.
__global__ __launch_bounds__(D3_NT) void k_d3_chunkpos(const int64_t *__restrict__ Xp,
                                                       const int64_t *__restrict__ Xj, int64_t nrows,
                                                       int ncp, uint32_t *__restrict__ pos)
{
    const int64_t total = nrows * ncp;
    for (int64_t t = (int64_t) blockIdx.x * D3_NT + threadIdx.x; t < total;
         t += (int64_t) gridDim.x * D3_NT)
    {
        const int64_t r = t / ncp;
        const int64_t c = t - r * ncp;
        const int64_t p0 = Xp[r], p1 = Xp[r + 1];
        const int64_t key = c << LC_LG;
        int64_t lo = p0, hi = p1;
        if (c == 0) hi = p0;
        else if (c == ncp - 1) lo = p1;
        while (lo < hi)
        {
            const int64_t mid = lo + ((hi - lo) >> 1);
            if (Xj[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        pos[t] = (uint32_t) (lo - p0);
    }
}

// one summary per row (off: the probe index offsets, or NULL when the
// operand has none; rows of length >= T have an index)
__global__ __launch_bounds__(D3_NT) void k_d3_desc(const int64_t *__restrict__ Xp,
                                                   const int64_t *__restrict__ Xj, int64_t nrows,
                                                   int64_t T, const int64_t *__restrict__ off,
                                                   RowDesc *__restrict__ out,
                                                   unsigned long long *maxspan)
{
    unsigned long long ms = 0;
    // every lane of a wavefront runs the same number of iterations
    for (int64_t r0 = (int64_t) blockIdx.x * D3_NT; r0 < nrows; r0 += (int64_t) gridDim.x * D3_NT)
    {
        const int64_t r = r0 + threadIdx.x;
        if (r >= nrows) continue;
        const int64_t p0 = Xp[r], len = Xp[r + 1] - p0;
        RowDesc d;
        d.p0 = p0;
        d.off = 0;
        d.len = (int32_t) len;
        d.lo = len > 0 ? (int32_t) Xj[p0] : 0;
        d.hi = len > 0 ? (int32_t) Xj[p0 + len - 1] : -1;
        d.kind = -1;
        if (off && len >= T)
        {
            d.off = off[r];
            d.kind = bm_eligible(0, len, d.lo, d.hi) ? 0 : hx_log2(len);
        }
        out[r] = d;
        const unsigned long long sp = len > 0 ? (unsigned long long) (d.hi - d.lo + 1) : 0ull;
        ms = sp > ms ? sp : ms;
    }
    // largest column span of a row: wavefront maximum, then an atomic only
    // where it raises the value seen so far
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
        const unsigned long long t = __shfl_xor(ms, o, 64);
        ms = t > ms ? t : ms;
    }
    if ((threadIdx.x & 63) == 0 &&
        ms > __hip_atomic_load(maxspan, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(maxspan, ms);
}

static inline unsigned grid_for(int64_t work_items, int64_t per_block, int64_t cap)
{
    int64_t g = (work_items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned) g;
}

// Decide which rows get bitmaps (longest first, within a budget of HBM),
// then build them.  A and B share bitmaps when they are the same matrix.
static int dot3_bitmaps(Dot3Args &g, const gm_csr *A, const gm_csr *B, bool need_rank,
                        int64_t minlen, DevBuf &offA, DevBuf &offB, DevBuf &words, DevBuf &rank,
                        DevBuf &list, DevBuf &ctr)
{
    hipStream_t st = stream();
    const bool shared = (A->p == B->p && A->j == B->j && A->nrows == B->nrows);
    if (!ctr.alloc(sizeof(unsigned long long) * (2 * BM_HIST + 8)))
        return set_error(GM_OUT_OF_MEMORY, "dot3: bitmap counters");
    unsigned long long *c = ctr.as<unsigned long long>();
    k_zero_u64<<<1, 64, 0, st>>>(c, 64);
    k_zero_u64<<<1, 64, 0, st>>>(c + 64, 64);
    k_zero_u64<<<1, 64, 0, st>>>(c + 128, 8);
    const int64_t gcap = (int64_t) cu_count() * 8;
    {
        ProfScope ps("dot3_bm_hist");
        k_d3_bm_hist<<<grid_for(A->nrows, D3_NT, gcap), D3_NT, 0, st>>>(A->p, A->j, A->nrows, minlen, c);
        if (!shared)
            k_d3_bm_hist<<<grid_for(B->nrows, D3_NT, gcap), D3_NT, 0, st>>>(B->p, B->j, B->nrows,
                                                                            minlen, c + BM_HIST);
        GM_HIP(hipGetLastError(), "k_d3_bm_hist");
    }
    unsigned long long *h = (unsigned long long *) pinned_scratch(sizeof(unsigned long long) * 2 * BM_HIST);
    if (!h) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(h, c, sizeof(unsigned long long) * 2 * BM_HIST, hipMemcpyDeviceToHost, st),
           "bitmap histogram");
    GM_HIP(hipStreamSynchronize(st), "bitmap histogram sync");
    const double wbytes = 8.0 + (need_rank ? 4.0 : 0.0);
    double budget = (double) device_hbm_bytes() / 16.0;
    if (budget > 16e9) budget = 16e9;
    // smallest length bucket whose cumulative words fit the budget
    int b0 = -1;
    double acc = 0;
    for (int b = BM_HIST - 1; b >= 0; b--)
    {
        const double wb = (double) (h[b] + h[BM_HIST + b]) * wbytes;
        if (wb == 0) continue;
        if (acc + wb > budget) break;
        acc += wb;
        b0 = b;
    }
    if (b0 < 0) return GM_SUCCESS;               // no row qualifies
    int64_t T = (int64_t) 1 << b0;
    if (T < minlen) T = minlen;
    const int64_t total_words = (int64_t) (acc / wbytes) + 1;
    if (!offA.alloc(sizeof(int64_t) * (A->nrows > 0 ? A->nrows : 1)) ||
        (!shared && !offB.alloc(sizeof(int64_t) * (B->nrows > 0 ? B->nrows : 1))) ||
        !words.alloc(sizeof(uint64_t) * total_words) ||
        (need_rank && !rank.alloc(sizeof(uint32_t) * total_words)) ||
        !list.alloc(sizeof(uint32_t) * ((A->nrows + (shared ? 0 : B->nrows)) > 0
                                        ? (A->nrows + (shared ? 0 : B->nrows)) : 1)))
        return set_error(GM_OUT_OF_MEMORY, "dot3: bitmaps (%lld words)", (long long) total_words);
    uint32_t *lst = list.as<uint32_t>();
    // c[128]: word bump shared by A and B; c[129] / c[130]: rows of A / B listed
    {
        ProfScope ps("dot3_bm_build");
        unsigned long long *dw = words.as<unsigned long long>();
        uint32_t *rk = need_rank ? rank.as<uint32_t>() : nullptr;
        k_d3_bm_assign<<<grid_for(A->nrows, D3_NT, gcap), D3_NT, 0, st>>>(
            A->p, A->j, A->nrows, T, offA.as<int64_t>(), c + 128, c + 129, lst);
        GM_HIP(hipGetLastError(), "k_d3_bm_assign");
        k_d3_bm_build<<<(unsigned) gcap, D3_NT, 0, st>>>(A->p, A->j, offA.as<int64_t>(), lst, c + 129,
                                                         dw, rk);
        GM_HIP(hipGetLastError(), "k_d3_bm_build");
        if (!shared)
        {
            k_d3_bm_assign<<<grid_for(B->nrows, D3_NT, gcap), D3_NT, 0, st>>>(
                B->p, B->j, B->nrows, T, offB.as<int64_t>(), c + 128, c + 130, lst + A->nrows);
            GM_HIP(hipGetLastError(), "k_d3_bm_assign");
            k_d3_bm_build<<<(unsigned) gcap, D3_NT, 0, st>>>(B->p, B->j, offB.as<int64_t>(),
                                                             lst + A->nrows, c + 130, dw, rk);
            GM_HIP(hipGetLastError(), "k_d3_bm_build");
        }
    }
    g.bA = RowBits{T, offA.as<int64_t>(), (const uint64_t *) words.ptr,
                   need_rank ? rank.as<uint32_t>() : nullptr};
    g.bB = shared ? g.bA
                  : RowBits{T, offB.as<int64_t>(), (const uint64_t *) words.ptr,
                            need_rank ? rank.as<uint32_t>() : nullptr};
    return GM_SUCCESS;
}

// Sort the team16, wave and workgroup bins by their grouping keys (radix sort
// of (key, entry) pairs) and run them with the XCD-regioned kernels.  The
// bins' keys must have been written by phase1 (g.bkeys != NULL).
template <class S>
static int dot3_grouped_teams(Dot3Args &g, const gm_csr *A, const gm_csr *B)
{
    const int64_t nkeys = A->nrows + B->nrows;
    hipStream_t st = stream();
    unsigned long long *hc = (unsigned long long *) pinned_scratch(6 * sizeof(unsigned long long));
    if (!hc) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(hc, g.bin_cnt, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
           "bin counts");
    // the operands' entry counts (32-bit column copies for the chunk kernel)
    GM_HIP(hipMemcpyAsync(hc + 4, A->p + A->nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz(A)");
    GM_HIP(hipMemcpyAsync(hc + 5, B->p + B->nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz(B)");
    GM_HIP(hipStreamSynchronize(st), "bin counts sync");
    const int64_t cnt[3] = {(int64_t) hc[0], (int64_t) hc[1], (int64_t) hc[2]};
    // every row spans at most L1_BITS columns: whole bitmaps fit in LDS
    const bool small = g.dA != nullptr && hc[3] > 0 && hc[3] <= (unsigned long long) L1_BITS;
    int64_t big = cnt[0] > cnt[1] ? cnt[0] : cnt[1];
    if (cnt[2] > big) big = cnt[2];
    if (big == 0) return GM_SUCCESS;
    int end_bit = 1;
    while (end_bit < 64 && ((int64_t) 1 << end_bit) <= nkeys) end_bit++;
    DevBuf tk, tv, tmp, ctr;
    if (!tk.alloc(sizeof(uint64_t) * big) || !tv.alloc(sizeof(uint64_t) * big) ||
        !ctr.alloc(40 * sizeof(unsigned long long)))
        return set_error(GM_OUT_OF_MEMORY, "dot3: grouping buffers");
    size_t tbytes = 0;
    for (int b = 0; b < 3; b++)
    {
        size_t tb = 0;
        GM_HIP(rocprim::radix_sort_pairs(nullptr, tb, (const uint64_t *) g.bkeys, tk.as<uint64_t>(),
                                         (const uint64_t *) g.bins, tv.as<uint64_t>(), (size_t) cnt[b],
                                         0, (unsigned) end_bit, st), "grouping sort size");
        if (tb > tbytes) tbytes = tb;
    }
    if (!tmp.alloc(tbytes)) return set_error(GM_OUT_OF_MEMORY, "dot3: grouping sort");
    k_zero_u64<<<1, 64, 0, st>>>(ctr.as<unsigned long long>(), 40);
    // the wavefront and workgroup bins probe bitmaps from LDS when the
    // semiring needs no positions and the row summaries exist
    const bool lds = S::pattern && g.dA != nullptr && g.bA.words != nullptr;
    // the chunk kernel reads every row's offsets at the chunk boundaries;
    // without room for them (very wide matrices) those bins take the
    // global-memory kernels
    DevBuf posA, posB;
    const int ncp = (int) ((A->ncols + LC_BITS - 1) >> LC_LG) + 1;
    const bool shared = (A->p == B->p && A->j == B->j && A->nrows == B->nrows);
    bool lds_chunk = false;
    if (lds && (cnt[2] > 0 || (cnt[1] > 0 && !small) || (LC_TEAM16 && cnt[0] > 0 && !small)))
    {
        const int64_t na = A->nrows * ncp, nb = shared ? 0 : B->nrows * ncp;
        if (na + nb <= LC_POS_MAX && posA.alloc(sizeof(uint32_t) * (na > 0 ? na : 1)) &&
            (shared || posB.alloc(sizeof(uint32_t) * (nb > 0 ? nb : 1))))
        {
            ProfScope ps("dot3_chunkpos");
            const unsigned cap = (unsigned) cu_count() * 32;
            k_d3_chunkpos<<<grid_for(na, D3_NT, cap), D3_NT, 0, st>>>(A->p, A->j, A->nrows, ncp,
                                                                      posA.as<uint32_t>());
            if (!shared)
                k_d3_chunkpos<<<grid_for(nb, D3_NT, cap), D3_NT, 0, st>>>(B->p, B->j, B->nrows, ncp,
                                                                          posB.as<uint32_t>());
            GM_HIP(hipGetLastError(), "k_d3_chunkpos");
            lds_chunk = true;
        }
    }
    const uint32_t *pA = posA.as<uint32_t>(), *pB = shared ? pA : posB.as<uint32_t>();
    // the short rows' columns as 32-bit words halve the chunk kernel's reads
    DevBuf narA, narB;
    bool n32 = false;
    if (LC_NARROW && g.cols32 && g.dA != nullptr &&
        ((small && (cnt[0] > 0 || cnt[1] > 0)) || (lds && (lds_chunk || cnt[0] > 0))))
    {
        const int64_t nza = (int64_t) hc[4], nzb = shared ? 0 : (int64_t) hc[5];
        if (narA.alloc(sizeof(uint32_t) * (nza > 0 ? nza : 1)) &&
            (shared || narB.alloc(sizeof(uint32_t) * (nzb > 0 ? nzb : 1))))
        {
            ProfScope ps("dot3_narrow");
            const unsigned cap = (unsigned) cu_count() * 32;
            k_d3_narrow<<<grid_for(nza, D3_NT, cap), D3_NT, 0, st>>>(A->j, nza, narA.as<uint32_t>());
            if (!shared)
                k_d3_narrow<<<grid_for(nzb, D3_NT, cap), D3_NT, 0, st>>>(B->j, nzb, narB.as<uint32_t>());
            GM_HIP(hipGetLastError(), "k_d3_narrow");
            n32 = true;
        }
    }
    const uint32_t *nA = narA.as<uint32_t>(), *nB = shared ? nA : narB.as<uint32_t>();
    for (int b = 0; b < 3; b++)
    {
        if (cnt[b] == 0) continue;
        {
            ProfScope ps("dot3_group_sort");
            size_t tb = tbytes;
            GM_HIP(rocprim::radix_sort_pairs(tmp.ptr, tb, (const uint64_t *) g.bkeys + (int64_t) b * g.mnz,
                                             tk.as<uint64_t>(),
                                             (const uint64_t *) g.bins + (int64_t) b * g.mnz,
                                             tv.as<uint64_t>(), (size_t) cnt[b], 0, (unsigned) end_bit, st),
                   "grouping sort");
        }
        if (b < 2 && small)
        {
            ProfScope ps(b == 0 ? "dot3_team16_lds1" : "dot3_wave_lds1");
            unsigned long long *c = ctr.as<unsigned long long>() + (b == 0 ? 0 : 8);
            const unsigned grid = (unsigned) (cu_count() * 8);
            if (b == 0 && n32)
                k_dot3_lds1<S, 16, true><<<grid, D3_NT, 0, st>>>(g, tv.as<uint64_t>(), tk.as<uint64_t>(),
                                                                   cnt[0], c, nA, nB);
            else if (b == 0)
                k_dot3_lds1<S, 16, false><<<grid, D3_NT, 0, st>>>(g, tv.as<uint64_t>(), tk.as<uint64_t>(),
                                                                    cnt[0], c, nA, nB);
            else if (n32)
                k_dot3_lds1<S, 64, true><<<grid, D3_NT, 0, st>>>(g, tv.as<uint64_t>(), tk.as<uint64_t>(),
                                                                  cnt[1], c, nA, nB);
            else
                k_dot3_lds1<S, 64, false><<<grid, D3_NT, 0, st>>>(g, tv.as<uint64_t>(), tk.as<uint64_t>(),
                                                                   cnt[1], c, nA, nB);
            GM_HIP(hipGetLastError(), "k_dot3_lds1");
        }
        else if (b == 0 && lds_chunk && LC_TEAM16)
        {
            if constexpr (S::pattern)
            {
                ProfScope ps("dot3_team16_lds");
                if (n32)
                    k_dot3_lds<S, 16, true><<<(unsigned) cu_count(), LC_NT, 0, st>>>(
                        g, tv.as<uint64_t>(), tk.as<uint64_t>(), cnt[0], ctr.as<unsigned long long>() + 16,
                        pA, pB, ncp, nA, nB);
                else
                    k_dot3_lds<S, 16, false><<<(unsigned) cu_count(), LC_NT, 0, st>>>(
                        g, tv.as<uint64_t>(), tk.as<uint64_t>(), cnt[0], ctr.as<unsigned long long>() + 16,
                        pA, pB, ncp, nA, nB);
                GM_HIP(hipGetLastError(), "k_dot3_lds");
            }
        }
        else if (b == 0 && lds)
        {
            // LDS hash tables of the long rows that pay for one (pattern
            // semirings); the rest of the bin on the global path
            if constexpr (S::pattern)
            {
                ProfScope ps("dot3_team16_lh");
                if (n32)
                    k_dot3_team_lh<S, 16, true><<<(unsigned) (cu_count() * 8), D3_NT, 0, st>>>(
                        g, tv.as<uint64_t>(), tk.as<uint64_t>(), cnt[0], ctr.as<unsigned long long>(), nA, nB);
                else
                    k_dot3_team_lh<S, 16, false><<<(unsigned) (cu_count() * 8), D3_NT, 0, st>>>(
                        g, tv.as<uint64_t>(), tk.as<uint64_t>(), cnt[0], ctr.as<unsigned long long>(), nA, nB);
                GM_HIP(hipGetLastError(), "k_dot3_team_lh");
            }
        }
        else if (b == 0)
        {
            // short rows: LDS chunks cost more than they save here
            ProfScope ps("dot3_team16");
            k_dot3_team_grp<S, 16><<<(unsigned) (cu_count() * 8), D3_NT, 0, st>>>(
                g, tv.as<uint64_t>(), cnt[0], ctr.as<unsigned long long>());
            GM_HIP(hipGetLastError(), "k_dot3_team_grp16");
        }
        else if (lds_chunk)
        {
            if constexpr (S::pattern)
            {
                ProfScope ps(b == 1 ? "dot3_wave_lds" : "dot3_block_lds");
                const unsigned grid = (unsigned) cu_count();
                unsigned long long *c = ctr.as<unsigned long long>() + (b == 1 ? 24 : 32);
                if (b == 1 && n32)
                    k_dot3_lds<S, LC_WAVE_G, true><<<grid, LC_NT, 0, st>>>(
                        g, tv.as<uint64_t>(), tk.as<uint64_t>(), cnt[b], c, pA, pB, ncp, nA, nB);
                else if (b == 1)
                    k_dot3_lds<S, LC_WAVE_G, false><<<grid, LC_NT, 0, st>>>(
                        g, tv.as<uint64_t>(), tk.as<uint64_t>(), cnt[b], c, pA, pB, ncp, nA, nB);
                else if (n32)
                    k_dot3_lds<S, 64, true><<<grid, LC_NT, 0, st>>>(
                        g, tv.as<uint64_t>(), tk.as<uint64_t>(), cnt[b], c, pA, pB, ncp, nA, nB);
                else
                    k_dot3_lds<S, 64, false><<<grid, LC_NT, 0, st>>>(
                        g, tv.as<uint64_t>(), tk.as<uint64_t>(), cnt[b], c, pA, pB, ncp, nA, nB);
                GM_HIP(hipGetLastError(), "k_dot3_lds");
            }
        }
        else if (b == 1)
        {
            ProfScope ps("dot3_wave");
            k_dot3_team_grp<S, 64><<<(unsigned) (cu_count() * 8), D3_NT, 0, st>>>(
                g, tv.as<uint64_t>(), cnt[1], ctr.as<unsigned long long>() + 8);
            GM_HIP(hipGetLastError(), "k_dot3_team_grp64");
        }
        else
        {
            ProfScope ps("dot3_block");
            k_dot3_block_grp<S><<<(unsigned) (cu_count() * 8), D3_NT, 0, st>>>(
                g, tv.as<uint64_t>(), cnt[2], ctr.as<unsigned long long>() + 16);
            GM_HIP(hipGetLastError(), "k_dot3_block_grp");
        }
    }
    return GM_SUCCESS;
}

template <class S>
static int dot3_run(gm_csr_out *C, const gm_csr *M, bool mask_struct, const gm_csr *A,
                    const gm_csr *B, const gm_dot3_params *prm, int64_t mnz_known = -1)
{
    using Z = typename S::Z;
    hipStream_t st = stream();
    const int64_t n = M->nrows;
    int64_t Mnz = (n > 0) ? mnz_known : 0;
    if (Mnz < 0)
    {
        // nnz(M) from its last row pointer
        int64_t *hp = (int64_t *) pinned_scratch(sizeof(int64_t));
        if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
        GM_HIP(hipMemcpyAsync(hp, M->p + n, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz(M)");
        GM_HIP(hipStreamSynchronize(st), "nnz(M) sync");
        Mnz = *hp;
    }
    if (Mnz >= ((int64_t) 1 << 32) || n >= ((int64_t) 1 << 32))
        return set_error(GM_NO_VALUE, "dot3: nnz(M) or nrows beyond 2^32");

    const size_t zs = sizeof(Z);
    DevBuf Cp, Ci, Cx, W, live, bins, cnt;
    if (!Cp.alloc(sizeof(int64_t) * (n + 1)) || !Ci.alloc(sizeof(int64_t) * (Mnz > 0 ? Mnz : 1)) ||
        !Cx.alloc(S::iso_out ? zs : zs * (Mnz > 0 ? Mnz : 1)))
        return set_error(GM_OUT_OF_MEMORY, "dot3: output");
    int64_t nvals = 0;
    if (Mnz > 0)
    {
        if (!W.alloc(S::iso_out ? 8 : zs * Mnz) || !live.alloc(Mnz) ||
            !bins.alloc(sizeof(uint64_t) * 3 * Mnz) || !cnt.alloc(4 * sizeof(unsigned long long)))
            return set_error(GM_OUT_OF_MEMORY, "dot3: workspace");
        Dot3Args g;
        g.Mp = M->p; g.Mj = M->j; g.Mx = mask_struct ? nullptr : M->x;
        g.msize = (int32_t) type_size(M->type); g.m_iso = M->iso; g.mask_struct = mask_struct ? 1 : 0;
        if (!mask_struct && M->x == nullptr) g.mask_struct = 1;
        g.m_nrows = n; g.mnz = Mnz;
        g.Ap = A->p; g.Aj = A->j; g.Ax = A->x; g.a_iso = A->iso;
        g.Bp = B->p; g.Bj = B->j; g.Bx = B->x; g.b_iso = B->iso;
        g.W = W.ptr; g.live = live.as<uint8_t>(); g.bins = bins.as<uint64_t>();
        g.bin_cnt = cnt.as<unsigned long long>();
        g.cols32 = (A->ncols < ((int64_t) 1 << 32)) ? 1 : 0;
        g.bkeys = nullptr;
        g.key_nA = A->nrows;
        // group deferred entries by probed row: explicit, or by default for
        // masks large enough that the sort pays for itself
        const gm_dot3_params tu = dot3_tuned();
        const int64_t order = (prm && prm->order) ? prm->order : tu.order;
        const bool grouped = (order == 2) || (order == 0 && Mnz >= ((int64_t) 1 << 20));
        DevBuf bkeys;
        if (grouped)
        {
            if (!bkeys.alloc(sizeof(uint64_t) * 3 * Mnz))
                return set_error(GM_OUT_OF_MEMORY, "dot3: grouping keys");
            g.bkeys = bkeys.as<uint64_t>();
        }
        g.thread_max = pick_param(prm ? prm->thread_max : 0, tu.thread_max, 8);
        g.team16_max = pick_param(prm ? prm->team16_max : 0, tu.team16_max, 512);
        g.wave_max = pick_param(prm ? prm->wave_max : 0, tu.wave_max, 4096);
        const int64_t cap = (int64_t) cu_count() * 16;
        g.bA = RowBits{INT64_MAX, nullptr, nullptr, nullptr};
        g.bB = g.bA;
        DevBuf bm_offA, bm_offB, bm_words, bm_rank, bm_list, bm_ctr;
        int64_t minlen = (prm && prm->bitmap_min_len != 0) ? prm->bitmap_min_len
                       : (tu.bitmap_min_len != 0 ? tu.bitmap_min_len : 128);
        if (A->ncols >= ((int64_t) 1 << 32)) minlen = -1;   // hash keys are 32-bit columns
        if (minlen > 0)
            GM_TRY(dot3_bitmaps(g, A, B, !S::pattern, minlen, bm_offA, bm_offB, bm_words, bm_rank,
                                bm_list, bm_ctr));
        k_zero_u64<<<1, 64, 0, st>>>(g.bin_cnt, 4);
        // row summaries for pair setup (columns must fit 31 bits); bin_cnt[3]
        // receives the largest column span of a row
        g.dA = nullptr;
        g.dB = nullptr;
        DevBuf descA, descB;
        if (A->ncols < ((int64_t) 1 << 31))
        {
            const bool shared = (A->p == B->p && A->j == B->j && A->nrows == B->nrows);
            if (descA.alloc(sizeof(RowDesc) * (A->nrows > 0 ? A->nrows : 1)) &&
                (shared || descB.alloc(sizeof(RowDesc) * (B->nrows > 0 ? B->nrows : 1))))
            {
                ProfScope ps("dot3_desc");
                k_d3_desc<<<grid_for(A->nrows, D3_NT, cap), D3_NT, 0, st>>>(
                    A->p, A->j, A->nrows, g.bA.T, g.bA.off, descA.as<RowDesc>(), g.bin_cnt + 3);
                if (!shared)
                    k_d3_desc<<<grid_for(B->nrows, D3_NT, cap), D3_NT, 0, st>>>(
                        B->p, B->j, B->nrows, g.bB.T, g.bB.off, descB.as<RowDesc>(), g.bin_cnt + 3);
                GM_HIP(hipGetLastError(), "k_d3_desc");
                g.dA = descA.as<RowDesc>();
                g.dB = shared ? g.dA : descB.as<RowDesc>();
            }
        }
        // whole rows to the row kernel: mask rows with at least row_min
        // entries whose A row spans fewer than RM_BITS columns
        const int64_t row_min = (prm && prm->row_min != 0) ? prm->row_min
                              : (tu.row_min != 0 ? tu.row_min : -1);
        g.rowmode = nullptr;
        DevBuf rm_flag, rm_rows, rm_cnt;
        if (row_min > 0 && g.cols32)
        {
            if (!rm_flag.alloc(n) || !rm_rows.alloc(sizeof(uint32_t) * n) || !rm_cnt.alloc(8))
                return set_error(GM_OUT_OF_MEMORY, "dot3: row mode");
            k_zero_u64<<<1, 64, 0, st>>>(rm_cnt.as<unsigned long long>(), 1);
            ProfScope ps("dot3_rowmode");
            k_d3_rowmode<<<grid_for(n, D3_NT, cap), D3_NT, 0, st>>>(
                M->p, A->p, A->j, n, row_min, 8, rm_flag.as<uint8_t>(), rm_rows.as<uint32_t>(),
                rm_cnt.as<unsigned long long>());
            GM_HIP(hipGetLastError(), "k_d3_rowmode");
            g.rowmode = rm_flag.as<uint8_t>();
        }
        {
            ProfScope ps("dot3_phase1");
            k_dot3_phase1<S><<<grid_for(Mnz, D3_P1_TILE, cap), D3_NT, 0, st>>>(g);
            GM_HIP(hipGetLastError(), "k_dot3_phase1");
        }
        if (grouped)
        {
            GM_TRY(dot3_grouped_teams<S>(g, A, B));
        }
        else
        {
            {
                ProfScope ps("dot3_team16");
                k_dot3_team<S, 16><<<(unsigned) (cu_count() * 8), D3_NT, 0, st>>>(g, 0);
                GM_HIP(hipGetLastError(), "k_dot3_team16");
            }
            {
                ProfScope ps("dot3_wave");
                k_dot3_team<S, 64><<<(unsigned) (cu_count() * 8), D3_NT, 0, st>>>(g, 1);
                GM_HIP(hipGetLastError(), "k_dot3_wave");
            }
        }
        if (!grouped)
        {
            ProfScope ps("dot3_block");
            k_dot3_block<S><<<(unsigned) (cu_count() * 8), D3_NT, 0, st>>>(g);
            GM_HIP(hipGetLastError(), "k_dot3_block");
        }
        if (g.rowmode)
        {
            ProfScope ps("dot3_rows");
            k_dot3_rows<S><<<(unsigned) (cu_count() * 8), D3_NT, 0, st>>>(
                g, rm_rows.as<uint32_t>(), rm_cnt.as<unsigned long long>());
            GM_HIP(hipGetLastError(), "k_dot3_rows");
        }
        // compaction
        const int64_t ntiles = (Mnz + D3_CT_TILE - 1) / D3_CT_TILE;
        DevBuf toff;
        if (!toff.alloc(sizeof(int64_t) * (ntiles + 1)))
            return set_error(GM_OUT_OF_MEMORY, "dot3: compaction workspace");
        {
            ProfScope ps("dot3_tile_count");
            k_dot3_tile_count<<<(unsigned) ntiles, D3_NT, 0, st>>>(live.as<uint8_t>(), Mnz,
                                                                   toff.as<int64_t>());
            GM_HIP(hipGetLastError(), "k_dot3_tile_count");
        }
        GM_TRY(scan_exclusive_i64(toff.as<int64_t>(), toff.as<int64_t>(), ntiles));
        {
            ProfScope ps("dot3_compact");
            k_dot3_compact<Z, S::iso_out><<<(unsigned) ntiles, D3_NT, 0, st>>>(
                M->p, M->j, n, Mnz, live.as<uint8_t>(), (const Z *) W.ptr, toff.as<int64_t>(),
                Cp.as<int64_t>(), Ci.as<int64_t>(), (Z *) Cx.ptr);
            GM_HIP(hipGetLastError(), "k_dot3_compact");
        }
        int64_t *hp = (int64_t *) pinned_scratch(sizeof(int64_t));
        GM_HIP(hipMemcpyAsync(hp, toff.as<int64_t>() + ntiles, sizeof(int64_t),
                              hipMemcpyDeviceToHost, st), "nvals");
        GM_HIP(hipStreamSynchronize(st), "dot3 sync");
        nvals = *hp;
    }
    else
    {
        k_fill_i64<<<grid_for(n + 1, 256, 1024), 256, 0, st>>>(Cp.as<int64_t>(), n + 1, 0);
        GM_HIP(hipGetLastError(), "k_fill_i64");
    }
    if (S::iso_out)
    {
        k_set_iso_one<<<1, 1, 0, st>>>(Cx.ptr, (int) zs);
        GM_HIP(hipGetLastError(), "k_set_iso_one");
    }
    C->nrows = n;
    C->ncols = B->nrows;
    C->nvals = nvals;
    C->p = (int64_t *) Cp.release();
    C->j = (int64_t *) Ci.release();
    C->x = Cx.release();
    C->iso = S::iso_out ? 1 : 0;
    C->type = 0;
    return GM_SUCCESS;
}

} // namespace gm

using namespace gm;


//------------------------------------------------------------------------------
// dot2: C = A*B' or C<!M> = A*B' with every dot product of the small result
// computed (the reference's bitmap dot product, Source/mxm/GB_AxB_dot2.c).
// Here every pair (i, j) of a nonempty row of A and a nonempty row of B that
// the mask allows becomes an entry of a structural mask, and dot3 computes
// them: the same dot products, in the same k order, and a compacted sparse C
// in place of the reference's bitmap.
//------------------------------------------------------------------------------

// f[r] = 1 when row r is nonempty
__global__ __launch_bounds__(D3_NT) void k_d2_nonempty(const int64_t *__restrict__ Xp, int64_t n,
                                                       int64_t *__restrict__ f)
{
    for (int64_t r = (int64_t) blockIdx.x * D3_NT + threadIdx.x; r < n; r += (int64_t) gridDim.x * D3_NT)
        f[r] = Xp[r + 1] > Xp[r] ? 1 : 0;
}

// list[pos[r]] = r for the nonempty rows
__global__ __launch_bounds__(D3_NT) void k_d2_list(const int64_t *__restrict__ Xp, int64_t n,
                                                   const int64_t *__restrict__ pos, int64_t *__restrict__ list)
{
    for (int64_t r = (int64_t) blockIdx.x * D3_NT + threadIdx.x; r < n; r += (int64_t) gridDim.x * D3_NT)
        if (Xp[r + 1] > Xp[r]) list[pos[r]] = r;
}

// entries of the full mask in row i: all nonempty rows of B when A(i,:) is
// nonempty, less the ones where M(i,j) is true (C<!M>)
__global__ __launch_bounds__(D3_NT) void k_d2_count(const int64_t *__restrict__ Ap, int64_t m, int64_t bnvec,
                                                    const int64_t *__restrict__ fB, const int64_t *Mp,
                                                    const int64_t *Mj, const void *Mx, int32_t msize,
                                                    int32_t m_iso, int32_t mask_struct, int64_t *__restrict__ cnt)
{
    for (int64_t i = (int64_t) blockIdx.x * D3_NT + threadIdx.x; i < m; i += (int64_t) gridDim.x * D3_NT)
    {
        int64_t c = 0;
        if (Ap[i + 1] > Ap[i])
        {
            c = bnvec;
            if (Mp)
                for (int64_t e = Mp[i]; e < Mp[i + 1]; e++)
                    if (fB[Mj[e]] && (mask_struct || mask_true(Mx, e, msize, m_iso))) c--;
        }
        cnt[i] = c;
    }
}

// one wavefront per row: the row's list of allowed columns, in order
__global__ __launch_bounds__(D3_NT) void k_d2_fill(const int64_t *__restrict__ Ap, int64_t m,
                                                   const int64_t *__restrict__ listB, int64_t bnvec,
                                                   const int64_t *Mp, const int64_t *Mj, const void *Mx,
                                                   int32_t msize, int32_t m_iso, int32_t mask_struct,
                                                   const int64_t *__restrict__ Fp, int64_t *__restrict__ Fj)
{
    const int lane = threadIdx.x & 63;
    const int64_t wv = ((int64_t) blockIdx.x * D3_NT + threadIdx.x) >> 6;
    const int64_t nwv = ((int64_t) gridDim.x * D3_NT) >> 6;
    for (int64_t i = wv; i < m; i += nwv)
    {
        if (Ap[i + 1] == Ap[i]) continue;
        int64_t out = Fp[i];
        const int64_t m0 = Mp ? Mp[i] : 0, m1 = Mp ? Mp[i + 1] : 0;
        for (int64_t t0 = 0; t0 < bnvec; t0 += 64)
        {
            const int64_t t = t0 + lane;
            bool keep = false;
            int64_t j = 0;
            if (t < bnvec)
            {
                j = listB[t];
                keep = true;
                if (m1 > m0)
                {
                    const int64_t q = lower_bound(Mj, m0, m1, j);
                    if (q < m1 && Mj[q] == j && (mask_struct || mask_true(Mx, q, msize, m_iso))) keep = false;
                }
            }
            const unsigned long long bal = __ballot(keep);
            if (keep) Fj[out + __popcll(bal & ((1ull << lane) - 1ull))] = j;
            out += __popcll(bal);
        }
    }
}

template <class S>
static int dot2_run(gm_csr_out *C, const gm_csr *M, bool mask_comp, bool mask_struct, const gm_csr *A,
                    const gm_csr *B, const gm_dot3_params *prm)
{
    (void) mask_comp;
    hipStream_t st = stream();
    const int64_t m = A->nrows, nb = B->nrows;
    const int64_t cap = (int64_t) cu_count() * 16;
    DevBuf fA, fB, posB, listB, cnt, Fp, Fj;
    if (!fA.alloc(sizeof(int64_t) * (m > 0 ? m : 1)) || !fB.alloc(sizeof(int64_t) * (nb > 0 ? nb : 1)) ||
        !posB.alloc(sizeof(int64_t) * (nb + 1)) || !cnt.alloc(sizeof(int64_t) * (m > 0 ? m : 1)) ||
        !Fp.alloc(sizeof(int64_t) * (m + 1)))
        return set_error(GM_OUT_OF_MEMORY, "dot2: workspace");
    if (nb > 0)
    {
        k_d2_nonempty<<<grid_for(nb, D3_NT, cap), D3_NT, 0, st>>>(B->p, nb, fB.as<int64_t>());
        GM_HIP(hipGetLastError(), "k_d2_nonempty");
    }
    GM_TRY(scan_exclusive_i64(fB.as<int64_t>(), posB.as<int64_t>(), nb));
    int64_t *hp = (int64_t *) pinned_scratch(2 * sizeof(int64_t));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(hp, posB.as<int64_t>() + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st), "bnvec");
    GM_HIP(hipStreamSynchronize(st), "bnvec sync");
    const int64_t bnvec = hp[0];
    if (!listB.alloc(sizeof(int64_t) * (bnvec > 0 ? bnvec : 1)))
        return set_error(GM_OUT_OF_MEMORY, "dot2: row list");
    if (nb > 0)
    {
        k_d2_list<<<grid_for(nb, D3_NT, cap), D3_NT, 0, st>>>(B->p, nb, posB.as<int64_t>(), listB.as<int64_t>());
        GM_HIP(hipGetLastError(), "k_d2_list");
    }
    const bool use_m = M != nullptr;
    const void *Mx = (use_m && !mask_struct) ? M->x : nullptr;
    const int32_t mstruct = (!use_m || mask_struct || !M->x) ? 1 : 0;
    const int32_t msize = use_m ? (int32_t) type_size(M->type) : 0;
    if (m > 0)
    {
        k_d2_count<<<grid_for(m, D3_NT, cap), D3_NT, 0, st>>>(
            A->p, m, bnvec, fB.as<int64_t>(), use_m ? M->p : nullptr, use_m ? M->j : nullptr, Mx, msize,
            use_m ? M->iso : 0, mstruct, cnt.as<int64_t>());
        GM_HIP(hipGetLastError(), "k_d2_count");
    }
    GM_TRY(scan_exclusive_i64(cnt.as<int64_t>(), Fp.as<int64_t>(), m));
    GM_HIP(hipMemcpyAsync(hp + 1, Fp.as<int64_t>() + m, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz");
    GM_HIP(hipStreamSynchronize(st), "nnz sync");
    const int64_t fnz = hp[1];
    if (fnz >= ((int64_t) 1 << 32))
        return set_error(GM_NO_VALUE, "dot2: %lld dot products exceed the mask index range", (long long) fnz);
    if (!Fj.alloc(sizeof(int64_t) * (fnz > 0 ? fnz : 1)))
        return set_error(GM_OUT_OF_MEMORY, "dot2: %lld dot products", (long long) fnz);
    if (m > 0 && fnz > 0)
    {
        k_d2_fill<<<grid_for(m * 64, D3_NT, cap), D3_NT, 0, st>>>(
            A->p, m, listB.as<int64_t>(), bnvec, use_m ? M->p : nullptr, use_m ? M->j : nullptr, Mx, msize,
            use_m ? M->iso : 0, mstruct, Fp.as<int64_t>(), Fj.as<int64_t>());
        GM_HIP(hipGetLastError(), "k_d2_fill");
    }
    gm_csr F{};
    F.nrows = m;
    F.ncols = nb;
    F.p = Fp.as<int64_t>();
    F.j = Fj.as<int64_t>();
    F.x = nullptr;
    F.iso = 0;
    F.type = GM_BOOL;
    return dot3_run<S>(C, &F, true, A, B, prm);
}

extern "C" bool gm_dot3_branch(const gm_csr *M, bool mask_struct, const gm_csr *A,
                               const gm_csr *B, const gm_semiring *sr)
{
    (void) mask_struct;
    if (!M || !A || !B || !sr) return false;
    if (!semiring_supported(sr)) return false;
    if (!operand_ok(sr, A) || !operand_ok(sr, B)) return false;
    // C<M> = A*B': M is nrows(A) x nrows(B), A and B share their column space
    if (M->nrows != A->nrows || M->ncols != B->nrows || A->ncols != B->ncols) return false;
    return true;
}

// multipliers that give the same value with their operands swapped
static bool mult_commutes(int32_t mult_op)
{
    return mult_op == GM_MULT_TIMES || mult_op == GM_MULT_PLUS || mult_op == GM_MULT_PAIR ||
           mult_op == GM_MULT_K;
}

extern "C" bool gm_AxB_dot3_branch(const gm_csr *M, bool mask_struct, const gm_csr *A,
                                   const gm_csr *B, const gm_semiring *sr, bool flipxy)
{
    if (!sr || (flipxy && !mult_commutes(sr->mult_op))) return false;
    return gm_dot3_branch(M, mask_struct, A, B, sr);
}

extern "C" int gm_AxB_dot3(gm_csr_out *C, const gm_csr *M, bool mask_struct, const gm_csr *A,
                           const gm_csr *B, const gm_semiring *sr, bool flipxy)
{
    if (!sr) return set_error(GM_INVALID_VALUE, "gm_AxB_dot3: NULL semiring");
    if (flipxy && !mult_commutes(sr->mult_op))
        return set_error(GM_NO_VALUE, "gm_AxB_dot3: flipxy with a multiplier that does not commute");
    return gm_dot3(C, M, mask_struct, A, B, sr, nullptr);
}

static int dot3_dispatch(gm_csr_out *C, const gm_csr *M, bool mask_struct, const gm_csr *A,
                         const gm_csr *B, const gm_semiring *sr, const gm_dot3_params *params,
                         int64_t mnz_known = -1)
{
    int rc = GM_NO_VALUE;
#define GM_DOT3_CASE(ADD, MULT, ZC, Z)                                         \
    if (sr->add_op == ADD && sr->mult_op == MULT && sr->ztype == ZC)           \
    {                                                                          \
        rc = dot3_run<SR<ADD, MULT, Z>>(C, M, mask_struct, A, B, params, mnz_known); \
        if (rc == GM_SUCCESS) C->type = ZC;                                    \
        return rc;                                                             \
    }
    GM_SEMIRING_LIST(GM_DOT3_CASE)
#undef GM_DOT3_CASE
    return rc;
}

// The deferred entries are packed as (row << 32 | mask position), so one run
// takes fewer than 2^32 mask entries and rows.  A larger mask is cut into row
// slices below that (g_dot3_slice_max entries; tests lower it), each run on
// its own (its mask row pointers re-based to 0, A's rows viewed in place), and
// the slices' results are joined.  No counterpart in the reference, whose
// CPU kernels index with 64-bit integers throughout.
static int64_t g_dot3_slice_max = (int64_t) 1 << 32;

extern "C" void gm_debug_dot3_slice_max(int64_t v)
{
    g_dot3_slice_max = (v > 0 && v < ((int64_t) 1 << 32)) ? v : ((int64_t) 1 << 32);
}

__global__ void k_d3_rebase(const int64_t *__restrict__ src, int64_t n, int64_t base, int64_t add,
                            int64_t *__restrict__ dst)
{
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t) gridDim.x * blockDim.x)
        dst[t] = src[t] - base + add;
}

static int dot3_sliced(gm_csr_out *C, const gm_csr *M, bool mask_struct, const gm_csr *A,
                       const gm_csr *B, const gm_semiring *sr, const gm_dot3_params *params,
                       int64_t slice_max)
{
    hipStream_t st = stream();
    const int64_t n = M->nrows;
    std::vector<int64_t> hp((size_t) n + 1);
    GM_HIP(hipMemcpyAsync(hp.data(), M->p, sizeof(int64_t) * (size_t) (n + 1), hipMemcpyDeviceToHost, st),
           "dot3 slices: Mp");
    GM_HIP(hipStreamSynchronize(st), "dot3 slices: Mp sync");
    const int64_t row_max = ((int64_t) 1 << 31);
    // row slices: as many rows as fit below slice_max entries (at least one)
    std::vector<int64_t> cut{0};
    while (cut.back() < n)
    {
        const int64_t r0 = cut.back();
        int64_t lo = r0 + 1, hi = (n < r0 + row_max) ? n : r0 + row_max;
        while (lo < hi)
        {
            const int64_t mid = lo + (hi - lo + 1) / 2;
            if (hp[(size_t) mid] - hp[(size_t) r0] < slice_max) lo = mid; else hi = mid - 1;
        }
        if (hp[(size_t) lo] - hp[(size_t) r0] >= ((int64_t) 1 << 32))
            return set_error(GM_NO_VALUE, "dot3: a mask row of 2^32 entries or more");
        cut.push_back(lo);
    }
    const size_t ns = cut.size() - 1;
    std::vector<gm_csr_out> part(ns);
    for (auto &o : part) memset(&o, 0, sizeof(o));
    auto free_parts = [&]() {
        for (auto &o : part)
        {
            gm_free(o.p);
            gm_free(o.j);
            gm_free(o.x);
        }
    };
    DevBuf mp;
    int64_t rmax = 0;
    for (size_t k = 0; k < ns; k++) rmax = std::max(rmax, cut[k + 1] - cut[k]);
    if (!mp.alloc(sizeof(int64_t) * (size_t) (rmax + 1))) return set_error(GM_OUT_OF_MEMORY, "dot3 slices");
    const size_t msz = type_size(M->type);
    for (size_t k = 0; k < ns; k++)
    {
        const int64_t r0 = cut[k], r1 = cut[k + 1], base = hp[(size_t) r0];
        k_d3_rebase<<<grid_for(r1 - r0 + 1, 256, (int64_t) cu_count() * 8), 256, 0, st>>>(
            M->p + r0, r1 - r0 + 1, base, 0, mp.as<int64_t>());
        GM_HIP(hipGetLastError(), "k_d3_rebase");
        gm_csr Ms = *M, As = *A;
        Ms.nrows = r1 - r0;
        Ms.p = mp.as<int64_t>();
        Ms.j = M->j + base;
        if (M->x && !M->iso) Ms.x = (const char *) M->x + (size_t) base * msz;
        Ms.aux = nullptr;
        As.nrows = r1 - r0;
        As.p = A->p + r0;               // positions stay absolute in A's arrays
        As.aux = nullptr;
        const int rc = dot3_dispatch(&part[k], &Ms, mask_struct, &As, B, sr, params);
        if (rc != GM_SUCCESS)
        {
            free_parts();
            return rc;
        }
    }
    // join: row pointers shifted by the entries of the slices before
    int64_t total = 0;
    for (auto &o : part) total += o.nvals;
    const bool iso = part[0].iso != 0;
    const size_t zs = type_size(part[0].type);
    gm_csr_out R;
    memset(&R, 0, sizeof(R));
    R.nrows = n;
    R.ncols = M->ncols;
    R.nvals = total;
    R.iso = iso ? 1 : 0;
    R.type = part[0].type;
    R.p = (int64_t *) gm_malloc(sizeof(int64_t) * (size_t) (n + 1));
    R.j = (int64_t *) gm_malloc(sizeof(int64_t) * (size_t) (total > 0 ? total : 1));
    R.x = gm_malloc(zs * (size_t) (iso ? 1 : (total > 0 ? total : 1)));
    if (!R.p || !R.j || !R.x)
    {
        gm_free(R.p);
        gm_free(R.j);
        gm_free(R.x);
        free_parts();
        return set_error(GM_OUT_OF_MEMORY, "dot3 slices: result");
    }
    int64_t off = 0;
    for (size_t k = 0; k < ns; k++)
    {
        const int64_t r0 = cut[k], r1 = cut[k + 1];
        const gm_csr_out &o = part[k];
        k_d3_rebase<<<grid_for(r1 - r0 + 1, 256, (int64_t) cu_count() * 8), 256, 0, st>>>(
            o.p, r1 - r0 + 1, 0, off, R.p + r0);
        GM_HIP(hipGetLastError(), "k_d3_rebase");
        if (o.nvals > 0)
        {
            GM_HIP(hipMemcpyAsync(R.j + off, o.j, sizeof(int64_t) * (size_t) o.nvals, hipMemcpyDeviceToDevice, st),
                   "dot3 slices: join j");
            if (!iso)
                GM_HIP(hipMemcpyAsync((char *) R.x + (size_t) off * zs, o.x, zs * (size_t) o.nvals,
                                      hipMemcpyDeviceToDevice, st), "dot3 slices: join x");
        }
        off += o.nvals;
    }
    if (iso) GM_HIP(hipMemcpyAsync(R.x, part[0].x, zs, hipMemcpyDeviceToDevice, st), "dot3 slices: iso");
    GM_HIP(hipStreamSynchronize(st), "dot3 slices: join sync");
    free_parts();
    *C = R;
    return GM_SUCCESS;
}

extern "C" int gm_dot3(gm_csr_out *C, const gm_csr *M, bool mask_struct, const gm_csr *A,
                       const gm_csr *B, const gm_semiring *sr, const gm_dot3_params *params)
{
    if (!C || !M || !A || !B || !sr) return set_error(GM_INVALID_VALUE, "gm_dot3: NULL argument");
    if (!gm_dot3_branch(M, mask_struct, A, B, sr))
        return set_error(GM_NO_VALUE, "gm_dot3: case not handled by the GPU kernels");
    GM_TRY(ensure_init());
    int64_t mnz = -1;
    if (M->nrows > 0 && M->p)
    {
        int64_t *hp = (int64_t *) pinned_scratch(sizeof(int64_t));
        if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
        hipStream_t st = stream();
        GM_HIP(hipMemcpyAsync(hp, M->p + M->nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz(M)");
        GM_HIP(hipStreamSynchronize(st), "nnz(M) sync");
        mnz = *hp;
        const int64_t slice_max = g_dot3_slice_max;
        if (mnz >= slice_max || M->nrows >= ((int64_t) 1 << 31))
            return dot3_sliced(C, M, mask_struct, A, B, sr, params, slice_max);
    }
    return dot3_dispatch(C, M, mask_struct, A, B, sr, params, mnz);
}

// sums over the mask entries (i,j) of min(|A(i,:)|, |B(j,:)|) and of
// |A(i,:)| + |B(j,:)|: the list lengths the probe and merge models of a dot3
// call read (one thread per mask row, a partial per workgroup)
__global__ __launch_bounds__(D3_NT) void k_d3_work(const int64_t *__restrict__ Mp,
                                                   const int64_t *__restrict__ Mj, int64_t n,
                                                   const int64_t *__restrict__ Ap,
                                                   const int64_t *__restrict__ Bp,
                                                   unsigned long long *__restrict__ part)
{
    __shared__ unsigned long long s[D3_NT / 64 + 1];
    unsigned long long smin = 0, ssum = 0;
    for (int64_t i = (int64_t) blockIdx.x * D3_NT + threadIdx.x; i < n; i += (int64_t) gridDim.x * D3_NT)
    {
        const int64_t a = Ap[i + 1] - Ap[i];
        for (int64_t p = Mp[i]; p < Mp[i + 1]; p++)
        {
            const int64_t j = Mj[p];
            const int64_t b = Bp[j + 1] - Bp[j];
            smin += (unsigned long long) (a < b ? a : b);
            ssum += (unsigned long long) (a + b);
        }
    }
    const unsigned long long tmin = block_sum<unsigned long long, D3_NT>(smin, s);
    const unsigned long long tsum = block_sum<unsigned long long, D3_NT>(ssum, s);
    if (threadIdx.x == 0)
    {
        atomicAdd(&part[0], tmin);
        atomicAdd(&part[1], tsum);
    }
}

extern "C" int gm_dot3_work(int64_t *min_sum, int64_t *len_sum, const gm_csr *M, const gm_csr *A,
                            const gm_csr *B)
{
    if (!min_sum || !len_sum || !M || !A || !B)
        return set_error(GM_INVALID_VALUE, "gm_dot3_work: NULL argument");
    if (M->nrows != A->nrows || M->ncols != B->nrows)
        return set_error(GM_INVALID_VALUE, "gm_dot3_work: M must be nrows(A) x nrows(B)");
    *min_sum = 0;
    *len_sum = 0;
    if (M->nrows == 0) return GM_SUCCESS;
    if (!M->p || !A->p || !B->p) return set_error(GM_INVALID_VALUE, "gm_dot3_work: NULL row pointers");
    GM_TRY(ensure_init());
    hipStream_t st = stream();
    DevBuf part;
    if (!part.alloc(2 * sizeof(unsigned long long))) return set_error(GM_OUT_OF_MEMORY, "gm_dot3_work");
    k_zero_u64<<<1, 64, 0, st>>>(part.as<unsigned long long>(), 2);
    k_d3_work<<<grid_for(M->nrows, D3_NT, (int64_t) cu_count() * 8), D3_NT, 0, st>>>(
        M->p, M->j, M->nrows, A->p, B->p, part.as<unsigned long long>());
    GM_HIP(hipGetLastError(), "k_d3_work");
    unsigned long long *hp = (unsigned long long *) pinned_scratch(2 * sizeof(unsigned long long));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(hp, part.ptr, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st), "work");
    GM_HIP(hipStreamSynchronize(st), "gm_dot3_work sync");
    *min_sum = (int64_t) hp[0];
    *len_sum = (int64_t) hp[1];
    return GM_SUCCESS;
}

extern "C" int gm_nonempty_rows(int64_t *count, const gm_csr *A)
{
    if (!count || !A) return set_error(GM_INVALID_VALUE, "gm_nonempty_rows: NULL argument");
    *count = 0;
    if (A->nrows <= 0) return GM_SUCCESS;
    if (!A->p) return set_error(GM_INVALID_VALUE, "gm_nonempty_rows: NULL row pointers");
    GM_TRY(ensure_init());
    DevBuf f, pos;
    if (!f.alloc(sizeof(int64_t) * A->nrows) || !pos.alloc(sizeof(int64_t) * (A->nrows + 1)))
        return set_error(GM_OUT_OF_MEMORY, "gm_nonempty_rows: workspace");
    hipStream_t st = stream();
    k_d2_nonempty<<<grid_for(A->nrows, D3_NT, (int64_t) cu_count() * 16), D3_NT, 0, st>>>(A->p, A->nrows,
                                                                                         f.as<int64_t>());
    GM_HIP(hipGetLastError(), "k_d2_nonempty");
    GM_TRY(scan_exclusive_i64(f.as<int64_t>(), pos.as<int64_t>(), A->nrows));
    int64_t *hp = (int64_t *) pinned_scratch(sizeof(int64_t));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(hp, pos.as<int64_t>() + A->nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st),
           "nonempty count");
    GM_HIP(hipStreamSynchronize(st), "nonempty sync");
    *count = hp[0];
    return GM_SUCCESS;
}

extern "C" int gm_dot2(gm_csr_out *C, const gm_csr *M, bool mask_comp, bool mask_struct, const gm_csr *A,
                       const gm_csr *B, const gm_semiring *sr, const gm_dot3_params *params)
{
    if (!C || !A || !B || !sr) return set_error(GM_INVALID_VALUE, "gm_dot2: NULL argument");
    if (M && !mask_comp)
        return set_error(GM_NO_VALUE, "gm_dot2: a mask that is not complemented goes to dot3");
    if (!semiring_supported(sr) || !operand_ok(sr, A) || !operand_ok(sr, B))
        return set_error(GM_NO_VALUE, "gm_dot2: case not handled by the GPU kernels");
    if (A->ncols != B->ncols || (M && (M->nrows != A->nrows || M->ncols != B->nrows)))
        return set_error(GM_INVALID_VALUE, "gm_dot2: dimensions");
    if (A->nrows >= ((int64_t) 1 << 32)) return set_error(GM_NO_VALUE, "gm_dot2: nrows beyond 2^32");
    GM_TRY(ensure_init());
#define GM_DOT2_CASE(ADD, MULT, ZC, Z)                                                      \
    if (sr->add_op == ADD && sr->mult_op == MULT && sr->ztype == ZC)                        \
    {                                                                                       \
        int rc = dot2_run<SR<ADD, MULT, Z>>(C, M, mask_comp, mask_struct, A, B, params);    \
        if (rc == GM_SUCCESS) C->type = ZC;                                                 \
        return rc;                                                                          \
    }
    GM_SEMIRING_LIST(GM_DOT2_CASE)
#undef GM_DOT2_CASE
    return set_error(GM_NO_VALUE, "gm_dot2: semiring not instantiated");
}
