// gm_dense.hip: bitmap and full matrices, and the products whose result is
// one of them.
//
// A bitmap or full matrix is held by rows: entry (i,j) at i*ncols + j, with a
// byte b[i*ncols + j] != 0 for a present entry (bitmap) or every entry
// present (full, b absent).  Reference: the bitmap/full formats of
// Source/builtin/include/GB_matrix.h and the dense-result products
//   saxbit  C = A*B, C<M> = A*B, C<!M> = A*B with C bitmap
//           (Source/mxm/GB_AxB_saxbit.c),
//   saxpy4  C += A*B, C full, A sparse, B bitmap/full (GB_AxB_saxpy4.c),
//   saxpy5  C += A*B, C full, A bitmap/full, B sparse (GB_AxB_saxpy5.c),
// both "+=" forms with the semiring's monoid as the accumulator and no mask.
//
// One kernel covers all three: a workgroup owns row i of C and a tile of
// DN_TW columns, keeps the tile's accumulators, presence flags and mask in
// LDS, and applies the row's terms A(i,k) B(k,j) for k ascending (the
// reference's sequential order, so fp64 sums are bit-identical to it):
//   B bitmap/full: each thread owns the tile's columns t, t+256, ... and
//                  reads B(k, those columns) directly, so no two threads
//                  touch one accumulator and no barrier is needed;
//   B sparse:      the threads take the entries of B(k,:) inside the tile
//                  (found by binary search), one barrier per k keeps the
//                  updates of one column in k order.
// ANY keeps the last writer (largest k), as the saxpy3 kernels do.

#include <cstring>
#include "gm_internal.h"
#include "gm_block.h"

namespace gm
{

constexpr int DN_NT = 256;
constexpr int DN_TW = 2048;                 // columns of C per tile

//------------------------------------------------------------------------------
// format conversions
//------------------------------------------------------------------------------

// present entries of each row (one workgroup per row, grid-stride)
__global__ __launch_bounds__(DN_NT) void k_dn_count(const int8_t *__restrict__ b, int64_t m, int64_t n,
                                                    int64_t *__restrict__ cnt)
{
    __shared__ int64_t s[DN_NT / 64 + 1];
    for (int64_t i = blockIdx.x; i < m; i += gridDim.x)
    {
        int64_t c = 0;
        if (b)
            for (int64_t j = threadIdx.x; j < n; j += DN_NT) c += b[i * n + j] != 0;
        else
            c = (threadIdx.x == 0) ? n : 0;
        const int64_t tot = block_sum<int64_t, DN_NT>(c, s);
        if (threadIdx.x == 0) cnt[i] = tot;
    }
}

// CSR of the present entries, in column order within each row
__global__ __launch_bounds__(DN_NT) void k_dn_fill(const int8_t *__restrict__ b, const uint8_t *__restrict__ x,
                                                   int64_t m, int64_t n, int32_t zs, int32_t iso,
                                                   const int64_t *__restrict__ Cp, int64_t *__restrict__ Cj,
                                                   uint8_t *__restrict__ Cx)
{
    __shared__ int64_t s[DN_NT / 64 + 1];
    for (int64_t i = blockIdx.x; i < m; i += gridDim.x)
    {
        int64_t base = Cp[i];
        for (int64_t j0 = 0; j0 < n; j0 += DN_NT)
        {
            const int64_t j = j0 + threadIdx.x;
            const int f = (j < n && (!b || b[i * n + j] != 0)) ? 1 : 0;
            int64_t tot;
            const int64_t pre = block_exclusive_scan<int64_t, DN_NT>(f, s, &tot);
            if (f)
            {
                const int64_t q = base + pre;
                Cj[q] = j;
                if (!iso)
                    for (int u = 0; u < zs; u++) Cx[q * zs + u] = x[(i * n + j) * zs + u];
            }
            base += tot;
        }
    }
}

// scatter a CSR matrix into bitmap/full arrays (b zeroed beforehand; b NULL
// for full, whose every entry the caller has checked to be present)
__global__ __launch_bounds__(DN_NT) void k_dn_scatter(const int64_t *__restrict__ Ap,
                                                      const int64_t *__restrict__ Aj,
                                                      const uint8_t *__restrict__ Ax, int64_t m, int64_t n,
                                                      int32_t zs, int32_t iso, int8_t *__restrict__ b,
                                                      uint8_t *__restrict__ x)
{
    for (int64_t i = blockIdx.x; i < m; i += gridDim.x)
    {
        for (int64_t p = Ap[i] + threadIdx.x; p < Ap[i + 1]; p += DN_NT)
        {
            const int64_t d = i * n + Aj[p];
            if (b) b[d] = 1;
            if (!iso)
                for (int u = 0; u < zs; u++) x[d * zs + u] = Ax[p * zs + u];
        }
    }
}

__global__ void k_dn_zero(uint8_t *x, int64_t n)
{
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t) gridDim.x * blockDim.x)
        x[t] = 0;
}

//------------------------------------------------------------------------------
// dense-result product
//------------------------------------------------------------------------------

struct DnArgs
{
    int64_t m, n, kdim;
    // A: sparse (Ap/Aj/Ax) or bitmap/full (Ab/Ax, Ab NULL for full)
    const int64_t *Ap, *Aj;
    const int8_t *Ab;
    const void *Ax;
    int32_t a_iso;
    const int64_t *Bp, *Bj;
    const int8_t *Bb;
    const void *Bx;
    int32_t b_iso;
    // mask: none (mkind 0), sparse (1) or bitmap/full (2)
    int32_t mkind, mcomp, mstruct, msize, m_iso;
    const int64_t *Mp, *Mj;
    const int8_t *Mb;
    const void *Mx;
    // C += A*B: Cin full values, else NULL
    const void *Cin;
    int8_t *Cb;                     // bitmap of C (NULL: full)
    void *Cx;
    unsigned long long *nvals;
};

__device__ __forceinline__ bool dn_mask_value(const void *Mx, int64_t p, int32_t msize, int32_t iso)
{
    return mask_true(Mx, p, msize, iso);
}

template <class S>
__device__ __forceinline__ void dn_apply(typename S::Z *acc, uint8_t *flag, int c, typename S::Z t)
{
    if (!flag[c]) { acc[c] = t; flag[c] = 1; }
    else acc[c] = S::upd(acc[c], t);
}

// ASP / BSP: A / B sparse (else bitmap or full)
template <class S, bool ASP, bool BSP>
__global__ __launch_bounds__(DN_NT) void k_dn_mxm(DnArgs g)
{
    using Z = typename S::Z;
    __shared__ Z s_acc[DN_TW];
    __shared__ uint8_t s_flag[DN_TW];
    __shared__ uint8_t s_mask[DN_TW];
    __shared__ int64_t s_cnt[DN_NT / 64 + 1];
    const Z *Ax = (const Z *) g.Ax;
    const Z *Bx = (const Z *) g.Bx;
    const int64_t ntiles = (g.n + DN_TW - 1) / DN_TW;
    const int64_t items = g.m * ntiles;
    for (int64_t it = blockIdx.x; it < items; it += gridDim.x)
    {
        const int64_t i = it / ntiles;
        const int64_t c0 = (it - i * ntiles) * DN_TW;
        const int w = (int) ((g.n - c0) < DN_TW ? (g.n - c0) : DN_TW);
        // tile state: nothing yet; the mask's allowed columns
        for (int c = threadIdx.x; c < DN_TW; c += DN_NT)
        {
            s_flag[c] = 0;
            uint8_t allow = 1;
            if (g.mkind == 2 && c < w)
            {
                const int64_t d = i * g.n + c0 + c;
                const bool present = !g.Mb || g.Mb[d] != 0;
                allow = (present && (g.mstruct || dn_mask_value(g.Mx, d, g.msize, g.m_iso))) ? 1 : 0;
            }
            else if (g.mkind == 1) allow = 0;
            s_mask[c] = allow;
        }
        __syncthreads();
        if (g.mkind == 1)
        {
            // the mask row's entries inside the tile
            const int64_t m0 = g.Mp[i], m1 = g.Mp[i + 1];
            const int64_t lo = lower_bound(g.Mj, m0, m1, c0);
            const int64_t hi = lower_bound(g.Mj, lo, m1, c0 + w);
            for (int64_t p = lo + threadIdx.x; p < hi; p += DN_NT)
                if (g.mstruct || dn_mask_value(g.Mx, p, g.msize, g.m_iso)) s_mask[g.Mj[p] - c0] = 1;
            __syncthreads();
        }
        if (g.mcomp)
        {
            for (int c = threadIdx.x; c < DN_TW; c += DN_NT) s_mask[c] = s_mask[c] ? 0 : 1;
            __syncthreads();
        }
        // the row's terms, k ascending
        int64_t pa = 0, pa_end = 0;
        if constexpr (ASP) { pa = g.Ap[i]; pa_end = g.Ap[i + 1]; }
        else { pa = 0; pa_end = g.kdim; }
        for (int64_t q = pa; q < pa_end; q++)
        {
            int64_t k;
            Z a;
            if constexpr (ASP)
            {
                k = g.Aj[q];
                a = S::geta(Ax, q, g.a_iso);
            }
            else
            {
                k = q;
                const int64_t d = i * g.kdim + k;
                if (g.Ab && g.Ab[d] == 0) continue;          // uniform across the workgroup
                a = S::geta(Ax, d, g.a_iso);
            }
            if constexpr (!BSP)
            {
                // B(k, tile) by owned columns: no barrier
                for (int c = threadIdx.x; c < w; c += DN_NT)
                {
                    const int64_t d = k * g.n + c0 + c;
                    if (g.Bb && g.Bb[d] == 0) continue;
                    if (!s_mask[c]) continue;
                    dn_apply<S>(s_acc, s_flag, c, S::mul(a, S::geta(Bx, d, g.b_iso), k));
                }
            }
            else
            {
                const int64_t b0 = g.Bp[k], b1 = g.Bp[k + 1];
                const int64_t lo = (b0 < b1 && g.Bj[b0] >= c0) ? b0 : lower_bound(g.Bj, b0, b1, c0);
                const int64_t hi = lower_bound(g.Bj, lo, b1, c0 + w);
                for (int64_t p = lo + threadIdx.x; p < hi; p += DN_NT)
                {
                    const int c = (int) (g.Bj[p] - c0);
                    if (!s_mask[c]) continue;
                    dn_apply<S>(s_acc, s_flag, c, S::mul(a, S::geta(Bx, p, g.b_iso), k));
                }
                __syncthreads();
            }
        }
        __syncthreads();
        // the tile of C
        int64_t cnt = 0;
        for (int c = threadIdx.x; c < w; c += DN_NT)
        {
            const int64_t d = i * g.n + c0 + c;
            const bool f = s_flag[c] != 0;
            if (g.Cin)
            {
                // C full += T with the monoid
                const Z cin = ((const Z *) g.Cin)[d];
                ((Z *) g.Cx)[d] = f ? S::upd(cin, s_acc[c]) : cin;
                cnt++;
            }
            else
            {
                g.Cb[d] = f ? 1 : 0;
                if constexpr (!S::iso_out) ((Z *) g.Cx)[d] = f ? s_acc[c] : (Z) 0;
                cnt += f ? 1 : 0;
            }
        }
        const int64_t tot = block_sum<int64_t, DN_NT>(cnt, s_cnt);
        if (threadIdx.x == 0 && tot) atomicAdd(g.nvals, (unsigned long long) tot);
        __syncthreads();
    }
}

__global__ void k_dn_zero_u64(unsigned long long *x) { if (threadIdx.x == 0 && blockIdx.x == 0) *x = 0; }
__global__ void k_dn_iso_one(void *x, int size)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        uint8_t *b = (uint8_t *) x;
        for (int t = 0; t < size; t++) b[t] = 0;
        b[0] = 1;
    }
}

static unsigned dn_grid(int64_t items)
{
    const int64_t cap = (int64_t) cu_count() * 8;
    int64_t g = items < cap ? items : cap;
    return (unsigned) (g < 1 ? 1 : g);
}

template <class S>
static int dn_mxm_run(gm_dense_out *C, const gm_csr *Ms, const gm_dense *Md, bool mask_comp,
                      bool mask_struct, const gm_csr *As, const gm_dense *Ad, const gm_csr *Bs,
                      const gm_dense *Bd, const gm_dense *Cin)
{
    using Z = typename S::Z;
    hipStream_t st = stream();
    DnArgs g;
    memset(&g, 0, sizeof(g));
    g.m = As ? As->nrows : Ad->nrows;
    g.kdim = As ? As->ncols : Ad->ncols;
    g.n = Bs ? Bs->ncols : Bd->ncols;
    if (As) { g.Ap = As->p; g.Aj = As->j; g.Ax = As->x; g.a_iso = As->iso; }
    else { g.Ab = Ad->b; g.Ax = Ad->x; g.a_iso = Ad->iso; }
    if (Bs) { g.Bp = Bs->p; g.Bj = Bs->j; g.Bx = Bs->x; g.b_iso = Bs->iso; }
    else { g.Bb = Bd->b; g.Bx = Bd->x; g.b_iso = Bd->iso; }
    g.mcomp = mask_comp ? 1 : 0;
    if (Ms)
    {
        g.mkind = 1; g.Mp = Ms->p; g.Mj = Ms->j; g.Mx = Ms->x;
        g.msize = (int32_t) type_size(Ms->type); g.m_iso = Ms->iso;
        g.mstruct = (mask_struct || !Ms->x) ? 1 : 0;
    }
    else if (Md)
    {
        g.mkind = 2; g.Mb = Md->b; g.Mx = Md->x;
        g.msize = (int32_t) type_size(Md->type); g.m_iso = Md->iso;
        g.mstruct = (mask_struct || !Md->x) ? 1 : 0;
    }
    const int64_t mn = g.m * g.n;
    const bool accum = (Cin != nullptr);
    DevBuf Cb, Cx, cnt;
    if ((!accum && !Cb.alloc(mn > 0 ? mn : 1)) ||
        !Cx.alloc(S::iso_out && !accum ? sizeof(Z) : sizeof(Z) * (mn > 0 ? mn : 1)) ||
        !cnt.alloc(sizeof(unsigned long long)))
        return set_error(GM_OUT_OF_MEMORY, "dense product: output");
    g.Cin = accum ? Cin->x : nullptr;
    g.Cb = accum ? nullptr : Cb.as<int8_t>();
    g.Cx = Cx.ptr;
    g.nvals = cnt.as<unsigned long long>();
    k_dn_zero_u64<<<1, 64, 0, st>>>(g.nvals);
    if (mn > 0)
    {
        const int64_t items = g.m * ((g.n + DN_TW - 1) / DN_TW);
        ProfScope ps(accum ? (As ? "saxpy4" : "saxpy5") : "saxbit");
        if (As && Bs) k_dn_mxm<S, true, true><<<dn_grid(items), DN_NT, 0, st>>>(g);
        else if (As) k_dn_mxm<S, true, false><<<dn_grid(items), DN_NT, 0, st>>>(g);
        else if (Bs) k_dn_mxm<S, false, true><<<dn_grid(items), DN_NT, 0, st>>>(g);
        else k_dn_mxm<S, false, false><<<dn_grid(items), DN_NT, 0, st>>>(g);
        GM_HIP(hipGetLastError(), "k_dn_mxm");
    }
    if (S::iso_out && !accum)
    {
        k_dn_iso_one<<<1, 1, 0, st>>>(Cx.ptr, (int) sizeof(Z));
        GM_HIP(hipGetLastError(), "k_dn_iso_one");
    }
    unsigned long long *hp = (unsigned long long *) pinned_scratch(sizeof(unsigned long long));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(hp, g.nvals, sizeof(unsigned long long), hipMemcpyDeviceToHost, st), "nvals");
    GM_HIP(hipStreamSynchronize(st), "dense product sync");
    C->nrows = g.m;
    C->ncols = g.n;
    C->nvals = (int64_t) *hp;
    C->b = accum ? nullptr : (int8_t *) Cb.release();
    C->x = Cx.release();
    C->iso = (S::iso_out && !accum) ? 1 : 0;
    C->type = 0;
    return GM_SUCCESS;
}

static bool dn_ok(const gm_dense *D)
{
    return D && D->nrows >= 0 && D->ncols >= 0 && (D->x || D->nrows * D->ncols == 0);
}

} // namespace gm

using namespace gm;

extern "C" int gm_dense_to_csr(gm_csr_out *R, const gm_dense *D)
{
    if (!R || !dn_ok(D)) return set_error(GM_INVALID_VALUE, "gm_dense_to_csr: bad argument");
    GM_TRY(ensure_init());
    hipStream_t st = stream();
    const int64_t m = D->nrows, n = D->ncols;
    const size_t zs = type_size(D->type);
    if (zs == 0) return set_error(GM_INVALID_VALUE, "gm_dense_to_csr: type");
    DevBuf cnt, Cp, Cj, Cx;
    if (!cnt.alloc(sizeof(int64_t) * (m > 0 ? m : 1)) || !Cp.alloc(sizeof(int64_t) * (m + 1)))
        return set_error(GM_OUT_OF_MEMORY, "gm_dense_to_csr: workspace");
    if (m > 0)
    {
        k_dn_count<<<dn_grid(m), DN_NT, 0, st>>>(D->b, m, n, cnt.as<int64_t>());
        GM_HIP(hipGetLastError(), "k_dn_count");
    }
    GM_TRY(scan_exclusive_i64(cnt.as<int64_t>(), Cp.as<int64_t>(), m));
    int64_t *hp = (int64_t *) pinned_scratch(sizeof(int64_t));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(hp, Cp.as<int64_t>() + m, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz");
    GM_HIP(hipStreamSynchronize(st), "gm_dense_to_csr sync");
    const int64_t nz = *hp;
    if (!Cj.alloc(sizeof(int64_t) * (nz > 0 ? nz : 1)) ||
        !Cx.alloc(D->iso ? zs : zs * (size_t) (nz > 0 ? nz : 1)))
        return set_error(GM_OUT_OF_MEMORY, "gm_dense_to_csr: output");
    if (m > 0 && nz > 0)
    {
        k_dn_fill<<<dn_grid(m), DN_NT, 0, st>>>(D->b, (const uint8_t *) D->x, m, n, (int32_t) zs, D->iso,
                                                Cp.as<int64_t>(), Cj.as<int64_t>(), Cx.as<uint8_t>());
        GM_HIP(hipGetLastError(), "k_dn_fill");
    }
    if (D->iso) GM_HIP(hipMemcpyAsync(Cx.ptr, D->x, zs, hipMemcpyDeviceToDevice, st), "iso value");
    GM_HIP(hipStreamSynchronize(st), "gm_dense_to_csr sync");
    R->nrows = m;
    R->ncols = n;
    R->nvals = nz;
    R->p = (int64_t *) Cp.release();
    R->j = (int64_t *) Cj.release();
    R->x = Cx.release();
    R->iso = D->iso;
    R->type = D->type;
    return GM_SUCCESS;
}

extern "C" int gm_csr_to_dense(gm_dense_out *D, const gm_csr *A, bool full)
{
    if (!D || !A || !A->p) return set_error(GM_INVALID_VALUE, "gm_csr_to_dense: bad argument");
    GM_TRY(ensure_init());
    hipStream_t st = stream();
    const int64_t m = A->nrows, n = A->ncols, mn = m * n;
    const size_t zs = type_size(A->type);
    if (zs == 0) return set_error(GM_INVALID_VALUE, "gm_csr_to_dense: type");
    int64_t *hp = (int64_t *) pinned_scratch(sizeof(int64_t));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(hp, A->p + m, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz");
    GM_HIP(hipStreamSynchronize(st), "gm_csr_to_dense sync");
    const int64_t nz = *hp;
    if (full && nz != mn) return set_error(GM_INVALID_VALUE, "gm_csr_to_dense: not every entry is present");
    if (nz > 0 && !A->x && !A->iso) return set_error(GM_INVALID_VALUE, "gm_csr_to_dense: no values");
    DevBuf b, x;
    if ((!full && !b.alloc(mn > 0 ? mn : 1)) || !x.alloc(A->iso ? zs : zs * (size_t) (mn > 0 ? mn : 1)))
        return set_error(GM_OUT_OF_MEMORY, "gm_csr_to_dense: output");
    if (!full && mn > 0)
    {
        k_dn_zero<<<dn_grid((mn + 255) / 256), 256, 0, st>>>(b.as<uint8_t>(), mn);
        GM_HIP(hipGetLastError(), "k_dn_zero b");
    }
    if (!A->iso && mn > 0)
    {
        k_dn_zero<<<dn_grid((mn * (int64_t) zs + 255) / 256), 256, 0, st>>>(x.as<uint8_t>(), mn * (int64_t) zs);
        GM_HIP(hipGetLastError(), "k_dn_zero x");
    }
    if (m > 0 && nz > 0)
    {
        k_dn_scatter<<<dn_grid(m), DN_NT, 0, st>>>(A->p, A->j, (const uint8_t *) A->x, m, n, (int32_t) zs,
                                                   A->iso, full ? nullptr : b.as<int8_t>(), x.as<uint8_t>());
        GM_HIP(hipGetLastError(), "k_dn_scatter");
    }
    if (A->iso) GM_HIP(hipMemcpyAsync(x.ptr, A->x, zs, hipMemcpyDeviceToDevice, st), "iso value");
    GM_HIP(hipStreamSynchronize(st), "gm_csr_to_dense sync");
    D->nrows = m;
    D->ncols = n;
    D->nvals = nz;
    D->b = full ? nullptr : (int8_t *) b.release();
    D->x = x.release();
    D->iso = A->iso;
    D->type = A->type;
    return GM_SUCCESS;
}

extern "C" int gm_saxbit(gm_dense_out *C, const gm_csr *Ms, const gm_dense *Md, bool mask_comp,
                         bool mask_struct, const gm_csr *As, const gm_dense *Ad, const gm_csr *Bs,
                         const gm_dense *Bd, const gm_semiring *sr, const gm_dense *Cin)
{
    if (!C || !sr || (!As == !Ad) || (!Bs == !Bd) || (Ms && Md))
        return set_error(GM_INVALID_VALUE, "gm_saxbit: one of each operand pair must be given");
    if (!semiring_supported(sr)) return set_error(GM_NO_VALUE, "gm_saxbit: semiring not handled");
    const int64_t m = As ? As->nrows : Ad->nrows, kd = As ? As->ncols : Ad->ncols;
    const int64_t kb = Bs ? Bs->nrows : Bd->nrows, n = Bs ? Bs->ncols : Bd->ncols;
    if (kd != kb) return set_error(GM_INVALID_VALUE, "gm_saxbit: inner dimensions differ");
    if ((Ms && (Ms->nrows != m || Ms->ncols != n)) || (Md && (Md->nrows != m || Md->ncols != n)))
        return set_error(GM_INVALID_VALUE, "gm_saxbit: mask dimensions");
    if (Cin && (Cin->nrows != m || Cin->ncols != n || Cin->b || Cin->iso || Cin->type != sr->ztype ||
                Ms || Md || sr->add_op == GM_ADD_ANY))
        return set_error(GM_NO_VALUE, "gm_saxbit: C += A*B needs a full, non-iso C of the monoid's "
                                      "type, no mask and a monoid other than ANY");
    if (!sr || (sr->mult_op != GM_MULT_PAIR && sr->mult_op != GM_MULT_K))
    {
        const int32_t at = As ? As->type : Ad->type, bt = Bs ? Bs->type : Bd->type;
        const void *ax = As ? As->x : Ad->x, *bx = Bs ? Bs->x : Bd->x;
        if (at != sr->ztype || bt != sr->ztype || !ax || !bx)
            return set_error(GM_NO_VALUE, "gm_saxbit: operand types must be the semiring's");
    }
    GM_TRY(ensure_init());
    int rc = GM_NO_VALUE;
#define GM_DN_CASE(ADD, MULT, ZC, Z)                                                        \
    if (sr->add_op == ADD && sr->mult_op == MULT && sr->ztype == ZC)                        \
    {                                                                                       \
        rc = dn_mxm_run<SR<ADD, MULT, Z>>(C, Ms, Md, mask_comp, mask_struct, As, Ad, Bs, Bd, \
                                          Cin);                                             \
        if (rc == GM_SUCCESS) C->type = ZC;                                                 \
        return rc;                                                                          \
    }
    GM_SEMIRING_LIST(GM_DN_CASE)
#undef GM_DN_CASE
    return rc;
}
