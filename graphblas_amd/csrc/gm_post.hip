// gm_post.hip: the steps around the product kernels that GrB_mxm needs for
// its full contract, kept on the device:
//
//   gm_transpose    CSR of A' from the CSR of A (the reference's
//                   GB_transpose, Source/transpose/GB_transpose.c, used when
//                   the orientation rules of GB_AxB_meta.c:683-694,718-723
//                   leave an explicit transpose).  Entries are keyed
//                   (column << 32 | row) and stably radix-sorted (rocPRIM)
//                   on the column bits with their positions: the entries
//                   enter in row order, so the rows of A' come out sorted; row
//                   pointers by binary search of each row's first key, then
//                   a gather of the values.
//   gm_accum_mask   C<M, replace> = accum (C, T): the reference's
//                   GB_accum_mask (Source/mask/GB_accum_mask.c) for sparse
//                   C, M and T of one type.  One thread per row merges the
//                   sorted rows of C and T and walks the sorted mask row:
//                   count pass, exclusive scan of the counts, fill pass.

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include "gm_internal.h"

namespace gm
{

static inline unsigned post_grid(int64_t n, int nt, int64_t cap)
{
    int64_t g = (n + nt - 1) / nt;
    if (g > cap) g = cap;
    return (unsigned) (g > 0 ? g : 1);
}

__global__ void k_post_zero(int64_t *x, int64_t n)
{
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) x[t] = 0;
}

//------------------------------------------------------------------------------
// transpose
//------------------------------------------------------------------------------

// key of entry p: (column << 32) | row; pos[p] = p.  A workgroup takes
// TR_TILE consecutive entries: the rows holding its first and last entry are
// found once, and each entry searches only between them.
constexpr int TR_TILE = 256 * 16;

__device__ __forceinline__ int64_t tr_row_of(const int64_t *__restrict__ Ap, int64_t lo, int64_t hi, int64_t p)
{
    // last i in [lo, hi] with Ap[i] <= p
    while (lo < hi)
    {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (Ap[mid] <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_tr_keys(const int64_t *__restrict__ Ap, const int64_t *__restrict__ Aj, int64_t m,
                          int64_t nnz, uint64_t *__restrict__ keys, int64_t *__restrict__ pos)
{
    __shared__ int64_t s_r[2];
    for (int64_t t0 = (int64_t) blockIdx.x * TR_TILE; t0 < nnz; t0 += (int64_t) gridDim.x * TR_TILE)
    {
        const int64_t t1 = (t0 + TR_TILE < nnz) ? t0 + TR_TILE : nnz;
        if (threadIdx.x < 2) s_r[threadIdx.x] = tr_row_of(Ap, 0, m - 1, threadIdx.x == 0 ? t0 : t1 - 1);
        __syncthreads();
        const int64_t r0 = s_r[0], r1 = s_r[1];
        __syncthreads();
        for (int64_t p = t0 + threadIdx.x; p < t1; p += blockDim.x)
        {
            const int64_t i = tr_row_of(Ap, r0, r1, p);
            keys[p] = ((uint64_t) Aj[p] << 32) | (uint64_t) i;
            pos[p] = p;
        }
    }
}

__device__ __forceinline__ int64_t lb_u64(const uint64_t *a, int64_t n, uint64_t v)
{
    int64_t lo = 0, hi = n;
    while (lo < hi)
    {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// T.p[r] = first sorted key of column r; T.j = row; T.x = A.x[pos]
__global__ void k_tr_rows(const uint64_t *__restrict__ keys, int64_t nnz, int64_t n,
                          int64_t *__restrict__ Tp)
{
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    for (int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r <= n; r += stride)
        Tp[r] = (r == n) ? nnz : lb_u64(keys, nnz, (uint64_t) r << 32);
}

__global__ void k_tr_fill(const uint64_t *__restrict__ keys, const int64_t *__restrict__ pos,
                          int64_t nnz, const uint8_t *__restrict__ Ax, int32_t xsize,
                          int64_t *__restrict__ Tj, uint8_t *__restrict__ Tx)
{
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    for (int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += stride)
    {
        Tj[q] = (int64_t) (keys[q] & 0xffffffffull);
        if (Tx)
        {
            const int64_t p = pos[q];
            for (int b = 0; b < xsize; b++) Tx[q * xsize + b] = Ax[p * xsize + b];
        }
    }
}

//------------------------------------------------------------------------------
// accum + mask
//------------------------------------------------------------------------------

template <class Z>
__device__ __forceinline__ Z accum_apply(int op, Z c, Z t)
{
    if constexpr (std::is_same<Z, double>::value)
    {
        switch (op)
        {
            case 1: return __dadd_rn(c, t);
            case 2: return __dsub_rn(c, t);
            case 3: return __dmul_rn(c, t);
            case 4: return fmin(c, t);
            case 5: return fmax(c, t);
            case 6: return c;
            case 9: return 1.0;
            case 10: return (c != 0.0 || t != 0.0) ? 1.0 : 0.0;
            case 11: return (c != 0.0 && t != 0.0) ? 1.0 : 0.0;
            default: return t;          // SECOND, ANY
        }
    }
    else if constexpr (std::is_same<Z, uint8_t>::value)
    {
        // bool: PLUS and MAX and LOR are "or"; TIMES, MIN and LAND are "and";
        // MINUS is "xor" (GrB's bool arithmetic)
        const bool a = c != 0, b = t != 0;
        switch (op)
        {
            case 1: case 5: case 10: return (uint8_t) (a || b);
            case 2: return (uint8_t) (a != b);
            case 3: case 4: case 11: return (uint8_t) (a && b);
            case 6: return (uint8_t) a;
            case 9: return 1;
            default: return (uint8_t) b;
        }
    }
    else
    {
        using U = typename std::make_unsigned<Z>::type;   // wrap-around arithmetic
        switch (op)
        {
            case 1: return (Z) ((U) c + (U) t);
            case 2: return (Z) ((U) c - (U) t);
            case 3: return (Z) ((U) c * (U) t);
            case 4: return c < t ? c : t;
            case 5: return c > t ? c : t;
            case 6: return c;
            case 9: return 1;
            case 10: return (Z) ((c != 0) || (t != 0));
            case 11: return (Z) ((c != 0) && (t != 0));
            default: return t;
        }
    }
}

struct AccArgs
{
    int64_t m;
    const int64_t *Cp, *Cj; const void *Cx; int32_t c_iso;
    const int64_t *Tp, *Tj; const void *Tx; int32_t t_iso;
    const int64_t *Mp, *Mj; const void *Mx; int32_t m_iso, msize;
    int32_t has_mask, mask_comp, mask_struct, replace, op;
    int64_t *Op;          // counts (pass 0), then offsets (pass 1)
    int64_t *Oj;
    void *Ox;
};

template <class Z, int PASS>
__global__ void k_accum_mask(AccArgs g)
{
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    const Z *Cx = (const Z *) g.Cx, *Tx = (const Z *) g.Tx;
    Z *Ox = (Z *) g.Ox;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < g.m; i += stride)
    {
        int64_t pc = g.Cp ? g.Cp[i] : 0, pc1 = g.Cp ? g.Cp[i + 1] : 0;
        int64_t pt = g.Tp[i], pt1 = g.Tp[i + 1];
        int64_t pm = g.has_mask ? g.Mp[i] : 0, pm1 = g.has_mask ? g.Mp[i + 1] : 0;
        int64_t out = PASS == 0 ? 0 : g.Op[i];
        while (pc < pc1 || pt < pt1)
        {
            const int64_t jc = pc < pc1 ? g.Cj[pc] : INT64_MAX;
            const int64_t jt = pt < pt1 ? g.Tj[pt] : INT64_MAX;
            const int64_t j = jc < jt ? jc : jt;
            const bool cin = (jc == j), tin = (jt == j);
            bool allow = true;
            if (g.has_mask)
            {
                while (pm < pm1 && g.Mj[pm] < j) pm++;
                bool mij = (pm < pm1 && g.Mj[pm] == j) &&
                           (g.mask_struct || mask_true(g.Mx, pm, g.msize, g.m_iso));
                allow = g.mask_comp ? !mij : mij;
            }
            bool present;
            Z z = Z(0);
            if (allow)
            {
                if (g.op)
                {
                    present = cin || tin;
                    if (PASS == 1 && present)
                    {
                        const Z cv = cin ? Cx[g.c_iso ? 0 : pc] : Z(0);
                        const Z tv = tin ? Tx[g.t_iso ? 0 : pt] : Z(0);
                        z = (cin && tin) ? accum_apply<Z>(g.op, cv, tv) : (cin ? cv : tv);
                    }
                }
                else
                {
                    present = tin;
                    if (PASS == 1 && present) z = Tx[g.t_iso ? 0 : pt];
                }
            }
            else
            {
                present = cin && !g.replace;
                if (PASS == 1 && present) z = Cx[g.c_iso ? 0 : pc];
            }
            if (present)
            {
                if (PASS == 1)
                {
                    g.Oj[out] = j;
                    Ox[out] = z;
                }
                out++;
            }
            if (cin) pc++;
            if (tin) pt++;
        }
        if (PASS == 0) g.Op[i] = out;
    }
}

template <class Z>
static int accum_mask_run(gm_csr_out *O, const gm_csr *C, const gm_csr *M, bool mask_comp,
                          bool mask_struct, bool replace, const gm_csr *T, int32_t op)
{
    hipStream_t st = stream();
    const int64_t m = T->nrows;
    AccArgs g;
    memset(&g, 0, sizeof(g));
    g.m = m;
    if (C && C->p) { g.Cp = C->p; g.Cj = C->j; g.Cx = C->x; g.c_iso = C->iso; }
    g.Tp = T->p; g.Tj = T->j; g.Tx = T->x; g.t_iso = T->iso;
    if (M && M->p)
    {
        g.has_mask = 1;
        g.Mp = M->p; g.Mj = M->j; g.Mx = mask_struct ? nullptr : M->x;
        g.m_iso = M->iso; g.msize = (int32_t) type_size(M->type);
    }
    g.mask_comp = mask_comp ? 1 : 0;
    g.mask_struct = (mask_struct || !M || !M->x) ? 1 : 0;
    g.replace = replace ? 1 : 0;
    g.op = op;
    DevBuf Op;
    if (!Op.alloc(sizeof(int64_t) * (m + 1))) return set_error(GM_OUT_OF_MEMORY, "accum: row counts");
    g.Op = Op.as<int64_t>();
    const unsigned grid = post_grid(m, 256, (int64_t) cu_count() * 16);
    {
        ProfScope ps("accum_mask_count");
        k_accum_mask<Z, 0><<<grid, 256, 0, st>>>(g);
        GM_HIP(hipGetLastError(), "k_accum_mask count");
    }
    GM_TRY(scan_exclusive_i64(g.Op, g.Op, m));
    int64_t *hp = (int64_t *) pinned_scratch(sizeof(int64_t));
    if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    GM_HIP(hipMemcpyAsync(hp, g.Op + m, sizeof(int64_t), hipMemcpyDeviceToHost, st), "accum nnz");
    GM_HIP(hipStreamSynchronize(st), "accum nnz sync");
    const int64_t nz = *hp;
    DevBuf Oj, Ox;
    if (!Oj.alloc(sizeof(int64_t) * (nz > 0 ? nz : 1)) || !Ox.alloc(sizeof(Z) * (nz > 0 ? nz : 1)))
        return set_error(GM_OUT_OF_MEMORY, "accum: output (%lld entries)", (long long) nz);
    g.Oj = Oj.as<int64_t>();
    g.Ox = Ox.ptr;
    if (nz > 0)
    {
        ProfScope ps("accum_mask_fill");
        k_accum_mask<Z, 1><<<grid, 256, 0, st>>>(g);
        GM_HIP(hipGetLastError(), "k_accum_mask fill");
    }
    GM_HIP(hipStreamSynchronize(st), "accum sync");
    O->nrows = m;
    O->ncols = T->ncols;
    O->nvals = nz;
    O->p = (int64_t *) Op.release();
    O->j = (int64_t *) Oj.release();
    O->x = Ox.release();
    O->iso = 0;
    O->type = T->type;
    return GM_SUCCESS;
}

} // namespace gm

using namespace gm;

extern "C" int gm_transpose(gm_csr_out *T, const gm_csr *A)
{
    if (!T || !A || !A->p) return set_error(GM_INVALID_VALUE, "gm_transpose: NULL argument");
    if (A->nrows >= ((int64_t) 1 << 32) || A->ncols >= ((int64_t) 1 << 32))
        return set_error(GM_NO_VALUE, "gm_transpose: dimensions beyond 2^32");
    GM_TRY(ensure_init());
    hipStream_t st = stream();
    const int64_t m = A->nrows, n = A->ncols;
    int64_t nnz = 0;
    {
        int64_t *hp = (int64_t *) pinned_scratch(sizeof(int64_t));
        if (!hp) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
        GM_HIP(hipMemcpyAsync(hp, A->p + m, sizeof(int64_t), hipMemcpyDeviceToHost, st), "nnz(A)");
        GM_HIP(hipStreamSynchronize(st), "nnz(A) sync");
        nnz = *hp;
    }
    const size_t xs = A->x ? type_size(A->type) : 0;
    const bool iso = A->iso != 0;
    DevBuf Tp, Tj, Tx, k0, k1, p0, p1, tmp;
    if (!Tp.alloc(sizeof(int64_t) * (n + 1)) || !Tj.alloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1)) ||
        (xs && !Tx.alloc(iso ? xs : xs * (nnz > 0 ? nnz : 1))))
        return set_error(GM_OUT_OF_MEMORY, "gm_transpose: output");
    if (nnz > 0)
    {
        if (!k0.alloc(8 * nnz) || !k1.alloc(8 * nnz) || !p0.alloc(8 * nnz) || !p1.alloc(8 * nnz))
            return set_error(GM_OUT_OF_MEMORY, "gm_transpose: sort buffers");
        const unsigned grid = post_grid(nnz, 256, (int64_t) cu_count() * 16);
        {
            ProfScope ps("transpose_keys");
            k_tr_keys<<<post_grid(nnz, TR_TILE, (int64_t) cu_count() * 16), 256, 0, st>>>(
                A->p, A->j, m, nnz, k0.as<uint64_t>(), p0.as<int64_t>());
            GM_HIP(hipGetLastError(), "k_tr_keys");
        }
        // the keys enter in row order and the radix sort is stable, so
        // sorting by the column bits alone leaves each column's rows ascending
        int cbits = 1;
        while (cbits < 32 && ((int64_t) 1 << cbits) < n) cbits++;
        const unsigned begin_bit = 32u, end_bit = 32u + (unsigned) cbits;
        size_t tb = 0;
        GM_HIP(rocprim::radix_sort_pairs(nullptr, tb, k0.as<uint64_t>(), k1.as<uint64_t>(),
                                         p0.as<int64_t>(), p1.as<int64_t>(), (size_t) nnz, begin_bit, end_bit,
                                         st), "transpose sort size");
        if (!tmp.alloc(tb)) return set_error(GM_OUT_OF_MEMORY, "gm_transpose: sort storage");
        {
            ProfScope ps("transpose_sort");
            GM_HIP(rocprim::radix_sort_pairs(tmp.ptr, tb, k0.as<uint64_t>(), k1.as<uint64_t>(),
                                             p0.as<int64_t>(), p1.as<int64_t>(), (size_t) nnz, begin_bit,
                                             end_bit, st), "transpose sort");
        }
        {
            ProfScope ps("transpose_fill");
            k_tr_rows<<<post_grid(n + 1, 256, (int64_t) cu_count() * 16), 256, 0, st>>>(
                k1.as<uint64_t>(), nnz, n, Tp.as<int64_t>());
            GM_HIP(hipGetLastError(), "k_tr_rows");
            k_tr_fill<<<grid, 256, 0, st>>>(k1.as<uint64_t>(), p1.as<int64_t>(), nnz,
                                            (const uint8_t *) A->x, (int32_t) xs, Tj.as<int64_t>(),
                                            (xs && !iso) ? Tx.as<uint8_t>() : nullptr);
            GM_HIP(hipGetLastError(), "k_tr_fill");
        }
    }
    else
    {
        k_post_zero<<<post_grid(n + 1, 256, 1024), 256, 0, st>>>(Tp.as<int64_t>(), n + 1);
        GM_HIP(hipGetLastError(), "k_post_zero");
    }
    if (xs && iso) GM_HIP(hipMemcpyAsync(Tx.ptr, A->x, xs, hipMemcpyDeviceToDevice, st), "iso value");
    GM_HIP(hipStreamSynchronize(st), "transpose sync");
    T->nrows = n;
    T->ncols = m;
    T->nvals = nnz;
    T->p = (int64_t *) Tp.release();
    T->j = (int64_t *) Tj.release();
    T->x = xs ? Tx.release() : nullptr;
    T->iso = iso ? 1 : 0;
    T->type = A->type;
    return GM_SUCCESS;
}

extern "C" int gm_accum_mask(gm_csr_out *O, const gm_csr *C, const gm_csr *M, bool mask_comp,
                             bool mask_struct, bool replace, const gm_csr *T, int32_t accum_op)
{
    if (!O || !T || !T->p) return set_error(GM_INVALID_VALUE, "gm_accum_mask: NULL argument");
    if (C && C->p && (C->nrows != T->nrows || C->ncols != T->ncols || C->type != T->type))
        return set_error(GM_INVALID_VALUE, "gm_accum_mask: C and T differ in shape or type");
    if (M && M->p && (M->nrows != T->nrows || M->ncols != T->ncols))
        return set_error(GM_INVALID_VALUE, "gm_accum_mask: mask shape");
    if (accum_op < 0 || accum_op > 11)
        return set_error(GM_INVALID_VALUE, "gm_accum_mask: accumulator %d", accum_op);
    if ((C && C->p && !C->x) || !T->x)
        return set_error(GM_INVALID_VALUE, "gm_accum_mask: C and T need values");
    GM_TRY(ensure_init());
    switch (T->type)
    {
        case GM_BOOL: return accum_mask_run<uint8_t>(O, C, M, mask_comp, mask_struct, replace, T, accum_op);
        case GM_INT32: return accum_mask_run<int32_t>(O, C, M, mask_comp, mask_struct, replace, T, accum_op);
        case GM_INT64: return accum_mask_run<int64_t>(O, C, M, mask_comp, mask_struct, replace, T, accum_op);
        case GM_FP64: return accum_mask_run<double>(O, C, M, mask_comp, mask_struct, replace, T, accum_op);
        default: return set_error(GM_NO_VALUE, "gm_accum_mask: type %d not supported", T->type);
    }
}
