// gm_internal.h: shared declarations of the gfx950 kernel layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <type_traits>
#include <vector>
#include "gm_kernels.h"

// derived forms of a CSR operand (built by gm_csr_aux_build, gm_saxpy3.hip)
struct gm_csr_aux
{
    int64_t nrows;
    int32_t rec_lg;     // row records: row k at k << rec_lg (-1: none)
    void *rec;          // nrows << rec_lg records of 16 bytes (column, value bits)
};

//------------------------------------------------------------------------------
// runtime (gm_runtime.hip)
//------------------------------------------------------------------------------

namespace gm
{

hipStream_t stream();                       // the library stream
int ensure_init();                          // gm_init() result, cached
bool is_initialised();                      // gm_init() has succeeded
int reduce_scratch_setup();                 // device address of the reduction scratch (gm_reduce.hip)
int set_error(int code, const char *fmt, ...);
int check_hip(hipError_t e, const char *what);
int cu_count();
int64_t device_hbm_bytes();                // total HBM of the current device

// timed launch bracket: when profiling is on, events are recorded around the
// launch on the library stream and attributed to `name`
struct ProfScope
{
    int slot;
    explicit ProfScope(const char *name);
    ~ProfScope();
};

// device scratch owned by one call; released on scope exit
struct DevBuf
{
    void *ptr = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (ptr) gm_free(ptr); }
    bool alloc(size_t bytes) { ptr = gm_malloc(bytes ? bytes : 8); return ptr != nullptr; }
    template <class T> T *as() const { return (T *) ptr; }
    void *release() { void *p = ptr; ptr = nullptr; return p; }
};

// pinned host staging for small device->host reads
void *pinned_scratch(size_t bytes);

// One piece of a stacked operand: rows [0, nrows) of the CSR (p, j, x), whose
// row pointers may start anywhere (a slice of a larger matrix), placed at
// stacked row row_base with col_add added to every column index.
struct BatchPart
{
    const int64_t *p, *j;
    const void *x;
    int64_t nrows;
    int64_t row_base;       // first stacked row
    int64_t col_add;        // added to every column index
    int64_t ent_base;       // first stacked entry (filled on the device)
    int32_t iso;
    int32_t pad;
};

// Stacks the parts into one CSR of total_rows rows on the library stream
// (P, J and X receive the arrays; values of `type` when `values`, iso parts
// expanded).  One device-to-host read of the total entry count.
int bt_stack(const std::vector<BatchPart> &hparts, int64_t total_rows, int64_t ncols, bool values,
             int32_t type, DevBuf &dparts, DevBuf &P, DevBuf &J, DevBuf &X, gm_csr *out);

// exclusive scan of n int64 values in d_in into d_out (d_out[n] = total);
// d_in and d_out may alias.  Uses the library stream.
int scan_exclusive_i64(const int64_t *d_in, int64_t *d_out, int64_t n);

size_t type_size(int32_t type);

// process-wide tuned values (gm_dot3_set_defaults / gm_saxpy3_set_defaults /
// gm_load_tuning); a 0 field means "built-in value"
gm_dot3_params dot3_tuned();
gm_saxpy3_params saxpy3_tuned();

// explicit value if set (> 0), else the tuned one if set, else the built-in
inline int64_t pick_param(int64_t explicit_v, int64_t tuned_v, int64_t builtin)
{
    return explicit_v > 0 ? explicit_v : (tuned_v > 0 ? tuned_v : builtin);
}

} // namespace gm

#define GM_TRY(call)                     \
    do {                                 \
        int gm_rc_ = (call);             \
        if (gm_rc_ != GM_SUCCESS) return gm_rc_; \
    } while (0)

#define GM_HIP(call, what) GM_TRY(gm::check_hip((call), (what)))

//------------------------------------------------------------------------------
// device helpers
//------------------------------------------------------------------------------

#define GM_WAVE 64

namespace gm
{

// mask value test: a valued mask entry is true iff its bytes are not all zero
// (the reference's GB_MCAST, Source/mask/include/GB_mcast.h)
__device__ __forceinline__ bool mask_true(const void *Mx, int64_t p, int32_t msize, int32_t iso)
{
    if (Mx == nullptr) return true;
    const int64_t q = iso ? 0 : p;
    switch (msize)
    {
        case 1: return ((const uint8_t *) Mx)[q] != 0;
        case 2: return ((const uint16_t *) Mx)[q] != 0;
        case 4: return ((const uint32_t *) Mx)[q] != 0;
        case 8: return ((const uint64_t *) Mx)[q] != 0;
        default:
        {
            const uint8_t *b = (const uint8_t *) Mx + q * msize;
            for (int t = 0; t < msize; t++) if (b[t]) return true;
            return false;
        }
    }
}

__device__ __forceinline__ int ilog2_ceil(int64_t x)
{
    // ceil(log2(x)) for x >= 1
    return (x <= 1) ? 0 : 64 - __clzll((unsigned long long) (x - 1));
}

// first position q in [lo, hi) with a[q] >= key (hi if none)
__device__ __forceinline__ int64_t lower_bound(const int64_t *a, int64_t lo, int64_t hi, int64_t key)
{
    while (lo < hi)
    {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

//------------------------------------------------------------------------------
// semiring traits
//------------------------------------------------------------------------------

template <class Z> struct Lim;
template <> struct Lim<int32_t>
{
    static __device__ __forceinline__ int32_t min() { return INT32_MIN; }
    static __device__ __forceinline__ int32_t max() { return INT32_MAX; }
};
template <> struct Lim<int64_t>
{
    static __device__ __forceinline__ int64_t min() { return INT64_MIN; }
    static __device__ __forceinline__ int64_t max() { return INT64_MAX; }
};
template <> struct Lim<double>
{
    static __device__ __forceinline__ double min() { return -__builtin_inf(); }
    static __device__ __forceinline__ double max() { return __builtin_inf(); }
};
template <> struct Lim<uint8_t>
{
    static __device__ __forceinline__ uint8_t min() { return 0; }
    static __device__ __forceinline__ uint8_t max() { return 1; }
};

// wrap-around integer arithmetic (the reference's C integer ops wrap on
// two's complement hardware); IEEE round-to-nearest for fp64 without fused
// multiply-add, matching the reference's separate MULT and ADD steps
template <class Z> __device__ __forceinline__ Z add_wrap(Z a, Z b)
{
    using U = typename std::make_unsigned<Z>::type;
    return (Z) ((U) a + (U) b);
}
template <> __device__ __forceinline__ double add_wrap<double>(double a, double b) { return __dadd_rn(a, b); }
template <class Z> __device__ __forceinline__ Z mul_wrap(Z a, Z b)
{
    using U = typename std::make_unsigned<Z>::type;
    return (Z) ((U) a * (U) b);
}
template <> __device__ __forceinline__ double mul_wrap<double>(double a, double b) { return __dmul_rn(a, b); }

template <class Z> __device__ __forceinline__ Z zmin(Z a, Z b) { return a < b ? a : b; }
template <class Z> __device__ __forceinline__ Z zmax(Z a, Z b) { return a > b ? a : b; }
// fmin/fmax ignoring a NaN operand (the reference's GB_FMIN/GB_FMAX)
template <> __device__ __forceinline__ double zmin<double>(double a, double b) { return __builtin_fmin(a, b); }
template <> __device__ __forceinline__ double zmax<double>(double a, double b) { return __builtin_fmax(a, b); }

// SR<ADD, MULT, Z>: Z is the result type; valued multipliers read A and B
// as Z (the branch predicates require xtype == ytype == ztype for them).
template <int ADD, int MULT, class Z_>
struct SR
{
    using Z = Z_;
    static constexpr int add = ADD;
    static constexpr int mult = MULT;
    static constexpr bool pattern = (MULT == GM_MULT_PAIR || MULT == GM_MULT_K);
    static constexpr bool is_any = (ADD == GM_ADD_ANY);
    // plus_pair on a 64-bit or fp result: "exists iff cij != 0"
    // (Source/mxm/include/GB_AxB_dot_cij.h:40-49)
    static constexpr bool plus_pair_wide =
        (ADD == GM_ADD_PLUS && MULT == GM_MULT_PAIR && sizeof(Z) == 8);
    // ANY_PAIR: iso result, value 1 (Source/mxm/GB_AxB__any_pair_iso.c)
    static constexpr bool iso_out = (ADD == GM_ADD_ANY && MULT == GM_MULT_PAIR);
    // ANY with a positional k: partial results combine to the smallest k,
    // whatever order the lanes visited the k's in
    static constexpr bool any_min_k = (ADD == GM_ADD_ANY && MULT == GM_MULT_K);

    static __device__ __forceinline__ Z identity()
    {
        if constexpr (ADD == GM_ADD_PLUS) return (Z) 0;
        else if constexpr (ADD == GM_ADD_MIN) return Lim<Z>::max();
        else if constexpr (ADD == GM_ADD_MAX) return Lim<Z>::min();
        else return (Z) 0;
    }
    static __device__ __forceinline__ Z mul(Z a, Z b, int64_t k)
    {
        if constexpr (MULT == GM_MULT_TIMES) return mul_wrap<Z>(a, b);
        else if constexpr (MULT == GM_MULT_PLUS) return add_wrap<Z>(a, b);
        else if constexpr (MULT == GM_MULT_PAIR) return (Z) 1;
        else return (Z) k;
    }
    // z = z (+) t
    static __device__ __forceinline__ Z upd(Z z, Z t)
    {
        if constexpr (ADD == GM_ADD_PLUS) return add_wrap<Z>(z, t);
        else if constexpr (ADD == GM_ADD_MIN) return zmin<Z>(z, t);
        else if constexpr (ADD == GM_ADD_MAX) return zmax<Z>(z, t);
        else return t;
    }
    static __device__ __forceinline__ bool terminal(Z z)
    {
        if constexpr (ADD == GM_ADD_MIN) return z == Lim<Z>::min();
        else if constexpr (ADD == GM_ADD_MAX) return z == Lim<Z>::max();
        else if constexpr (ADD == GM_ADD_ANY) return true;
        else return false;
    }
    static __device__ __forceinline__ Z geta(const Z *x, int64_t p, int iso)
    {
        if constexpr (pattern) return (Z) 0;
        else return x[iso ? 0 : p];
    }
};

// ordered combine of two partial results (lower lane first): for ANY the
// lower lane's witness wins, so a reduction in lane order keeps the first
// match in k order
template <class S>
__device__ __forceinline__ void combine(bool &e, typename S::Z &z, bool e2, typename S::Z z2)
{
    if (!e2) return;
    if (!e) { e = true; z = z2; return; }
    if constexpr (S::any_min_k) z = (z2 < z) ? z2 : z;
    else if constexpr (!S::is_any) z = S::upd(z, z2);
}

// shuffle helpers for 64-bit and 8-bit payloads
template <class Z> __device__ __forceinline__ Z shfl_down(Z v, int off, int width)
{
    if constexpr (sizeof(Z) == 8)
    {
        int64_t t;
        __builtin_memcpy(&t, &v, 8);
        int lo = __shfl_down((int) (t & 0xffffffff), off, width);
        int hi = __shfl_down((int) (t >> 32), off, width);
        int64_t r = ((int64_t) (uint32_t) lo) | ((int64_t) hi << 32);
        Z out;
        __builtin_memcpy(&out, &r, 8);
        return out;
    }
    else if constexpr (sizeof(Z) == 4)
    {
        int t;
        __builtin_memcpy(&t, &v, 4);
        t = __shfl_down(t, off, width);
        Z out;
        __builtin_memcpy(&out, &t, 4);
        return out;
    }
    else
    {
        return (Z) __shfl_down((int) v, off, width);
    }
}

// lane src's value (index shuffle) for any payload of 1, 2, 4 or 8 bytes
template <class Z> __device__ __forceinline__ Z shfl_any(Z v, int src, int width)
{
    if constexpr (sizeof(Z) == 8)
    {
        int64_t t;
        __builtin_memcpy(&t, &v, 8);
        int lo = __shfl((int) (t & 0xffffffff), src, width);
        int hi = __shfl((int) (t >> 32), src, width);
        int64_t r = ((int64_t) (uint32_t) lo) | ((int64_t) hi << 32);
        Z out;
        __builtin_memcpy(&out, &r, 8);
        return out;
    }
    else if constexpr (sizeof(Z) == 4)
    {
        int t;
        __builtin_memcpy(&t, &v, 4);
        t = __shfl(t, src, width);
        Z out;
        __builtin_memcpy(&out, &t, 4);
        return out;
    }
    else
    {
        return (Z) __shfl((int) v, src, width);
    }
}

// in-order tree reduction over a team of G lanes (G a power of 2 <= 64);
// lane 0 of the team holds the result
template <class S, int G>
__device__ __forceinline__ void team_reduce(bool &e, typename S::Z &z)
{
    using Z = typename S::Z;
    #pragma unroll
    for (int off = 1; off < G; off <<= 1)
    {
        // pair lanes (2m*off, 2m*off + off): the lower one absorbs the upper
        int e2 = __shfl_down((int) e, off, G);
        Z z2 = shfl_down<Z>(z, off, G);
        const int lane = threadIdx.x & (G - 1);
        if ((lane & (2 * off - 1)) == 0) combine<S>(e, z, e2 != 0, z2);
    }
}

} // namespace gm

// semiring instance list: X(ADD, MULT, ZTYPE_CODE, Z)
#define GM_SEMIRING_LIST(X)                          \
    X(GM_ADD_PLUS, GM_MULT_PAIR, GM_INT64, int64_t)  \
    X(GM_ADD_PLUS, GM_MULT_PAIR, GM_INT32, int32_t)  \
    X(GM_ADD_ANY, GM_MULT_PAIR, GM_BOOL, uint8_t)    \
    X(GM_ADD_PLUS, GM_MULT_TIMES, GM_FP64, double)   \
    X(GM_ADD_MIN, GM_MULT_PLUS, GM_INT32, int32_t)   \
    X(GM_ADD_ANY, GM_MULT_K, GM_INT64, int64_t)      \
    X(GM_ADD_PLUS, GM_MULT_TIMES, GM_INT64, int64_t) \
    X(GM_ADD_MIN, GM_MULT_PLUS, GM_INT64, int64_t)   \
    X(GM_ADD_MAX, GM_MULT_PLUS, GM_INT32, int32_t)   \
    X(GM_ADD_MIN, GM_MULT_PLUS, GM_FP64, double)

namespace gm
{
// true if (add, mult, ztype) is in GM_SEMIRING_LIST
bool semiring_supported(const gm_semiring *sr);
// operand value types acceptable for the semiring
bool operand_ok(const gm_semiring *sr, const gm_csr *X);
bool csr_ok(const gm_csr *X);
}
