// gm_reduce.hip: reduce a value array to a scalar with a monoid.
// Reference: Source/reduce/template/GB_reduce_to_scalar_template.c:40-95
// (the GrB_reduce that closes the triangle count).  Two passes: a fixed
// grid of partial results, then one workgroup; the combine order depends only
// on n, so fp64 sums are reproducible run to run.

#include "gm_internal.h"
#include "gm_block.h"
#include <limits>
#include <mutex>
#include <type_traits>

namespace gm
{

constexpr int RD_NT = 256;
constexpr int RD_GRID = 1024;

template <int ADD, class Z>
__device__ __forceinline__ Z rd_op(Z a, Z b)
{
    if constexpr (ADD == GM_ADD_PLUS) return add_wrap<Z>(a, b);
    else if constexpr (ADD == GM_ADD_MIN) return zmin<Z>(a, b);
    else if constexpr (ADD == GM_ADD_MAX) return zmax<Z>(a, b);
    else return a;
}

template <int ADD, class Z>
__device__ __forceinline__ Z rd_identity()
{
    if constexpr (ADD == GM_ADD_PLUS) return (Z) 0;
    else if constexpr (ADD == GM_ADD_MIN) return Lim<Z>::max();
    else if constexpr (ADD == GM_ADD_MAX) return Lim<Z>::min();
    else return (Z) 0;
}

template <int ADD, class Z>
__global__ __launch_bounds__(RD_NT) void k_reduce(const Z *__restrict__ x, int64_t n,
                                                  Z *__restrict__ out)
{
    __shared__ Z s[RD_NT];
    Z acc = rd_identity<ADD, Z>();
    bool any = false;
    for (int64_t t = (int64_t) blockIdx.x * RD_NT + threadIdx.x; t < n;
         t += (int64_t) gridDim.x * RD_NT)
    {
        if constexpr (ADD == GM_ADD_ANY)
        {
            if (!any) { acc = x[t]; any = true; }
        }
        else acc = rd_op<ADD, Z>(acc, x[t]);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        Z r = s[0];
        for (int t = 1; t < RD_NT; t++) r = rd_op<ADD, Z>(r, s[t]);
        out[blockIdx.x] = r;
    }
}

template <int ADD, class Z>
static int reduce_run(void *result, const void *x, int64_t n, int32_t iso)
{
    hipStream_t st = stream();
    Z *hr = (Z *) pinned_scratch(sizeof(Z) * RD_GRID);
    if (!hr) return set_error(GM_OUT_OF_MEMORY, "pinned scratch");
    if (n == 0)
    {
        Z id = (Z) 0;
        if constexpr (ADD == GM_ADD_MIN)
            id = std::is_floating_point<Z>::value ? std::numeric_limits<Z>::infinity()
                                                  : std::numeric_limits<Z>::max();
        else if constexpr (ADD == GM_ADD_MAX)
            id = std::is_floating_point<Z>::value ? -std::numeric_limits<Z>::infinity()
                                                  : std::numeric_limits<Z>::lowest();
        if (std::is_same<Z, uint8_t>::value) id = (ADD == GM_ADD_MIN) ? (Z) 1 : (Z) 0;
        *(Z *) result = id;
        return GM_SUCCESS;
    }
    if (iso)
    {
        GM_HIP(hipMemcpyAsync(hr, x, sizeof(Z), hipMemcpyDeviceToHost, st), "reduce iso");
        GM_HIP(hipStreamSynchronize(st), "reduce iso sync");
        Z v = hr[0], r = v;
        if constexpr (ADD == GM_ADD_PLUS)
        {
            if constexpr (std::is_floating_point<Z>::value) r = (Z) ((double) v * (double) n);
            else r = (Z) ((uint64_t) v * (uint64_t) n);
        }
        *(Z *) result = r;
        return GM_SUCCESS;
    }
    DevBuf part;
    if (!part.alloc(sizeof(Z) * RD_GRID)) return set_error(GM_OUT_OF_MEMORY, "reduce workspace");
    {
        ProfScope ps("reduce");
        k_reduce<ADD, Z><<<RD_GRID, RD_NT, 0, st>>>((const Z *) x, n, part.as<Z>());
        GM_HIP(hipGetLastError(), "k_reduce");
        k_reduce<ADD, Z><<<1, RD_NT, 0, st>>>(part.as<Z>(), RD_GRID, part.as<Z>());
        GM_HIP(hipGetLastError(), "k_reduce final");
    }
    GM_HIP(hipMemcpyAsync(hr, part.ptr, sizeof(Z), hipMemcpyDeviceToHost, st), "reduce d2h");
    GM_HIP(hipStreamSynchronize(st), "reduce sync");
    *(Z *) result = hr[0];
    return GM_SUCCESS;
}

//------------------------------------------------------------------------------
// allocation-free form: partials in a fixed device array, result on the device
//------------------------------------------------------------------------------

// The first pass's RD_GRID partial results (at most 8 bytes each) live in
// `scratch`, one copy per GPU, held until the process ends.  In host code the
// name is only a stand-in: kernels receive the device address that
// hipGetSymbolAddress returned for the current device (g_scratch_dev).
// Every partial is written by its workgroup before the second pass reads it,
// so nothing needs clearing between calls; two calls that overlap on one GPU
// (on different streams) race on the array.
constexpr size_t RD_SCRATCH_BYTES = (size_t) RD_GRID * 8;
static __device__ __align__(256) unsigned char scratch[RD_SCRATCH_BYTES];
constexpr int RD_MAX_DEVICES = 64;
static void *g_scratch_dev[RD_MAX_DEVICES];
static std::mutex g_scratch_mu;

// looks up the current device's scratch address once (the library's init
// calls it with the device current; a call on another device does it on
// first use)
int reduce_scratch_setup()
{
    int dev = 0;
    GM_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= RD_MAX_DEVICES)
        return set_error(GM_PANIC, "reduce scratch: device %d beyond %d", dev, RD_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[dev]) return GM_SUCCESS;
    void *ptr = nullptr;
    GM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) return set_error(GM_PANIC, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_dev[dev] = ptr;
    return GM_SUCCESS;
}

// the monoid's value for an empty array, as reduce_run returns it
template <int ADD, class Z>
__host__ __device__ constexpr Z rd_empty()
{
    if constexpr (std::is_same<Z, uint8_t>::value) return (ADD == GM_ADD_MIN) ? (Z) 1 : (Z) 0;
    else if constexpr (ADD == GM_ADD_MIN)
        return std::is_floating_point<Z>::value ? std::numeric_limits<Z>::infinity()
                                                : std::numeric_limits<Z>::max();
    else if constexpr (ADD == GM_ADD_MAX)
        return std::is_floating_point<Z>::value ? -std::numeric_limits<Z>::infinity()
                                                : std::numeric_limits<Z>::lowest();
    else return (Z) 0;
}

// n == 0 (identity) or an iso array (x[0] combined n times), on the device
template <int ADD, class Z>
__global__ void k_reduce_small(const Z *__restrict__ x, int64_t n, int32_t iso, Z *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (n == 0) { *out = rd_empty<ADD, Z>(); return; }
    const Z v = x[0];
    Z r = v;
    if constexpr (ADD == GM_ADD_PLUS)
    {
        if constexpr (std::is_floating_point<Z>::value) r = (Z) ((double) v * (double) n);
        else r = (Z) ((uint64_t) v * (uint64_t) n);
    }
    (void) iso;
    *out = r;
}

template <int ADD, class Z>
static int reduce_run_noalloc(void *result, const void *x, int64_t n, int32_t iso, hipStream_t st,
                              void *scr)
{
    static_assert((size_t) RD_GRID * sizeof(Z) <= RD_SCRATCH_BYTES, "partials fit the scratch");
    if (n == 0 || iso)
    {
        k_reduce_small<ADD, Z><<<1, 64, 0, st>>>((const Z *) x, n, iso, (Z *) result);
        GM_HIP(hipGetLastError(), "k_reduce_small");
        return GM_SUCCESS;
    }
    k_reduce<ADD, Z><<<RD_GRID, RD_NT, 0, st>>>((const Z *) x, n, (Z *) scr);
    GM_HIP(hipGetLastError(), "k_reduce (noalloc)");
    k_reduce<ADD, Z><<<1, RD_NT, 0, st>>>((const Z *) scr, RD_GRID, (Z *) result);
    GM_HIP(hipGetLastError(), "k_reduce final (noalloc)");
    return GM_SUCCESS;
}

} // namespace gm

using namespace gm;

extern "C" int gm_reduce_scalar_noalloc(void *result, int32_t add_op, int32_t type, const void *x,
                                        int64_t n, int32_t iso, void *stream)
{
    if (!result || (n > 0 && !x) || n < 0)
        return set_error(GM_INVALID_VALUE, "gm_reduce_scalar_noalloc: NULL pointer or n < 0");
    // no implicit initialisation here: gm_init creates the stream and looks
    // up the scratch address, and must run at the caller's set-up
    if (!is_initialised())
        return set_error(GM_PANIC, "gm_reduce_scalar_noalloc: call gm_init() first");
    int dev = 0;
    GM_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= RD_MAX_DEVICES)
        return set_error(GM_PANIC, "gm_reduce_scalar_noalloc: device %d beyond %d", dev,
                         RD_MAX_DEVICES);
    void *scr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        scr = g_scratch_dev[dev];
    }
    if (!scr)
    {
        GM_TRY(reduce_scratch_setup());
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        scr = g_scratch_dev[dev];
    }
    hipStream_t st = stream ? (hipStream_t) stream : gm::stream();
#define GM_RDN(ADD, TC, Z) \
    if (add_op == ADD && type == TC) return reduce_run_noalloc<ADD, Z>(result, x, n, iso, st, scr);
#define GM_RDN_T(TC, Z) GM_RDN(GM_ADD_PLUS, TC, Z) GM_RDN(GM_ADD_MIN, TC, Z) \
                        GM_RDN(GM_ADD_MAX, TC, Z) GM_RDN(GM_ADD_ANY, TC, Z)
    GM_RDN_T(GM_INT32, int32_t)
    GM_RDN_T(GM_INT64, int64_t)
    GM_RDN_T(GM_FP64, double)
    GM_RDN(GM_ADD_MAX, GM_BOOL, uint8_t)   // LOR
    GM_RDN(GM_ADD_MIN, GM_BOOL, uint8_t)   // LAND
    GM_RDN(GM_ADD_ANY, GM_BOOL, uint8_t)
#undef GM_RDN_T
#undef GM_RDN
    return set_error(GM_NOT_IMPLEMENTED, "gm_reduce_scalar_noalloc: monoid/type not supported");
}

extern "C" int gm_reduce_scalar(void *result, int32_t add_op, int32_t type, const void *x,
                                int64_t n, int32_t iso)
{
    if (!result || (n > 0 && !x)) return set_error(GM_INVALID_VALUE, "gm_reduce_scalar: NULL");
    GM_TRY(ensure_init());
#define GM_RD(ADD, TC, Z) \
    if (add_op == ADD && type == TC) return reduce_run<ADD, Z>(result, x, n, iso);
#define GM_RD_T(TC, Z) GM_RD(GM_ADD_PLUS, TC, Z) GM_RD(GM_ADD_MIN, TC, Z) \
                       GM_RD(GM_ADD_MAX, TC, Z) GM_RD(GM_ADD_ANY, TC, Z)
    GM_RD_T(GM_INT32, int32_t)
    GM_RD_T(GM_INT64, int64_t)
    GM_RD_T(GM_FP64, double)
    GM_RD(GM_ADD_MAX, GM_BOOL, uint8_t)   // LOR
    GM_RD(GM_ADD_MIN, GM_BOOL, uint8_t)   // LAND
    GM_RD(GM_ADD_ANY, GM_BOOL, uint8_t)
#undef GM_RD_T
#undef GM_RD
    return set_error(GM_NOT_IMPLEMENTED, "gm_reduce_scalar: monoid/type not supported");
}
