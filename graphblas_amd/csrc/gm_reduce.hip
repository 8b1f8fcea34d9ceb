// gm_reduce.hip: reduce a value array to a scalar with a monoid.
// Reference: Source/reduce/template/GB_reduce_to_scalar_template.c:40-95
// (the GrB_reduce that closes the triangle count).  Two passes: a fixed
// grid of partial results, then one workgroup; the combine order depends only
// on n, so fp64 sums are reproducible run to run.

#include "gm_internal.h"
#include "gm_block.h"
#include <limits>
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

} // namespace gm

using namespace gm;

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
