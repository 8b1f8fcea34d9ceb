// gm_block.h: wavefront / workgroup scan primitives for gfx950 (64-wide waves).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gm
{

// inclusive scan across the 64 lanes of a wavefront
template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v)
{
    const int lane = threadIdx.x & 63;
    #pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        T u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// exclusive scan across a workgroup of NT threads (NT a multiple of 64,
// NT <= 1024).  lds must hold NT/64 + 1 values.  Returns the exclusive
// prefix of v; *total receives the workgroup sum.  All threads must call.
template <class T, int NT>
__device__ __forceinline__ T block_exclusive_scan(T v, T *lds, T *total)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    T inc = wave_inclusive_scan<T>(v);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        T s = 0;
        for (int w = 0; w < NW; w++) { T t = lds[w]; lds[w] = s; s += t; }
        lds[NW] = s;
    }
    __syncthreads();
    T r = inc - v + lds[wid];
    *total = lds[NW];
    __syncthreads();
    return r;
}

// sum across a workgroup (all threads get the result)
template <class T, int NT>
__device__ __forceinline__ T block_sum(T v, T *lds)
{
    T tot;
    (void) block_exclusive_scan<T, NT>(v, lds, &tot);
    return tot;
}

} // namespace gm
