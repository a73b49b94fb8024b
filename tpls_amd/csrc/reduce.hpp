// Deterministic block reductions for 256-thread blocks (4 wavefronts of 64).
//
// Each block reduces its values in registers (wavefront shuffles, then LDS
// across the 4 waves) and writes one partial.  The last block to arrive (an
// agent-scope atomic counter, fenced on both sides so that partials written
// on other XCDs are visible) sums all partials in block order.  The result
// therefore depends only on the launch shape, never on scheduling.
#pragma once

#include <hip/hip_runtime.h>

#include "tpls_kernels.hpp"

namespace tpls {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over a 256-thread block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum256(double v, double* sm)
{
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    v = wave_sum(v);
    __syncthreads();
    if ((tid & 63) == 0) sm[tid >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (tid == 0) r = (sm[0] + sm[1]) + (sm[2] + sm[3]);
    return r;
}

// Two sums at once (halves the synchronisations of fused CG kernels).
__device__ __forceinline__ void block_sum256x2(double& a, double& b, double* sm)
{
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    a = wave_sum(a);
    b = wave_sum(b);
    __syncthreads();
    if ((tid & 63) == 0) {
        sm[tid >> 6] = a;
        sm[4 + (tid >> 6)] = b;
    }
    __syncthreads();
    if (tid == 0) {
        a = (sm[0] + sm[1]) + (sm[2] + sm[3]);
        b = (sm[4] + sm[5]) + (sm[6] + sm[7]);
    }
}

// Publish the block partial(s); returns true in every thread of the last
// block, after which red.part holds all partials.
__device__ __forceinline__ bool publish_partials(const Reduce& red, int nval, const double* vals, int nblocks,
                                                 int bid)
{
    __shared__ int last;
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    if (tid == 0) {
        for (int q = 0; q < nval; ++q) red.part[q * nblocks + bid] = vals[q];
        __threadfence();
        unsigned prev = atomicAdd(red.counter, 1u);
        last = (prev == unsigned(nblocks - 1));
    }
    __syncthreads();
    if (last) __threadfence();
    return last != 0;
}

// Sum nblocks partials (in block order per thread stripe, then the fixed
// block tree).  Valid in thread 0.  Call only from the last block.
__device__ __forceinline__ double sum_partials(const Reduce& red, int which, int nblocks, double* sm)
{
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    double t = 0.0;
    const volatile double* p = red.part + (long)which * nblocks;
    for (int q = tid; q < nblocks; q += 256) t += p[q];
    return block_sum256(t, sm);
}

// Single-value reduction finishing into *red.out (used by diagnostics).
__device__ __forceinline__ void block_reduce_finish(double v, const Reduce& red)
{
    __shared__ double sm[8];
    const int nblocks = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    double s = block_sum256(v, sm);
    if (publish_partials(red, 1, &s, nblocks, bid)) {
        double t = sum_partials(red, 0, nblocks, sm);
        const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
        if (tid == 0) {
            *red.out = t;
            *red.counter = 0u;
        }
    }
}

}  // namespace tpls
