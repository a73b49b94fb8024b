// Deterministic grid reductions without fences or atomics.
//
// A producing kernel reduces its values per block (wavefront shuffles, then
// LDS across the block's 4 wavefronts) and writes one partial per block and
// value: part[q * nblocks + block].  A one-block finishing kernel then sums
// the partials in a fixed order.  On gfx950 a device-scope release fence per
// block (the classic "last block finishes" pattern) writes back the L2 and
// costs tens of nanoseconds per block, far more than one extra launch.
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

__device__ __forceinline__ int tid256() { return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z); }

// Sum of nv values over a 256-thread block, written as partials of block bid.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* part, int nblocks, int bid)
{
    __shared__ double sm[4 * NV];
    const int tid = tid256();
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if ((tid & 63) == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) sm[4 * q + (tid >> 6)] = v[q];
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) part[(long)q * nblocks + bid] = (sm[4 * q] + sm[4 * q + 1]) + (sm[4 * q + 2] + sm[4 * q + 3]);
    }
}

// In a one-block (256-thread) finishing kernel: totals of nv partial rows,
// valid in thread 0.
template <int NV>
__device__ __forceinline__ void finish_sums(const double* part, int nblocks, double (&tot)[NV])
{
    __shared__ double sm[4 * NV];
    const int tid = tid256();
    double t[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) t[q] = 0.0;
    for (int b = tid; b < nblocks; b += 256) {
#pragma unroll
        for (int q = 0; q < NV; ++q) t[q] += part[(long)q * nblocks + b];
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) t[q] = wave_sum(t[q]);
    if ((tid & 63) == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) sm[4 * q + (tid >> 6)] = t[q];
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) tot[q] = (sm[4 * q] + sm[4 * q + 1]) + (sm[4 * q + 2] + sm[4 * q + 3]);
    }
}

}  // namespace tpls
