// Block-level FP64 reduction shared by the reduction kernels: wave sums via
// cross-lane shuffles, then a fixed-order sum over the block's waves, one
// partial per block (bit-reproducible; no atomics).
#pragma once

#include "common.hpp"

namespace iqg {

template <int M>
__device__ __forceinline__ void block_sum_store(double (&v)[M], int m_used, double* __restrict__ part) {
    __shared__ double sh[kRedThreads / 64][M];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        if (m < m_used) {
            double x = v[m];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
            if (lane == 0) sh[wave][m] = x;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < m_used) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kRedThreads / 64; ++w) s += sh[w][threadIdx.x];
        part[(uint64_t)blockIdx.x * m_used + threadIdx.x] = s;
    }
}

}  // namespace iqg
