// Deterministic FP64 reductions: a fixed grid of kReduceBlocks blocks, each
// summing a fixed contiguous range with a fixed thread stride and a fixed LDS
// tree, then one block summing the partials in a fixed order.  Repeated runs
// give identical bits, independent of scheduling.
#pragma once

#include "prh_workspace.hpp"

namespace prh {

constexpr int kReduceThreads = 256;

// Sums `value(i)` over i in [0, total) into partial[blockIdx.x * stride + slot].
template <int NV, typename F>
__device__ __forceinline__ void block_reduce_range(size_t total, F value, double* partial,
                                                   int slot_stride) {
  __shared__ double lds[NV][kReduceThreads];
  const size_t per = (total + gridDim.x - 1) / gridDim.x;
  const size_t beg = blockIdx.x * per;
  const size_t end = beg + per < total ? beg + per : total;
  double acc[NV];
  for (int q = 0; q < NV; q++) acc[q] = 0.0;
  // eight values per thread are loaded before they are added, in the order of
  // the one-at-a-time loop (same sums, more loads in flight)
  constexpr int kU = 8;
  size_t i = beg + threadIdx.x;
  for (; i + (kU - 1) * static_cast<size_t>(kReduceThreads) < end; i += kU * static_cast<size_t>(kReduceThreads)) {
    double v[kU][NV];
#pragma unroll
    for (int u = 0; u < kU; u++) value(i + u * static_cast<size_t>(kReduceThreads), v[u]);
#pragma unroll
    for (int u = 0; u < kU; u++)
      for (int q = 0; q < NV; q++) acc[q] += v[u][q];
  }
  for (; i < end; i += kReduceThreads) {
    double v[NV];
    value(i, v);
    for (int q = 0; q < NV; q++) acc[q] += v[q];
  }
  for (int q = 0; q < NV; q++) lds[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = kReduceThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
      for (int q = 0; q < NV; q++) lds[q][threadIdx.x] += lds[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; q++) partial[q * slot_stride + blockIdx.x] = lds[q][0];
}

// One block of 1024 threads: out[q] = sum of partial[q * stride + 0 .. nblocks).
template <int NV>
__global__ __launch_bounds__(1024) void k_reduce_final(const double* __restrict__ partial,
                                                       int nblocks, int stride,
                                                       double* __restrict__ out) {
  __shared__ double lds[NV][1024];
  for (int q = 0; q < NV; q++) {
    double a = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += 1024) a += partial[q * stride + i];
    lds[q][threadIdx.x] = a;
  }
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
      for (int q = 0; q < NV; q++) lds[q][threadIdx.x] += lds[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; q++) out[q] = lds[q][0];
}

}  // namespace prh
