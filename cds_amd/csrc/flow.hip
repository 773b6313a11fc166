// (c) Result accounting of the exact flow stage: k-cliques whose members
// all lie in a vertex set (the min-cut side of one component, or the final
// answer set over the whole graph), as the oracle counts them
// (cpuCode/c++/src/helpers.cpp:1210-1230).  The max flow itself is in
// flow_grid.hip.
#include "dev.h"
#include "kernels.h"

namespace cds {
namespace {

__global__ void k_count_in_reach(uint32_t T, uint32_t k,
                                 const uint32_t *__restrict__ mem,
                                 uint32_t vbase, const int *__restrict__ reach,
                                 unsigned long long *out, uint32_t ms) {
  unsigned long long c = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T;
       t += gridDim.x * blockDim.x) {
    bool all = true;
    for (uint32_t j = 0; j < k; ++j)
      all &= reach[mem[(uint64_t)t * ms + j] - vbase] != 0;
    c += all;
  }
  c = wave_sum64(c);
  if (lane() == 0 && c) atomicAdd(out, c);
}

__global__ void k_count_in_mask(uint64_t T, uint32_t k,
                                const uint32_t *__restrict__ cl,
                                const uint8_t *__restrict__ mask,
                                unsigned long long *out, uint32_t ms) {
  unsigned long long c = 0;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T;
       t += (uint64_t)gridDim.x * blockDim.x) {
    bool all = true;
    for (uint32_t j = 0; j < k; ++j) all &= mask[cl[t * ms + j]] != 0;
    c += all;
  }
  c = wave_sum64(c);
  if (lane() == 0 && c) atomicAdd(out, c);
}

}  // namespace

hipError_t count_cliques_in_reach(uint32_t T, uint32_t k, const uint32_t *mem,
                                  uint32_t vbase, const int *reach,
                                  unsigned long long *out, hipStream_t s, uint32_t ms) {
  if (ms == 0) ms = k;
  hipError_t e = hipMemsetAsync(out, 0, sizeof(*out), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_count_in_reach, dim3(grid_for(T, 256, 8192)), dim3(256),
                     0, s, T, k, mem, vbase, reach, out, ms);
  return hipGetLastError();
}

hipError_t count_cliques_in_mask(uint64_t T, uint32_t k, const uint32_t *cl,
                                 const uint8_t *mask, unsigned long long *out,
                                 hipStream_t s, uint32_t ms) {
  if (ms == 0) ms = k;
  hipError_t e = hipMemsetAsync(out, 0, sizeof(*out), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_count_in_mask, dim3(grid_for(T, 256, 8192)), dim3(256),
                     0, s, T, k, cl, mask, out, ms);
  return hipGetLastError();
}

}  // namespace cds
