// Device scratch of the zero-allocation entry point cds_count_cliques_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the clique count
// of cds_count_cliques with every scratch buffer carved out of one static
// device array of kNoallocScratchBytes.  Each GPU the library runs on holds
// its own copy of the array until the process ends.  The host finds the
// copy of each device with hipGetSymbolAddress once, at engine creation
// (noalloc_attach), and passes that address to the kernels; the array's name
// in host code is only a host-side stand-in for the symbol.  Every call on
// one GPU shares the array: calls that overlap on one GPU race.
#include <mutex>

#include "dev.h"
#include "kernels.h"

namespace cds {

namespace {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices];  // per device: address of its copy of `scratch`

__global__ void k_fill_u64(uint64_t *a, size_t n, uint64_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// out = sum of in[0..n); *out must be zero before the launch.
__global__ void k_sum_u32(const uint32_t *__restrict__ in, size_t n,
                          unsigned long long *out) {
  unsigned long long s = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    s += in[i];
  s = wave_sum64(s);
  if (lane() == 0 && s) atomicAdd(out, s);
}

}  // namespace

hipError_t noalloc_attach(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch[device] = p;
  return hipSuccess;
}

void *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_scratch[device];
}

hipError_t fill_u64(uint64_t *a, size_t n, uint64_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, a, n, v);
  return hipGetLastError();
}

hipError_t sum_u32(const uint32_t *in, size_t n, unsigned long long *out, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s,
                     reinterpret_cast<uint64_t *>(out), (size_t)1, (uint64_t)0);
  hipLaunchKernelGGL(k_sum_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, in, n, out);
  return hipGetLastError();
}

}  // namespace cds
