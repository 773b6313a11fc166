// Prefix sums, reductions and the stream-ordered device arena.
//
// Replaces the Thrust scans / reductions of the reference
// (cudaCode/main.cu:45-47, :67, :365-368, :426-429, :548-551, :669-681;
// cudaCode/utils/cuda_utils.cuh:46-70) with one hand-written scan:
// 256-thread blocks each own 2048 consecutive elements (8 per thread,
// wave64 shuffle scan, then a 4-wave LDS combine); block totals are
// scanned recursively and added back.
#include "dev.h"
#include "kernels.h"

namespace cds {

DeviceArena::~DeviceArena() { trim(); }

void DeviceArena::trim() {
  for (const Block &b : cache_) (void)hipFreeAsync(b.p, stream_);
  cache_.clear();
  cached_ = 0;
}

hipError_t DeviceArena::alloc(void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  size_t real = bytes;
  *p = nullptr;
  if (bytes >= kArenaCacheMin) {
    // smallest cached block that fits and is not more than twice as large
    size_t best = cache_.size();
    for (size_t i = 0; i < cache_.size(); ++i)
      if (cache_[i].real >= bytes && cache_[i].real / 2 <= bytes &&
          (best == cache_.size() || cache_[i].real < cache_[best].real))
        best = i;
    if (best < cache_.size()) {
      *p = cache_[best].p;
      real = cache_[best].real;
      cached_ -= real;
      cache_[best] = cache_.back();
      cache_.pop_back();
    }
  }
  if (!*p) {
    hipError_t e = hipMallocAsync(p, bytes, stream_);
    if (e != hipSuccess && !cache_.empty()) {
      (void)hipGetLastError();
      trim();
      e = hipMallocAsync(p, bytes, stream_);
    }
    if (e != hipSuccess) return e;
  }
  live_ += bytes;
  if (live_ > peak_) peak_ = live_;
  live_blocks_.push_back({*p, bytes, real});
  return hipSuccess;
}

void DeviceArena::release(void *p) {
  if (!p) return;
  for (size_t i = 0; i < live_blocks_.size(); ++i)
    if (live_blocks_[i].p == p) {
      const Block b = live_blocks_[i];
      live_ -= b.size;
      live_blocks_[i] = live_blocks_.back();
      live_blocks_.pop_back();
      if (b.real >= kArenaCacheMin && cached_ + b.real <= cache_cap_) {
        cache_.push_back(b);
        cached_ += b.real;
        return;
      }
      break;
    }
  (void)hipFreeAsync(p, stream_);
}

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 8;
constexpr int kScanPerBlock = kScanThreads * kScanPerThread;

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kScanThreads) void k_scan_local(
    const Tin *__restrict__ in, Tout *__restrict__ out, Tout *__restrict__ bsum,
    size_t n) {
  __shared__ Tout wave_total[kScanThreads / kWave];
  const size_t base =
      (size_t)blockIdx.x * kScanPerBlock + (size_t)threadIdx.x * kScanPerThread;
  Tout v[kScanPerThread];
  Tout sum = 0;
#pragma unroll
  for (int i = 0; i < kScanPerThread; ++i) {
    const size_t idx = base + i;
    v[i] = idx < n ? (Tout)in[idx] : (Tout)0;
    sum += v[i];
  }
  Tout incl = sum;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    Tout t = __shfl_up(incl, off, kWave);
    if ((int)lane() >= off) incl += t;
  }
  const int wid = threadIdx.x / kWave;
  if (lane() == kWave - 1) wave_total[wid] = incl;
  __syncthreads();
  Tout wprefix = 0;
  for (int w = 0; w < wid; ++w) wprefix += wave_total[w];
  Tout run = wprefix + incl - sum;
#pragma unroll
  for (int i = 0; i < kScanPerThread; ++i) {
    const size_t idx = base + i;
    if (idx < n) out[idx] = run;
    run += v[i];
  }
  if (threadIdx.x == kScanThreads - 1) bsum[blockIdx.x] = wprefix + incl;
}

template <typename Tout>
__global__ void k_scan_add(Tout *__restrict__ out, const Tout *__restrict__ boff,
                           size_t n) {
  const size_t base =
      (size_t)blockIdx.x * kScanPerBlock + (size_t)threadIdx.x * kScanPerThread;
  const Tout add = boff[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanPerThread; ++i) {
    const size_t idx = base + i;
    if (idx < n) out[idx] += add;
  }
}

template <typename Tout>
__global__ void k_set_total(Tout *out, size_t n, const Tout *src) {
  out[n] = *src;
}

template <typename Tin, typename Tout>
hipError_t scan_impl(const Tin *in, Tout *out, size_t n, DeviceArena &a) {
  hipStream_t s = a.stream();
  if (n == 0) return hipMemsetAsync(out, 0, sizeof(Tout), s);
  const size_t nb = (n + kScanPerBlock - 1) / kScanPerBlock;
  if (nb > 0x7FFFFFFFull) return hipErrorInvalidValue;
  Tout *bsum = nullptr, *bscan = nullptr;
  hipError_t e = a.alloc((void **)&bsum, nb * sizeof(Tout));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_scan_local<Tin, Tout>), dim3((unsigned)nb),
                     dim3(kScanThreads), 0, s, in, out, bsum, n);
  e = hipGetLastError();
  if (e == hipSuccess) {
    if (nb == 1) {
      hipLaunchKernelGGL((k_set_total<Tout>), dim3(1), dim3(1), 0, s, out, n,
                         bsum);
      e = hipGetLastError();
    } else {
      e = a.alloc((void **)&bscan, (nb + 1) * sizeof(Tout));
      if (e == hipSuccess) e = scan_impl<Tout, Tout>(bsum, bscan, nb, a);
      if (e == hipSuccess) {
        hipLaunchKernelGGL((k_scan_add<Tout>), dim3((unsigned)nb),
                           dim3(kScanThreads), 0, s, out, bscan, n);
        e = hipGetLastError();
      }
      if (e == hipSuccess) {
        hipLaunchKernelGGL((k_set_total<Tout>), dim3(1), dim3(1), 0, s, out, n,
                           bscan + nb);
        e = hipGetLastError();
      }
      a.release(bscan);
    }
  }
  a.release(bsum);
  return e;
}

template <typename Tout>
__global__ void k_zero_one(Tout *out) {
  *out = 0;
}

// scan_impl with its block sums in a caller-given workspace instead of the
// arena (scan_ws_words(n) words of Tout): bsum (nb), bscan (nb + 1), then
// the workspace of the recursive scan of the block sums.
template <typename Tin, typename Tout>
hipError_t scan_ws_impl(const Tin *in, Tout *out, size_t n, Tout *ws, size_t ws_words,
                        hipStream_t s) {
  if (ws_words < scan_ws_words(n)) return hipErrorInvalidValue;
  if (n == 0) {
    hipLaunchKernelGGL((k_zero_one<Tout>), dim3(1), dim3(1), 0, s, out);
    return hipGetLastError();
  }
  const size_t nb = (n + kScanPerBlock - 1) / kScanPerBlock;
  Tout *bsum = ws;
  hipLaunchKernelGGL((k_scan_local<Tin, Tout>), dim3((unsigned)nb), dim3(kScanThreads), 0,
                     s, in, out, bsum, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (nb == 1) {
    hipLaunchKernelGGL((k_set_total<Tout>), dim3(1), dim3(1), 0, s, out, n, bsum);
    return hipGetLastError();
  }
  Tout *bscan = ws + nb;
  e = scan_ws_impl<Tout, Tout>(bsum, bscan, nb, ws + 2 * nb + 1, ws_words - 2 * nb - 1, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_scan_add<Tout>), dim3((unsigned)nb), dim3(kScanThreads), 0, s, out,
                     bscan, n);
  hipLaunchKernelGGL((k_set_total<Tout>), dim3(1), dim3(1), 0, s, out, n, bscan + nb);
  return hipGetLastError();
}

__global__ void k_reduce_max_u32(const uint32_t *__restrict__ in, size_t n,
                                 uint32_t *out) {
  uint32_t m = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    m = max(m, in[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor(m, off, kWave));
  if (lane() == 0 && m) atomicMax(out, m);
}

}  // namespace

hipError_t scan_u32(const uint32_t *in, uint64_t *out, size_t n,
                    DeviceArena &a) {
  return scan_impl<uint32_t, uint64_t>(in, out, n, a);
}
hipError_t scan_u64(const uint64_t *in, uint64_t *out, size_t n,
                    DeviceArena &a) {
  return scan_impl<uint64_t, uint64_t>(in, out, n, a);
}
hipError_t scan_u32_32(const uint32_t *in, uint32_t *out, size_t n,
                       DeviceArena &a) {
  return scan_impl<uint32_t, uint32_t>(in, out, n, a);
}

size_t scan_ws_words(size_t n) {
  const size_t nb = (n + kScanPerBlock - 1) / kScanPerBlock;
  if (nb <= 1) return 1;
  return 2 * nb + 1 + scan_ws_words(nb);
}
hipError_t scan_u32_ws(const uint32_t *in, uint64_t *out, size_t n, uint64_t *ws,
                       size_t ws_words, hipStream_t s) {
  return scan_ws_impl<uint32_t, uint64_t>(in, out, n, ws, ws_words, s);
}
hipError_t scan_u32_32_ws(const uint32_t *in, uint32_t *out, size_t n, uint32_t *ws,
                          size_t ws_words, hipStream_t s) {
  return scan_ws_impl<uint32_t, uint32_t>(in, out, n, ws, ws_words, s);
}

hipError_t reduce_max_u32(const uint32_t *in, size_t n, uint32_t *out,
                          hipStream_t s) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce_max_u32, dim3(grid_for(n, 256, 4096)), dim3(256),
                     0, s, in, n, out);
  return hipGetLastError();
}

}  // namespace cds
