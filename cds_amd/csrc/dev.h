// Device-side helpers shared by the HIP translation units (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace cds {

constexpr uint32_t kAlive = 0xFFFFFFFFu;  // "not yet removed / killed"
constexpr int kWave = 64;                 // CDNA wavefront width

__device__ __forceinline__ unsigned lane() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanes_below() {
  return (1ull << lane()) - 1ull;
}

// Order-preserving wave compaction: returns this lane's slot among the
// lanes with `pred` set and writes the wave's total to *count.
__device__ __forceinline__ unsigned wave_slot(bool pred, unsigned *count) {
  const uint64_t mask = __ballot(pred);
  *count = __popcll(mask);
  return __popcll(mask & lanes_below());
}

// Wave-aggregated atomicAdd(&ctr[key], 1) for every lane with `active`:
// lanes with equal keys share one atomic; returns this lane's 0-based slot
// within its key's reservation (order among equal keys = lane order).
__device__ __forceinline__ uint32_t aggregated_reserve(uint32_t *ctr,
                                                       uint32_t key,
                                                       bool active) {
  uint64_t pending = __ballot(active);
  uint32_t slot = 0;
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint32_t lkey = __shfl(key, leader, kWave);
    const uint64_t same = __ballot(active && key == lkey) & pending;
    uint32_t base = 0;
    if ((int)lane() == leader) base = atomicAdd(&ctr[lkey], (uint32_t)__popcll(same));
    base = __shfl(base, leader, kWave);
    if (active && key == lkey) slot = base + __popcll(same & lanes_below());
    pending &= ~same;
  }
  return slot;
}

// Wave-wide sum of a 32-bit value.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ uint64_t wave_min64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}

// Index of `key` in the ascending array a[0..len), or -1.
__device__ __forceinline__ int find_sorted(const uint32_t *a, int len,
                                           uint32_t key) {
  int lo = 0, hi = len;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    uint32_t x = a[mid];
    if (x < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < len && a[lo] == key) ? lo : -1;
}

// Same for 64-bit indexed arrays in global memory.
__device__ __forceinline__ int64_t find_sorted64(const uint32_t *a, int64_t len,
                                                 uint32_t key) {
  int64_t lo = 0, hi = len;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    uint32_t x = a[mid];
    if (x < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < len && a[lo] == key) ? lo : -1;
}

inline unsigned grid_for(size_t work, unsigned block, unsigned cap = 65535u * 16u) {
  size_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace cds
