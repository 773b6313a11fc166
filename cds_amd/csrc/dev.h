// Device-side helpers shared by the HIP translation units (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdlib>

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

// Lane l's 64-bit value, in every lane (l wave-uniform).
__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)l);
  return ((uint64_t)hi << 32) | lo;
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

// ---- DPP lane exchanges (a VALU operand modifier: no LDS round trip) ----
// Inside rows of 16 lanes: xor 1 and xor 2 (quad permutes), the mirror
// within 8 lanes and within 16 lanes; row shifts right by 1, 2, 4, 8 (lanes
// shifted in from outside the row read the `old` operand, 0); and the GFX9
// row broadcasts (lane 15 of each row into the next row; lane 31 into rows
// 2 and 3).  After xor 1, xor 2, half mirror and mirror every lane of a row
// holds a commutative reduction over the row.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E;
constexpr int kDppShr1 = 0x111, kDppShr2 = 0x112, kDppShr4 = 0x114, kDppShr8 = 0x118;
constexpr int kDppMirror = 0x140, kDppHalfMirror = 0x141;
constexpr int kDppBcast15 = 0x142, kDppBcast31 = 0x143;

template <int kCtrl, int kRowMask = 0xF>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, kRowMask, 0xF, false);
}

template <int kCtrl>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
  const uint32_t lo = dpp<kCtrl>((uint32_t)v), hi = dpp<kCtrl>((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Sum over aligned groups of G lanes (G = 8, 16, 32 or 64), in every lane
// of the group.  Every lane of the wave must be active.
template <uint32_t G>
__device__ __forceinline__ uint32_t group_sum(uint32_t v) {
  static_assert(G == 8 || G == 16 || G == 32 || G == 64, "groups of 8 to 64 lanes");
  v += dpp<kDppXor1>(v);
  v += dpp<kDppXor2>(v);
  v += dpp<kDppHalfMirror>(v);
  if (G >= 16) v += dpp<kDppMirror>(v);
  if (G >= 32) v += __shfl_xor(v, 16, kWave);
  if (G >= 64) v += __shfl_xor(v, 32, kWave);
  return v;
}

// Smallest 64-bit key of the wave with a 32-bit payload from the lane that
// holds it, in every lane (keys distinct, or equal keys with equal
// payloads).  Every lane of the wave must be active.
template <int kCtrl>
__device__ __forceinline__ void min_step(uint64_t &key, uint32_t &pay) {
  const uint64_t ok = dpp64<kCtrl>(key);
  const uint32_t op = dpp<kCtrl>(pay);
  if (ok < key) {
    key = ok;
    pay = op;
  }
}
__device__ __forceinline__ void wave_argmin64(uint64_t &key, uint32_t &pay) {
  min_step<kDppXor1>(key, pay);
  min_step<kDppXor2>(key, pay);
  min_step<kDppHalfMirror>(key, pay);
  min_step<kDppMirror>(key, pay);
#pragma unroll
  for (int off = 16; off <= 32; off <<= 1) {
    const uint64_t ok = __shfl_xor(key, off, kWave);
    const uint32_t op = __shfl_xor(pay, off, kWave);
    if (ok < key) {
      key = ok;
      pay = op;
    }
  }
}
// The same over the first 16 lanes (each row of 16 reduces on its own).
__device__ __forceinline__ void row_argmin64(uint64_t &key, uint32_t &pay) {
  min_step<kDppXor1>(key, pay);
  min_step<kDppXor2>(key, pay);
  min_step<kDppHalfMirror>(key, pay);
  min_step<kDppMirror>(key, pay);
}

// Inclusive prefix sum over the wave's 64 lanes.  Every lane of the wave
// must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += dpp<kDppShr1>(v);
  v += dpp<kDppShr2>(v);
  v += dpp<kDppShr4>(v);
  v += dpp<kDppShr8>(v);
  v += dpp<kDppBcast15, 0xA>(v);
  v += dpp<kDppBcast31, 0xC>(v);
  return v;
}

// Last index i < count with key[i] <= x, for non-decreasing key[] with
// key[0] <= x.  Wave-cooperative 64-ary search: each step every lane loads
// one sample, so the dependent loads number ceil(log64(count)) instead of
// log2(count).  Every lane of the wave calls it with the same arguments.
__device__ __forceinline__ uint32_t wave_last_le(const uint64_t *key,
                                                 uint32_t count, uint64_t x) {
  uint32_t lo = 0, hi = count - 1;
  while (lo < hi) {
    const uint32_t step = (hi - lo) / kWave + 1;  // ceil((hi-lo+1)/64)
    const uint32_t i = lo + lane() * step;
    const uint64_t b = __ballot(i <= hi && key[i] <= x);
    const uint32_t j = 63 - __clzll(b);  // lane 0 samples lo: always set
    lo += j * step;
    hi = min(hi, lo + step - 1);
  }
  return lo;
}

// Member of pair q (clique q / k, position q % k) in a member array that
// holds ms words per clique (ms = k: plain lists; ms > k: clique records).
__device__ __forceinline__ uint32_t member_at(const uint32_t *mem, uint64_t q, uint32_t k,
                                              uint32_t ms) {
  if (ms == k) return mem[q];
  const uint64_t c = k == 2 ? q >> 1 : k == 4 ? q >> 2 : k == 3 ? q / 3 : q / k;
  return mem[c * ms + (q - c * k)];
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

// Barrier across all workgroups of a cooperative launch (B: GridBarrier).
// Arrivals go to one of eight group counters (blockIdx % 8); the last
// arrival of a group bumps the top counter and the last group flips the
// generation, so no counter sees more than gridDim/8 + 8 contending
// atomics.  One thread per workgroup releases the workgroup's stores
// device-wide before arriving (the fence writes back this XCD's L2) and
// acquires after the flip; the workgroup barriers around it order the
// other waves, which share the CU's caches.
// The wait is bounded: after b->limit wall-clock ticks (0: the default)
// the workgroup sets b->abort and ends (s_endpgm, every wave of it, right
// after the workgroup barrier); a workgroup that finds abort set ends the
// same way.  So a launch whose workgroups are not all resident finishes
// instead of hanging, its results invalid and b->abort set for the host.
// group_barrier is the same over the `nblocks` co-resident workgroups of
// one problem inside a larger launch (bid = this workgroup's index among
// them).
constexpr unsigned long long kBarrierDefaultTicks = 6000000000ull;  // 60 s at 100 MHz

// Thread 0's part of group_barrier: arrive, then wait for the generation
// to flip; returns 1 if the wait ran out or another workgroup's did.  Kept
// out of line: inlined at every barrier of the flow kernel it raised
// register pressure into scratch spills.
template <class B>
__device__ __noinline__ unsigned barrier_arrive_wait(B *b, unsigned nblocks, unsigned bid) {
  if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return 1;
  __threadfence();
  const unsigned gen = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned grp = bid & 7u;
  const unsigned gsz = (nblocks + 7u - grp) >> 3;
  const unsigned groups = nblocks < 8u ? nblocks : 8u;
  const unsigned a = __hip_atomic_fetch_add(&b->sub[grp * 32], 1u, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
  if ((a + 1u) % gsz == 0u) {
    const unsigned t =
        __hip_atomic_fetch_add(&b->top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((t + 1u) % groups == 0u)
      __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
  const unsigned long long t0 = wall_clock64();
  unsigned ab = 0;
  while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
    if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      ab = 1;
      break;
    }
    const unsigned long long lim =
        __hip_atomic_load(&b->limit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (wall_clock64() - t0 > (lim ? lim : kBarrierDefaultTicks)) {
      if (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __hip_atomic_store(&b->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ab = 1;
      }
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __threadfence();
  return ab;
}

template <class B>
__device__ __forceinline__ void group_barrier(B *b, unsigned nblocks, unsigned bid) {
  __shared__ unsigned s_abort;
  __syncthreads();
  // One workgroup: its waves share the CU's caches, so the workgroup
  // barrier (with its workgroup-scope fence) is the whole barrier.
  if (nblocks == 1) return;
  if (threadIdx.x == 0) s_abort = barrier_arrive_wait(b, nblocks, bid);
  __syncthreads();
  if (s_abort) __builtin_amdgcn_endpgm();
}

template <class B>
__device__ __forceinline__ void grid_barrier(B *b) {
  group_barrier(b, gridDim.x, blockIdx.x);
}

// Tick limit of one barrier wait: CDS_BARRIER_TIMEOUT_US (microseconds)
// if set, else 0 (the kernels' default, 60 s).  Read at every launch.
inline unsigned long long barrier_limit_ticks() {
  const char *v = getenv("CDS_BARRIER_TIMEOUT_US");
  if (!v || !*v) return 0;
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
      khz <= 0)
    khz = 100000;
  const unsigned long long t = strtoull(v, nullptr, 10) * (unsigned long long)khz / 1000ull;
  return t ? t : 1;
}

// Launches a kernel that synchronises its workgroups with grid_barrier:
// cooperatively (all workgroups co-resident) when there are several, as a
// plain launch when there is one, which needs no co-residency and lets
// independent single-workgroup solves overlap on other streams.
inline hipError_t launch_grid(const void *fn, unsigned grid, unsigned threads,
                              void **args, size_t lds, hipStream_t s) {
  if (grid == 1) return hipLaunchKernel(fn, dim3(1), dim3(threads), args, lds, s);
  return hipLaunchCooperativeKernel(fn, dim3(grid), dim3(threads), args, lds, s);
}

// The errors with which the runtime refuses a cooperative launch it cannot
// make co-resident (too many workgroups, or the device busy with another
// process); any other launch error is a real failure.
inline bool cooperative_refusal(hipError_t e) {
  return e == hipErrorCooperativeLaunchTooLarge || e == hipErrorIllegalState;
}

// launch_grid for kernels whose work split and barriers follow gridDim
// alone: a refused cooperative launch is retried at half the width, down
// to one workgroup, whose plain launch needs no co-residency.  *grid is
// the width asked for on entry and the width launched on return; any
// other error is returned as is.
inline hipError_t launch_grid_shrinking(const void *fn, unsigned *grid, unsigned threads,
                                        void **args, size_t lds, hipStream_t s) {
  hipError_t r = launch_grid(fn, *grid, threads, args, lds, s);
  while (cooperative_refusal(r) && *grid > 1) {
    (void)hipGetLastError();
    *grid = *grid > 2 ? *grid / 2 : 1;
    r = launch_grid(fn, *grid, threads, args, lds, s);
  }
  return r;
}

inline unsigned grid_for(size_t work, unsigned block, unsigned cap = 65535u * 16u) {
  size_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace cds
