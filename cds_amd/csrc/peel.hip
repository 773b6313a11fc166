// (b) k-clique-core peeling.
//
// Reference GPU path: cliqueCoreDecompose (cudaCode/main.cu:252-343) with
// selectNodes / processNodesByWarp (cudaCode/src/helpers.cu:611-750), which
// scans all T cliques per removed vertex.  Oracle: cliqueCoreDecompose
// (cpuCode/c++/src/helpers.cpp:16-150), a sequential removal of the
// (current clique degree, id)-minimum vertex.
//
// Here: a vertex -> clique incidence CSR is built once; the peel is
// level-synchronous and frontier-driven.  For level l, every alive vertex
// with degree <= l joins the frontier; each frontier vertex walks its
// incidence list, the first remover of a clique claims it with one CAS
// (kill[c] = l), and the claimer decrements the other members; the member
// whose decrement crosses l+1 -> l joins the next sub-frontier.  The result
// is the exact clique-core number of every vertex (order independent) and
// kill[c] = min core over c's members.
//
// The oracle's max peel-prefix density depends on its exact removal order
// only inside levels above ceil(max level-boundary density) (the engine
// proves which); replay_phase re-runs those levels in the oracle's order.
#include "dev.h"
#include "kernels.h"

namespace cds {
namespace {

__global__ void k_build_inc(uint64_t T, uint32_t k,
                            const uint32_t *__restrict__ cl,
                            const uint64_t *__restrict__ ioff,
                            uint32_t *__restrict__ icur,
                            uint32_t *__restrict__ inc) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
       c += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t j = 0; j < k; ++j) {
      const uint32_t x = cl[c * k + j];
      const uint32_t pos = atomicAdd(&icur[x], 1u);
      inc[ioff[x] + pos] = (uint32_t)c;
    }
}

__global__ void k_peel_init_v(uint32_t n, const uint64_t *__restrict__ cd,
                              uint32_t *__restrict__ deg,
                              uint32_t *__restrict__ core) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x) {
    deg[v] = (uint32_t)cd[v];
    core[v] = kAlive;
  }
}

__global__ void k_fill_u32(uint64_t n, uint32_t *__restrict__ a, uint32_t val) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = val;
}

__global__ void k_min_alive(uint32_t n, const uint32_t *__restrict__ deg,
                            const uint32_t *__restrict__ core, uint32_t *out) {
  uint32_t m = kAlive;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x)
    if (core[v] == kAlive) m = min(m, deg[v]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = min(m, (uint32_t)__shfl_xor(m, off, kWave));
  if (lane() == 0 && m != kAlive) atomicMin(out, m);
}

__global__ void k_frontier(uint32_t n, uint32_t level,
                           const uint32_t *__restrict__ deg,
                           uint32_t *__restrict__ core,
                           uint32_t *__restrict__ frontier, uint32_t *count) {
  for (uint32_t base = blockIdx.x * blockDim.x; base < n;
       base += gridDim.x * blockDim.x) {
    const uint32_t v = base + threadIdx.x;
    const bool take = v < n && core[v] == kAlive && deg[v] <= level;
    unsigned cnt;
    const unsigned slot = wave_slot(take, &cnt);
    uint32_t start = 0;
    if (lane() == 0 && cnt) start = atomicAdd(count, cnt);
    start = __shfl(start, 0, kWave);
    if (take) {
      core[v] = level;
      frontier[start + slot] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_peel_process(
    uint32_t k, uint32_t level, const uint32_t *__restrict__ frontier,
    const uint32_t *__restrict__ fcount, const uint64_t *__restrict__ ioff,
    const uint32_t *__restrict__ inc, const uint32_t *__restrict__ cl,
    uint32_t *__restrict__ deg, uint32_t *__restrict__ core,
    uint32_t *__restrict__ kill, uint32_t *__restrict__ next,
    uint32_t *__restrict__ ncount, unsigned long long *__restrict__ killed) {
  const uint32_t nf = *fcount;
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  uint32_t nk = 0;
  for (uint32_t f = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       f < nf; f += waves) {
    const uint32_t v = frontier[f];
    const uint64_t e = ioff[v + 1];
    for (uint64_t q = ioff[v] + lane(); q < e; q += kWave) {
      const uint32_t c = inc[q];
      if (atomicCAS(&kill[c], kAlive, level) != kAlive) continue;
      ++nk;
      for (uint32_t t = 0; t < k; ++t) {
        const uint32_t u = cl[(uint64_t)c * k + t];
        if (u == v || core[u] != kAlive) continue;
        const uint32_t old = atomicSub(&deg[u], 1u);
        if (old == level + 1) {
          core[u] = level;
          next[atomicAdd(ncount, 1u)] = u;
        }
      }
    }
  }
  nk = wave_sum(nk);
  if (lane() == 0 && nk) atomicAdd(killed, (unsigned long long)nk);
}

// ---- exact replay of one level in the oracle's removal order ----

__global__ void k_collect_level(uint32_t n, uint32_t level,
                                const uint32_t *__restrict__ core,
                                uint32_t *__restrict__ pv,
                                uint32_t *__restrict__ posmap,
                                uint32_t *count) {
  for (uint32_t base = blockIdx.x * blockDim.x; base < n;
       base += gridDim.x * blockDim.x) {
    const uint32_t v = base + threadIdx.x;
    const bool take = v < n && core[v] == level;
    unsigned cnt;
    const unsigned slot = wave_slot(take, &cnt);
    uint32_t start = 0;
    if (lane() == 0 && cnt) start = atomicAdd(count, cnt);
    start = __shfl(start, 0, kWave);
    if (take) {
      pv[start + slot] = v;
      posmap[v] = start + slot;
    }
  }
}

__global__ __launch_bounds__(256) void k_phase_deg(
    uint32_t level, const uint32_t *__restrict__ count,
    const uint32_t *__restrict__ pv, const uint64_t *__restrict__ ioff,
    const uint32_t *__restrict__ inc, const uint32_t *__restrict__ kill,
    uint32_t *__restrict__ pdeg, uint8_t *__restrict__ palive) {
  const uint32_t M = *count;
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       i < M; i += waves) {
    const uint32_t v = pv[i];
    uint32_t d = 0;
    for (uint64_t q = ioff[v] + lane(); q < ioff[v + 1]; q += kWave)
      d += kill[inc[q]] == level;
    d = wave_sum(d);
    if (lane() == 0) {
      pdeg[i] = d;
      palive[i] = 1;
    }
  }
}

// a/b > c/d for b, d > 0, exact (128-bit products).
__device__ __forceinline__ bool frac_gt(uint64_t a, uint64_t b, uint64_t c,
                                        uint64_t d) {
  const uint64_t h1 = __umul64hi(a, d), l1 = a * d;
  const uint64_t h2 = __umul64hi(c, b), l2 = c * b;
  return h1 > h2 || (h1 == h2 && l1 > l2);
}

constexpr int kReplayThreads = 1024;

// One workgroup replays level `level` starting from the state C_level
// (T_start cliques, N_start vertices): repeatedly removes the alive phase
// vertex with the smallest (current degree, id), kills its alive phase
// cliques and decrements their other phase members; records the largest
// density of the intermediate states (after 1..M-1 removals).
__global__ __launch_bounds__(kReplayThreads) void k_replay(
    uint32_t k, uint32_t level, const uint32_t *__restrict__ count,
    const uint32_t *__restrict__ pv, uint32_t *__restrict__ pdeg,
    const uint32_t *__restrict__ posmap, uint8_t *__restrict__ palive,
    const uint32_t *__restrict__ core, const uint32_t *__restrict__ kill,
    const uint64_t *__restrict__ ioff, const uint32_t *__restrict__ inc,
    const uint32_t *__restrict__ cl, uint32_t *__restrict__ rdead,
    uint64_t T_start, uint64_t N_start, ReplayOut *out) {
  __shared__ uint64_t s_wmin[kReplayThreads / kWave];
  __shared__ uint64_t s_key;
  const uint32_t M = *count;
  const uint32_t tid = threadIdx.x, wid = tid / kWave;
  uint64_t R = 0, best_num = 0, best_den = 1;
  uint32_t found = 0, steps = 0;
  for (uint32_t j = 0; j < M; ++j) {
    uint64_t key = ~0ull;
    for (uint32_t i = tid; i < M; i += kReplayThreads)
      if (palive[i]) {
        const uint64_t kk = ((uint64_t)pdeg[i] << 32) | pv[i];
        key = kk < key ? kk : key;
      }
    key = wave_min64(key);
    if (lane() == 0) s_wmin[wid] = key;
    __syncthreads();
    if (tid == 0) {
      uint64_t m = s_wmin[0];
      for (int w = 1; w < kReplayThreads / kWave; ++w)
        m = s_wmin[w] < m ? s_wmin[w] : m;
      s_key = m;
    }
    __syncthreads();
    const uint64_t kk = s_key;
    if (kk == ~0ull) break;  // inconsistent counts; reported via steps
    const uint32_t v = (uint32_t)kk;
    if (tid == 0) {
      palive[posmap[v]] = 0;
      R += kk >> 32;
      ++steps;
      if (j + 1 < M) {
        const uint64_t num = T_start - R, den = N_start - (j + 1);
        if (!found || frac_gt(num, den, best_num, best_den)) {
          best_num = num;
          best_den = den;
          found = 1;
        }
      }
    }
    const uint64_t e = ioff[v + 1];
    for (uint64_t q = ioff[v] + tid; q < e; q += kReplayThreads) {
      const uint32_t c = inc[q];
      if (kill[c] != level || rdead[c] == level + 1) continue;
      rdead[c] = level + 1;
      for (uint32_t t = 0; t < k; ++t) {
        const uint32_t u = cl[(uint64_t)c * k + t];
        if (u != v && core[u] == level) atomicSub(&pdeg[posmap[u]], 1u);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    out->best_num = best_num;
    out->best_den = best_den;
    out->steps = steps;
    out->found = found;
  }
}

}  // namespace

hipError_t build_incidence(uint64_t T, uint32_t k, const uint32_t *cliques,
                           const uint64_t *ioff, uint32_t *icur,
                           uint32_t *inc, hipStream_t s) {
  hipLaunchKernelGGL(k_build_inc, dim3(grid_for(T, 256, 65535)), dim3(256), 0,
                     s, T, k, cliques, ioff, icur, inc);
  return hipGetLastError();
}

hipError_t peel_init(uint32_t n, uint64_t T, const uint64_t *cd,
                     uint32_t *deg, uint32_t *core, uint32_t *kill,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_peel_init_v, dim3(grid_for(n, 256, 65535)), dim3(256),
                     0, s, n, cd, deg, core);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(T, 256, 65535)), dim3(256), 0,
                     s, T, kill, kAlive);
  return hipGetLastError();
}

hipError_t peel_min_alive(uint32_t n, const uint32_t *deg,
                          const uint32_t *core, uint32_t *out, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, 1ull, out, kAlive);
  hipLaunchKernelGGL(k_min_alive, dim3(grid_for(n, 256, 4096)), dim3(256), 0,
                     s, n, deg, core, out);
  return hipGetLastError();
}

hipError_t peel_frontier(uint32_t n, uint32_t level, const uint32_t *deg,
                         uint32_t *core, uint32_t *frontier, uint32_t *count,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_frontier, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s,
                     n, level, deg, core, frontier, count);
  return hipGetLastError();
}

hipError_t peel_process(uint32_t k, uint32_t level, const uint32_t *frontier,
                        const uint32_t *fcount, uint32_t fmax,
                        const uint64_t *ioff, const uint32_t *inc,
                        const uint32_t *cliques, uint32_t *deg,
                        uint32_t *core, uint32_t *kill, uint32_t *next,
                        uint32_t *ncount, unsigned long long *killed,
                        hipStream_t s) {
  const unsigned grid = grid_for((size_t)(fmax ? fmax : 1) * kWave, 256, 16384);
  hipLaunchKernelGGL(k_peel_process, dim3(grid), dim3(256), 0, s, k, level,
                     frontier, fcount, ioff, inc, cliques, deg, core, kill,
                     next, ncount, killed);
  return hipGetLastError();
}

hipError_t replay_phase(uint32_t k, uint32_t level, uint32_t n,
                        const uint32_t *core, const uint32_t *kill,
                        const uint64_t *ioff, const uint32_t *inc,
                        const uint32_t *cliques, uint64_t T_start,
                        uint64_t N_start, uint32_t M, uint32_t *pv,
                        uint32_t *pdeg, uint32_t *posmap, uint8_t *palive,
                        uint32_t *rdead, uint32_t *pcount, ReplayOut *out,
                        hipStream_t s) {
  hipError_t e = hipMemsetAsync(pcount, 0, sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_collect_level, dim3(grid_for(n, 256, 8192)), dim3(256),
                     0, s, n, level, core, pv, posmap, pcount);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_phase_deg, dim3(grid_for((size_t)M * kWave, 256, 8192)),
                     dim3(256), 0, s, level, pcount, pv, ioff, inc, kill, pdeg,
                     palive);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_replay, dim3(1), dim3(kReplayThreads), 0, s, k, level,
                     pcount, pv, pdeg, posmap, palive, core, kill, ioff, inc,
                     cliques, rdead, T_start, N_start, out);
  return hipGetLastError();
}

}  // namespace cds
