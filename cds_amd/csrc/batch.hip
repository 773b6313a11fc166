// Batched solves: many independent graphs as one disjoint union.
//
// The reference solves one graph per process (cudaCode/main.cu:865-1104).
// Here B graphs of the same k are laid side by side: graph b owns vertices
// [vbase_b, vbase_b + n_b) of one union CSR, so the DAG, the listing, the
// incidence and the level-synchronous peel run once over the union (clique
// cores never cross graphs: every clique lies inside one graph, and a
// vertex's core number depends only on its own graph).  The kernels below
// are the union-specific pieces: building the union, per-graph totals, the
// per-graph level tables, the per-graph densest-core flags and the
// per-graph counts of the result.
#include <hipcub/hipcub.hpp>

#include "dev.h"
#include "kernels.h"

namespace cds {
namespace {

// One workgroup per graph (grid-stride): offsets shifted by the graph's
// first adjacency slot, neighbours by its first vertex.
__global__ void k_union_build(uint32_t B, const UnionPart *__restrict__ parts,
                              uint32_t *__restrict__ off, uint32_t *__restrict__ adj,
                              uint32_t *__restrict__ gid) {
  for (uint32_t b = blockIdx.x; b < B; b += gridDim.x) {
    const UnionPart pt = parts[b];
    for (uint32_t i = threadIdx.x; i < pt.n; i += blockDim.x) {
      off[pt.vbase + i] = pt.abase + pt.off[i];
      gid[pt.vbase + i] = b;
    }
    const uint32_t na = pt.off[pt.n];
    for (uint32_t j = threadIdx.x; j < na; j += blockDim.x)
      adj[pt.abase + j] = pt.vbase + pt.adj[j];
  }
}

// Per-graph sum and maximum of the clique degrees (consecutive vertices
// share a graph, so a wave usually folds into one atomic pair).
__global__ void k_graph_cd(uint32_t n, const uint32_t *__restrict__ gid,
                           const uint64_t *__restrict__ cd,
                           unsigned long long *__restrict__ sum,
                           unsigned long long *__restrict__ mx) {
  for (uint32_t base = blockIdx.x * blockDim.x; base < n;
       base += gridDim.x * blockDim.x) {
    const uint32_t v = base + threadIdx.x;
    const bool in = v < n;
    const uint32_t b = in ? gid[v] : 0;
    const uint64_t x = in ? cd[v] : 0;
    const uint32_t b0 = __shfl(b, 0, kWave);
    const bool uniform = __all(!in || b == b0);
    if (uniform) {
      const uint64_t s = wave_sum64(x);
      uint64_t m = x;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const uint64_t t = __shfl_xor(m, o, kWave);
        m = t > m ? t : m;
      }
      if (lane() == 0) {
        atomicAdd(&sum[b0], (unsigned long long)s);
        atomicMax(&mx[b0], (unsigned long long)m);
      }
    } else if (in) {
      atomicAdd(&sum[b], (unsigned long long)x);
      atomicMax(&mx[b], (unsigned long long)x);
    }
  }
}

__global__ void k_level_keys(uint32_t n, const uint32_t *__restrict__ gid,
                             const uint32_t *__restrict__ core,
                             uint64_t *__restrict__ key,
                             uint32_t *__restrict__ val) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x) {
    key[v] = ((uint64_t)gid[v] << 32) | core[v];
    val[v] = v;
  }
}

// Index of `key` in the ascending array a[0..len) (it is present).
__device__ __forceinline__ uint32_t find_key(const uint64_t *a, uint32_t len,
                                             uint64_t key) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// killed[r] += 1 for every clique, r = run of (graph of the clique, its
// kill level).  Cliques of one listing item are consecutive, so lanes of a
// wave mostly share a run and reserve with one atomic.
__global__ void k_run_killed(uint64_t T, uint32_t k, const uint32_t *__restrict__ cl,
                             const uint32_t *__restrict__ kill,
                             const uint32_t *__restrict__ gid,
                             const uint64_t *__restrict__ ukeys, uint32_t nruns,
                             uint32_t *__restrict__ killed) {
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < T;
       base += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t c = base + threadIdx.x;
    const bool in = c < T;
    uint32_t r = 0;
    if (in) {
      const uint64_t key = ((uint64_t)gid[cl[c * k]] << 32) | kill[c];
      r = find_key(ukeys, nruns, key);
    }
    (void)aggregated_reserve(killed, r, in);
  }
}

__global__ void k_posmap(uint32_t n, const uint32_t *__restrict__ sval,
                         uint32_t *__restrict__ posmap) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    posmap[sval[i]] = i;
}

// Densest-core flags: 1 if the vertex's core (the clique's kill level)
// reaches its graph's k'.
__global__ void k_core_flags01(uint32_t n, const uint32_t *__restrict__ gid,
                               const uint32_t *__restrict__ core,
                               const uint32_t *__restrict__ kp,
                               uint32_t *__restrict__ out) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x)
    out[v] = core[v] >= kp[gid[v]] ? 1u : 0u;
}

__global__ void k_kill_flags01(uint64_t T, uint32_t k, const uint32_t *__restrict__ cl,
                               const uint32_t *__restrict__ gid,
                               const uint32_t *__restrict__ kill,
                               const uint32_t *__restrict__ kp,
                               uint32_t *__restrict__ out) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
       c += (uint64_t)gridDim.x * blockDim.x)
    out[c] = kill[c] >= kp[gid[cl[c * k]]] ? 1u : 0u;
}

// Per graph: adjacency entries between flagged vertices (core edges x 2)
// and marked adjacency entries (pruned edges x 2).
__global__ void k_graph_edges(uint32_t n, const uint32_t *__restrict__ gid,
                              const uint32_t *__restrict__ off,
                              const uint32_t *__restrict__ adj,
                              const uint32_t *__restrict__ flag,
                              const uint8_t *__restrict__ mark,
                              unsigned long long *__restrict__ core_e,
                              unsigned long long *__restrict__ kept_e) {
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nw = gridDim.x * blockDim.x / kWave;
  for (uint32_t v = gw; v < n; v += nw) {
    uint64_t ce = 0, ke = 0;
    const bool fv = flag[v] != 0;
    for (uint32_t p = off[v] + lane(); p < off[v + 1]; p += kWave) {
      if (fv && flag[adj[p]]) ++ce;
      if (mark && mark[p]) ++ke;
    }
    ce = wave_sum64(ce);
    ke = wave_sum64(ke);
    if (lane() == 0) {
      if (ce) atomicAdd(&core_e[gid[v]], (unsigned long long)ce);
      if (ke) atomicAdd(&kept_e[gid[v]], (unsigned long long)ke);
    }
  }
}

// Per graph: cliques whose members all have mask != 0.
__global__ void k_graph_cliques_in_mask(uint64_t T, uint32_t k,
                                        const uint32_t *__restrict__ cl,
                                        const uint32_t *__restrict__ gid,
                                        const uint8_t *__restrict__ mask,
                                        uint32_t *__restrict__ out) {
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < T;
       base += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t c = base + threadIdx.x;
    bool all = c < T;
    if (all)
      for (uint32_t j = 0; j < k; ++j) all &= mask[cl[c * k + j]] != 0;
    const uint32_t b = all ? gid[cl[c * k]] : 0;
    (void)aggregated_reserve(out, b, all);
  }
}

}  // namespace

hipError_t union_build(uint32_t B, const UnionPart *parts, uint32_t n_total,
                       uint32_t a_total, uint32_t *off, uint32_t *adj,
                       uint32_t *gid, hipStream_t s) {
  hipLaunchKernelGGL(k_union_build, dim3(B < 8192 ? B : 8192), dim3(256), 0, s, B,
                     parts, off, adj, gid);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(off + n_total),
                           (int)a_total, 1, s);
}

hipError_t graph_clique_degrees(uint32_t n, const uint32_t *gid, const uint64_t *cd,
                                unsigned long long *sum, unsigned long long *mx,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_graph_cd, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n,
                     gid, cd, sum, mx);
  return hipGetLastError();
}

hipError_t level_runs(uint32_t n, const uint32_t *gid, const uint32_t *core,
                      uint64_t *key, uint32_t *val, uint64_t *skey, uint32_t *sval,
                      uint64_t *ukey, uint32_t *counts, uint32_t *nruns,
                      DeviceArena &arena, hipStream_t s) {
  hipLaunchKernelGGL(k_level_keys, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s,
                     n, gid, core, key, val);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t b1 = 0, b2 = 0;
  if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, b1, key, skey, val, sval,
                                              (int)n, 0, 64, s)) != hipSuccess)
    return e;
  if ((e = hipcub::DeviceRunLengthEncode::Encode(nullptr, b2, skey, ukey, counts,
                                                 nruns, (int)n, s)) != hipSuccess)
    return e;
  void *tmp = nullptr;
  const size_t bytes = b1 > b2 ? b1 : b2;
  if ((e = arena.alloc(&tmp, bytes ? bytes : 1)) != hipSuccess) return e;
  e = hipcub::DeviceRadixSort::SortPairs(tmp, b1, key, skey, val, sval, (int)n, 0,
                                         64, s);
  if (e == hipSuccess)
    e = hipcub::DeviceRunLengthEncode::Encode(tmp, b2, skey, ukey, counts, nruns,
                                              (int)n, s);
  arena.release(tmp);
  return e;
}

hipError_t run_killed(uint64_t T, uint32_t k, const uint32_t *cl,
                      const uint32_t *kill, const uint32_t *gid,
                      const uint64_t *ukeys, uint32_t nruns, uint32_t *killed,
                      hipStream_t s) {
  hipError_t e = hipMemsetAsync(killed, 0, (size_t)nruns * sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_run_killed, dim3(grid_for(T, 256, 16384)), dim3(256), 0, s,
                     T, k, cl, kill, gid, ukeys, nruns, killed);
  return hipGetLastError();
}

hipError_t positions_of(uint32_t n, const uint32_t *sval, uint32_t *posmap,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_posmap, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n,
                     sval, posmap);
  return hipGetLastError();
}

hipError_t densest_flags(uint32_t n, uint64_t T, uint32_t k, const uint32_t *cl,
                         const uint32_t *gid, const uint32_t *core,
                         const uint32_t *kill, const uint32_t *kp,
                         uint32_t *core01, uint32_t *kill01, hipStream_t s) {
  hipLaunchKernelGGL(k_core_flags01, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s,
                     n, gid, core, kp, core01);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_kill_flags01, dim3(grid_for(T, 256, 16384)), dim3(256), 0,
                     s, T, k, cl, gid, kill, kp, kill01);
  return hipGetLastError();
}

hipError_t graph_edge_counts(uint32_t n, const uint32_t *gid, const uint32_t *off,
                             const uint32_t *adj, const uint32_t *flag,
                             const uint8_t *mark, unsigned long long *core_e,
                             unsigned long long *kept_e, hipStream_t s) {
  hipLaunchKernelGGL(k_graph_edges, dim3(grid_for((size_t)n * kWave, 256, 8192)),
                     dim3(256), 0, s, n, gid, off, adj, flag, mark, core_e, kept_e);
  return hipGetLastError();
}

hipError_t graph_cliques_in_mask(uint64_t T, uint32_t k, const uint32_t *cl,
                                 const uint32_t *gid, const uint8_t *mask,
                                 uint32_t *out, hipStream_t s) {
  hipLaunchKernelGGL(k_graph_cliques_in_mask, dim3(grid_for(T, 256, 16384)),
                     dim3(256), 0, s, T, k, cl, gid, mask, out);
  return hipGetLastError();
}

}  // namespace cds
