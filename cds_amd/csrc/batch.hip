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

// ---- (graph, core) level runs: a stable LSD radix sort of 64-bit keys
// with 32-bit values, 8-bit digits over the key bits that vary (the low
// bits of the core word, then of the graph word), then run-length codes.
constexpr int kRsThreads = 256, kRsRounds = 8;
constexpr uint32_t kRsTile = kRsThreads * kRsRounds;  // items per workgroup
constexpr int kRsWaves = kRsThreads / kWave;

__device__ __forceinline__ uint32_t rs_digit(uint64_t key, uint32_t shift) {
  return (uint32_t)(key >> shift) & 255u;
}

// hist[d * nb + b] = items of tile b whose digit is d.
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(uint32_t n,
                                                        const uint64_t *__restrict__ key,
                                                        uint32_t shift,
                                                        uint32_t *__restrict__ hist) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t t0 = blockIdx.x * kRsTile;
  for (uint32_t i = t0 + threadIdx.x; i < n && i < t0 + kRsTile; i += kRsThreads)
    atomicAdd(&cnt[rs_digit(key[i], shift)], 1u);
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter of tile b to the positions the scanned histogram gives:
// items are taken in index order, 256 per round; an item's place among the
// round's items of its digit is its rank among the equal-digit lanes of its
// wave (a match over the digit's 8 bits) plus the counts of earlier waves.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    uint32_t n, const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, uint32_t shift,
    const uint32_t *__restrict__ hoff) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[kRsWaves][256];
  const uint32_t w = threadIdx.x / kWave;
  base[threadIdx.x] = hoff[threadIdx.x * gridDim.x + blockIdx.x];
  for (int x = 0; x < kRsWaves; ++x) wcnt[x][threadIdx.x] = 0;
  __syncthreads();
  const uint32_t t0 = blockIdx.x * kRsTile;
  for (int r = 0; r < kRsRounds; ++r) {
    const uint32_t i = t0 + (uint32_t)r * kRsThreads + threadIdx.x;
    const bool in = i < n;
    const uint64_t kv = in ? kin[i] : 0;
    const uint32_t d = rs_digit(kv, shift);
    uint64_t peers = __ballot(in);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t rank = __popcll(peers & lanes_below());
    if (in && rank == 0) wcnt[w][d] = __popcll(peers);
    __syncthreads();
    if (in) {
      uint32_t pos = base[d] + rank;
      for (uint32_t x = 0; x < w; ++x) pos += wcnt[x][d];
      kout[pos] = kv;
      vout[pos] = vin[i];
    }
    __syncthreads();
    uint32_t add = 0;
    for (int x = 0; x < kRsWaves; ++x) {
      add += wcnt[x][threadIdx.x];
      wcnt[x][threadIdx.x] = 0;
    }
    base[threadIdx.x] += add;
    __syncthreads();
  }
}

// Largest graph id and core: m[0] = max gid, m[1] = max core (m zeroed).
__global__ void k_key_max(uint32_t n, const uint32_t *__restrict__ gid,
                          const uint32_t *__restrict__ core, uint32_t *m) {
  uint32_t a = 0, b = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x) {
    a = max(a, gid[v]);
    b = max(b, core[v]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = max(a, (uint32_t)__shfl_xor(a, off, kWave));
    b = max(b, (uint32_t)__shfl_xor(b, off, kWave));
  }
  if (lane() == 0) {
    if (a) atomicMax(&m[0], a);
    if (b) atomicMax(&m[1], b);
  }
}

// head[i] = 1 where a run of equal sorted keys starts.
__global__ void k_run_heads(uint32_t n, const uint64_t *__restrict__ skey,
                            uint32_t *__restrict__ head) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    head[i] = (i == 0 || skey[i] != skey[i - 1]) ? 1u : 0u;
}

// Run r (idx = exclusive scan of head, idx[n] = runs): its key and its
// first position; starts[runs] = n.
__global__ void k_run_write(uint32_t n, const uint64_t *__restrict__ skey,
                            const uint32_t *__restrict__ head,
                            const uint32_t *__restrict__ idx, uint64_t *__restrict__ ukey,
                            uint32_t *__restrict__ starts, uint32_t *__restrict__ nruns) {
  const uint32_t runs = idx[n];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    if (!head[i]) continue;
    ukey[idx[i]] = skey[i];
    starts[idx[i]] = i;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *nruns = runs;
    starts[runs] = n;
  }
}

__global__ void k_run_lengths(const uint32_t *__restrict__ starts,
                              const uint32_t *__restrict__ nruns,
                              uint32_t *__restrict__ counts) {
  const uint32_t runs = *nruns;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < runs;
       r += gridDim.x * blockDim.x)
    counts[r] = starts[r + 1] - starts[r];
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
  if (n == 0) return hipMemsetAsync(nruns, 0, sizeof(uint32_t), s);
  hipLaunchKernelGGL(k_level_keys, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s,
                     n, gid, core, key, val);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the key bits that vary: low bits of the core, then of the graph id
  uint32_t *mx = nullptr;
  if ((e = arena.alloc((void **)&mx, 2 * sizeof(uint32_t))) != hipSuccess) return e;
  uint32_t hm[2] = {0, 0};
  e = hipMemsetAsync(mx, 0, 2 * sizeof(uint32_t), s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_key_max, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, n, gid,
                       core, mx);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hm, mx, sizeof(hm), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  arena.release(mx);
  if (e != hipSuccess) return e;
  auto bits = [](uint32_t x) {
    uint32_t b = 0;
    while (b < 32 && (x >> b)) ++b;
    return b;
  };
  std::vector<uint32_t> shifts;
  for (uint32_t sh = 0; sh < bits(hm[1]); sh += 8) shifts.push_back(sh);
  for (uint32_t sh = 0; sh < bits(hm[0]); sh += 8) shifts.push_back(32 + sh);
  const uint32_t nb = (n + kRsTile - 1) / kRsTile;
  uint32_t *hist = nullptr, *hoff = nullptr;
  if ((e = arena.alloc((void **)&hist, (size_t)256 * nb * sizeof(uint32_t))) != hipSuccess)
    return e;
  if ((e = arena.alloc((void **)&hoff, ((size_t)256 * nb + 1) * sizeof(uint32_t))) !=
      hipSuccess) {
    arena.release(hist);
    return e;
  }
  // ping-pong between (key, val) and (skey, sval); an even pass count
  // starts by copying so the last pass lands in (skey, sval)
  uint64_t *ka = key, *kb = skey;
  uint32_t *va = val, *vb = sval;
  if (shifts.size() % 2 == 0) {
    e = hipMemcpyAsync(skey, key, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(sval, val, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  for (size_t pi = 0; pi < shifts.size() && e == hipSuccess; ++pi) {
    hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kRsThreads), 0, s, n, ka, shifts[pi], hist);
    e = hipGetLastError();
    if (e == hipSuccess) e = scan_u32_32(hist, hoff, (size_t)256 * nb, arena);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kRsThreads), 0, s, n, ka, va, kb, vb,
                         shifts[pi], hoff);
      e = hipGetLastError();
    }
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  // runs of equal keys
  uint32_t *head = nullptr, *idx = nullptr, *starts = nullptr;
  if (e == hipSuccess) e = arena.alloc((void **)&head, (size_t)n * sizeof(uint32_t));
  if (e == hipSuccess) e = arena.alloc((void **)&idx, ((size_t)n + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = arena.alloc((void **)&starts, ((size_t)n + 1) * sizeof(uint32_t));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_run_heads, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, skey,
                       head);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = scan_u32_32(head, idx, n, arena);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_run_write, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, skey,
                       head, idx, ukey, starts, nruns);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_run_lengths, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, starts,
                       nruns, counts);
    e = hipGetLastError();
  }
  arena.release(starts);
  arena.release(idx);
  arena.release(head);
  arena.release(hoff);
  arena.release(hist);
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
