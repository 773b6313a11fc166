// Core pruning before materialising cliques, and the exact peel of a small
// dense core on adjacency bitsets.
//
// A graph with too many k-cliques to materialise (k T slots beyond 32 bits
// or beyond the device) is first cut to its tau-clique-core S: count
// passes over induced subgraphs drop every vertex with fewer than tau
// cliques until none is left to drop, and tau rises to the ceiling of the
// density of S while that grows (each S is a boundary state of the peel,
// so its density is a peel row and tau never passes k').  The rows of the
// peel above S (vertices with core numbers below tau) have densities at
// most max(rho(S), tau - 1): each removed one of them with at most tau - 1
// cliques.  So k' = ceil(max row density) is decided inside S.
//
// When S still holds too many cliques for the incidence of the persistent
// peel but has few vertices (a dense core), the peel of S runs on adjacency
// bitsets instead: the oracle's sequential removal of the vertex with the
// smallest (clique degree, id) (cpuCode/c++/src/helpers.cpp:16-150), one
// removal at a time, where removing v takes from each neighbour u the
// number of (k-2)-cliques in N(u) & N(v) & alive (1 for k = 2, a popcount
// for k = 3, the edges inside that set for k = 4).  Two launches per
// removal: a one-workgroup argmin that records the row, then one workgroup
// per vertex for the decrements.
#include "dev.h"
#include "kernels.h"

namespace cds {

namespace {

__global__ void k_keep_u32(uint32_t n, const uint8_t *__restrict__ keep,
                           uint32_t *__restrict__ k32) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x)
    k32[v] = keep[v] ? 1u : 0u;
}

// Kept neighbours of each kept vertex, one wave per vertex.
__global__ void k_induced_count(uint32_t n, const uint32_t *__restrict__ off,
                                const uint32_t *__restrict__ adj,
                                const uint8_t *__restrict__ keep,
                                const uint32_t *__restrict__ newid,
                                uint32_t *__restrict__ cnt) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; v < n; v += waves) {
    if (!keep[v]) continue;
    uint32_t c = 0;
    for (uint32_t p = off[v] + lane(); p < off[v + 1]; p += kWave) c += keep[adj[p]] ? 1u : 0u;
    c = wave_sum(c);
    if (lane() == 0) cnt[newid[v]] = c;
  }
}

// Kept neighbours in new ids, in adjacency order (ascending: new ids keep
// the order of old ones), one wave per vertex.
__global__ void k_induced_fill(uint32_t n, const uint32_t *__restrict__ off,
                               const uint32_t *__restrict__ adj,
                               const uint8_t *__restrict__ keep,
                               const uint32_t *__restrict__ newid,
                               const uint32_t *__restrict__ off2,
                               uint32_t *__restrict__ adj2, uint32_t *__restrict__ sel) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; v < n; v += waves) {
    if (!keep[v]) continue;
    const uint32_t nv = newid[v];
    if (lane() == 0) sel[nv] = v;
    uint32_t at = off2[nv];
    for (uint32_t p0 = off[v]; p0 < off[v + 1]; p0 += kWave) {
      const uint32_t p = p0 + lane();
      const bool in = p < off[v + 1] && keep[adj[p]];
      unsigned cnt = 0;
      const unsigned slot = wave_slot(in, &cnt);
      if (in) adj2[at + slot] = newid[adj[p]];
      at += cnt;
    }
  }
}

// keep[sel[i]] = 0 for the induced vertices i with fewer than tau cliques.
__global__ void k_drop_below(uint32_t ns, const uint64_t *__restrict__ cd, uint64_t tau,
                             const uint32_t *__restrict__ sel, uint8_t *__restrict__ keep,
                             uint32_t *__restrict__ dropped) {
  uint32_t d = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns;
       i += gridDim.x * blockDim.x)
    if (cd[i] < tau) {
      keep[sel[i]] = 0;
      ++d;
    }
  d = wave_sum(d);
  if (lane() == 0 && d) atomicAdd(dropped, d);
}

// Adjacency bitsets: row v holds W 64-bit words.
__global__ void k_bitset_rows(uint32_t ns, const uint32_t *__restrict__ off,
                              const uint32_t *__restrict__ adj, uint32_t W,
                              unsigned long long *__restrict__ bits) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; v < ns; v += waves) {
    unsigned long long *row = bits + (size_t)v * W;
    for (uint32_t w = lane(); w < W; w += kWave) row[w] = 0;
    for (uint32_t p = off[v] + lane(); p < off[v + 1]; p += kWave) {
      const uint32_t u = adj[p];
      atomicOr(&row[u >> 6], 1ull << (u & 63));
    }
  }
}

__global__ void k_alive_init(uint32_t ns, uint32_t W, unsigned long long *__restrict__ alive) {
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < W;
       w += gridDim.x * blockDim.x) {
    const uint32_t lo = w * 64;
    alive[w] = ns >= lo + 64 ? ~0ull : (ns > lo ? (1ull << (ns - lo)) - 1 : 0ull);
  }
}

constexpr int kDenseThreads = 1024;

// Removal i: the alive vertex of smallest (cur, id); records it and its
// current clique count, and takes it out of `alive`.
__global__ __launch_bounds__(kDenseThreads) void k_dense_argmin(
    uint32_t ns, uint32_t i, const unsigned long long *__restrict__ cur,
    unsigned long long *__restrict__ alive, uint32_t *__restrict__ step_v,
    uint32_t *__restrict__ rv, unsigned long long *__restrict__ rval) {
  constexpr uint32_t kNone = 0xFFFFFFFFu;
  __shared__ uint32_t s_id[kDenseThreads / kWave];
  __shared__ unsigned long long s_val[kDenseThreads / kWave];
  unsigned long long bval = 0;
  uint32_t bid = kNone;
  for (uint32_t v = threadIdx.x; v < ns; v += kDenseThreads) {
    if (!((alive[v >> 6] >> (v & 63)) & 1ull)) continue;
    const unsigned long long c = cur[v];
    if (bid == kNone || c < bval) {  // ids ascend in this loop: first wins ties
      bval = c;
      bid = v;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long ov = __shfl_xor(bval, o, kWave);
    const uint32_t oi = __shfl_xor(bid, o, kWave);
    if (oi != kNone && (bid == kNone || ov < bval || (ov == bval && oi < bid))) {
      bval = ov;
      bid = oi;
    }
  }
  const uint32_t wid = threadIdx.x / kWave;
  if (lane() == 0) {
    s_id[wid] = bid;
    s_val[wid] = bval;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t v = kNone;
    unsigned long long c = 0;
    for (int w = 0; w < kDenseThreads / kWave; ++w) {
      const uint32_t ow = s_id[w];
      if (ow == kNone) continue;
      if (v == kNone || s_val[w] < c || (s_val[w] == c && ow < v)) {
        v = ow;
        c = s_val[w];
      }
    }
    step_v[0] = v;
    rv[i] = v;
    rval[i] = c;
    if (v != kNone) alive[v >> 6] &= ~(1ull << (v & 63));
  }
}

// Decrements after removal of v = *step_v: workgroup u takes from cur[u]
// the cliques holding both u and v among the alive vertices.
__global__ __launch_bounds__(256) void k_dense_dec(
    uint32_t ns, uint32_t W, uint32_t k, const unsigned long long *__restrict__ bits,
    const unsigned long long *__restrict__ alive, const uint32_t *__restrict__ step_v,
    unsigned long long *__restrict__ cur) {
  __shared__ unsigned long long C[kDenseMaxVertices / 64];
  __shared__ unsigned long long s_part[256 / kWave];
  const uint32_t v = *step_v;
  for (uint32_t u = blockIdx.x; u < ns; u += gridDim.x) {
    if (v == 0xFFFFFFFFu || u == v) continue;
    if (!((alive[u >> 6] >> (u & 63)) & 1ull)) continue;
    const unsigned long long *Nu = bits + (size_t)u * W, *Nv = bits + (size_t)v * W;
    if (!((Nu[v >> 6] >> (v & 63)) & 1ull)) continue;  // u not adjacent to v
    unsigned long long part = 0;
    if (k == 2) {
      part = threadIdx.x == 0 ? 1 : 0;
    } else {
      for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) {
        C[w] = Nu[w] & Nv[w] & alive[w];
        if (k == 3) part += __popcll(C[w]);
      }
      if (k == 4) {
        __syncthreads();
        // edges inside C: each w in C counts |N(w) & C|, halved below
        for (uint32_t w = threadIdx.x; w < ns; w += blockDim.x) {
          if (!((C[w >> 6] >> (w & 63)) & 1ull)) continue;
          const unsigned long long *Nw = bits + (size_t)w * W;
          for (uint32_t x = 0; x < W; ++x) part += __popcll(Nw[x] & C[x]);
        }
      }
    }
    part = wave_sum64(part);
    if (lane() == 0) s_part[threadIdx.x / kWave] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long d = 0;
      for (int w = 0; w < 256 / kWave; ++w) d += s_part[w];
      if (k == 4) d >>= 1;
      cur[u] -= d;
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t induced_subgraph(uint32_t n, const uint32_t *off, const uint32_t *adj,
                            const uint8_t *keep, uint32_t *newid, uint32_t *cnt,
                            uint32_t *off2, uint32_t *adj2, uint32_t *sel, bool fill,
                            DeviceArena &a) {
  hipStream_t s = a.stream();
  if (!fill) {
    hipLaunchKernelGGL(k_keep_u32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, keep,
                       cnt);
    hipError_t e = scan_u32_32(cnt, newid, n, a);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_induced_count, dim3(grid_for((size_t)n * kWave, 256, 16384)),
                       dim3(256), 0, s, n, off, adj, keep, newid, cnt);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_induced_fill, dim3(grid_for((size_t)n * kWave, 256, 16384)),
                     dim3(256), 0, s, n, off, adj, keep, newid, off2, adj2, sel);
  return hipGetLastError();
}

hipError_t drop_below(uint32_t ns, const uint64_t *cd, uint64_t tau, const uint32_t *sel,
                      uint8_t *keep, uint32_t *dropped, hipStream_t s) {
  hipError_t e = hipMemsetAsync(dropped, 0, sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_drop_below, dim3(grid_for(ns, 256, 4096)), dim3(256), 0, s, ns, cd,
                     tau, sel, keep, dropped);
  return hipGetLastError();
}

hipError_t dense_peel(uint32_t ns, uint32_t k, const uint32_t *off, const uint32_t *adj,
                      const uint64_t *cd, unsigned long long *bits, unsigned long long *alive,
                      unsigned long long *cur, uint32_t *step_v, uint32_t *rv,
                      unsigned long long *rval, hipStream_t s) {
  if (ns == 0 || ns > kDenseMaxVertices || k < 2 || k > 4) return hipErrorInvalidValue;
  const uint32_t W = (ns + 63) / 64;
  hipLaunchKernelGGL(k_bitset_rows, dim3(grid_for((size_t)ns * kWave, 256, 8192)), dim3(256),
                     0, s, ns, off, adj, W, bits);
  hipLaunchKernelGGL(k_alive_init, dim3(1), dim3(256), 0, s, ns, W, alive);
  hipError_t e = hipMemcpyAsync(cur, cd, (size_t)ns * 8, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const unsigned grid = ns < 2048 ? ns : 2048;
  for (uint32_t i = 0; i < ns; ++i) {
    hipLaunchKernelGGL(k_dense_argmin, dim3(1), dim3(kDenseThreads), 0, s, ns, i, cur, alive,
                       step_v, rv, rval);
    hipLaunchKernelGGL(k_dense_dec, dim3(grid), dim3(256), 0, s, ns, W, k, bits, alive,
                       step_v, cur);
  }
  return hipGetLastError();
}

}  // namespace cds
