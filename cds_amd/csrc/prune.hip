// Core pruning before materialising cliques, and the exact peel of a small
// dense core on adjacency bitsets.
//
// A graph with too many k-cliques to materialise (k T slots beyond 32 bits
// or beyond the device) is first cut to its tau-clique-core S: every vertex
// with fewer than tau cliques in S is dropped until none is left to drop,
// and tau rises to the ceiling of the density of S while that grows (each
// S is a boundary state of the peel, so its density is a peel row and tau
// never passes k').  After a large drop the clique degrees come from a
// count pass over the rebuilt induced subgraph; after a small one (k = 3,
// 4) the cliques through the dropped vertices are enumerated from their
// neighbourhoods and taken from the degrees of the rest, in place.  The rows of the
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

// Drops in place (state per vertex of the current subgraph: 1 in S, 2
// dropped this round, 0 dropped before): vertices in S with fewer than tau
// cliques go to state 2 and to the list D, with the sum of their clique
// degrees (stats[0]) and their largest degree (stats[1]).
__global__ void k_prune_select(uint32_t ns, const uint64_t *__restrict__ cd, uint64_t tau,
                               const uint32_t *__restrict__ off, uint8_t *__restrict__ state,
                               uint32_t *__restrict__ D, uint32_t *__restrict__ nd,
                               unsigned long long *__restrict__ stats) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns;
       i += gridDim.x * blockDim.x) {
    if (state[i] != 1 || cd[i] >= tau) continue;
    state[i] = 2;
    D[atomicAdd(nd, 1u)] = i;
    atomicAdd(&stats[0], (unsigned long long)cd[i]);
    atomicMax(&stats[1], (unsigned long long)(off[i + 1] - off[i]));
  }
}

__global__ void k_prune_drop_degrees(uint32_t nd, const uint32_t *__restrict__ D,
                                     const uint32_t *__restrict__ off,
                                     uint32_t *__restrict__ deg) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nd;
       j += gridDim.x * blockDim.x)
    deg[j] = off[D[j] + 1] - off[D[j]];
}

// x may share a clique with the dropped vertex d that is counted from d:
// x is in S, or dropped this round with a larger id (each clique with
// dropped members is removed from its smallest one).
__device__ __forceinline__ bool unlink_eligible(uint8_t sx, uint32_t x, uint32_t d) {
  return sx == 1 || (sx == 2 && x > d);
}

// The K-cliques that contain a dropped vertex, removed from the clique
// degrees of their members in S.  One wave per (dropped vertex d, its
// neighbour a) pair with a eligible: the eligible neighbours b > a of d
// adjacent to a go to an LDS list (one region of `cap` words per wave,
// cap >= every dropped degree); K = 3: each b closes a triangle {d, a, b};
// K = 4: each pair b < c of the list with c in N(b) closes {d, a, b, c}.
// removed[0] += the cliques removed.
template <uint32_t K>
__global__ __launch_bounds__(256) void k_prune_unlink(
    uint32_t nd, const uint32_t *__restrict__ D, const uint32_t *__restrict__ poff,
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
    const uint8_t *__restrict__ state, unsigned long long *__restrict__ cd,
    unsigned long long *__restrict__ removed, uint32_t cap) {
  extern __shared__ uint32_t lds_list[];
  uint32_t *list = lds_list + (threadIdx.x / kWave) * cap;
  const uint32_t total = poff[nd];
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  unsigned long long rem = 0;
  for (uint32_t q = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; q < total; q += waves) {
    // the dropped vertex owning pair q: the last j with poff[j] <= q
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (poff[mid] <= q)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t d = D[lo];
    const uint32_t pa = off[d] + (q - poff[lo]), pend = off[d + 1];
    const uint32_t a = adj[pa];
    const uint8_t sa = state[a];
    if (!unlink_eligible(sa, a, d)) continue;
    const uint32_t a0 = off[a], alen = off[a + 1] - a0;
    uint32_t nl = 0;
    for (uint32_t p0 = pa + 1; p0 < pend; p0 += kWave) {
      const uint32_t p = p0 + lane();
      bool in = false;
      uint32_t b = 0;
      if (p < pend) {
        b = adj[p];
        in = unlink_eligible(state[b], b, d) && find_sorted(adj + a0, (int)alen, b) >= 0;
      }
      unsigned cnt = 0;
      const unsigned slot = wave_slot(in, &cnt);
      if (in) list[nl + slot] = b;
      nl += cnt;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    unsigned long long ca = 0;
    if (K == 3) {
      ca = nl;
      for (uint32_t i = lane(); i < nl; i += kWave) {
        const uint32_t b = list[i];
        if (state[b] == 1) atomicAdd(&cd[b], ~0ull);
      }
    } else {
      for (uint32_t i = 0; i + 1 < nl; ++i) {
        const uint32_t b = list[i];
        const uint32_t b0 = off[b], blen = off[b + 1] - b0;
        uint32_t hits = 0;
        for (uint32_t j = i + 1 + lane(); j < nl; j += kWave) {
          const uint32_t c = list[j];
          if (find_sorted(adj + b0, (int)blen, c) >= 0) {
            ++hits;
            if (state[c] == 1) atomicAdd(&cd[c], ~0ull);
          }
        }
        hits = wave_sum(hits);
        if (lane() == 0 && hits && state[b] == 1) atomicAdd(&cd[b], 0ull - hits);
        ca += hits;
      }
    }
    if (lane() == 0 && ca) {
      if (sa == 1) atomicAdd(&cd[a], 0ull - ca);
      rem += ca;
    }
    // the list is rewritten by the next pair of this wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (lane() == 0 && rem) atomicAdd(removed, rem);
}

__global__ void k_prune_finish(uint32_t nd, const uint32_t *__restrict__ D,
                               const uint32_t *__restrict__ sel, uint8_t *__restrict__ state,
                               uint8_t *__restrict__ keep) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nd;
       j += gridDim.x * blockDim.x) {
    state[D[j]] = 0;
    keep[sel[D[j]]] = 0;
  }
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

// src[p] = the vertex whose adjacency holds arc p, one wave per vertex.
__global__ void k_arc_src(uint32_t ns, const uint32_t *__restrict__ off,
                          uint32_t *__restrict__ src) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; v < ns; v += waves)
    for (uint32_t p = off[v] + lane(); p < off[v + 1]; p += kWave) src[p] = v;
}

// Clique degrees and clique count of a graph of ns <= kBitsCountMax
// vertices from its adjacency bitsets (W <= 64 words a row, G >= W lanes a
// group), every clique once from its two smallest members u < v, one wave
// per edge: C = N(u) & N(v) & {ids > v}, one word per lane of each group.
// K = 3: the triangles {u, v, y}, y in C.  K = 4: the 4-cliques {u, v, y,
// x}, y < x adjacent in C: |N(y) & C| of them through y (one member y per
// group at a time), e(C) = half their sum through u and through v.
// Counts gather in LDS (ns words) and are added to cd once per workgroup.
template <uint32_t K, uint32_t G>
__global__ __launch_bounds__(1024) void k_bits_count(
    uint32_t ns, uint32_t W, const uint32_t *__restrict__ src,
    const uint32_t *__restrict__ adj, uint32_t arcs,
    const unsigned long long *__restrict__ bits, unsigned long long *__restrict__ cd,
    unsigned long long *__restrict__ total) {
  extern __shared__ unsigned long long lcd[];
  for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) lcd[i] = 0;
  __syncthreads();
  constexpr uint32_t kGroups = kWave / G;
  const uint32_t j = lane() % G, gid = lane() / G;
  const uint32_t jj = j < W ? j : W - 1;
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  unsigned long long tot = 0;
  for (uint32_t p = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; p < arcs; p += waves) {
    const uint32_t u = src[p], v = adj[p];
    if (v <= u) continue;
    const unsigned long long x = bits[(size_t)u * W + jj] & bits[(size_t)v * W + jj];
    const uint32_t vw = v >> 6, vb = v & 63;
    const unsigned long long gt =
        j < vw ? 0ull : (j > vw ? ~0ull : (vb == 63 ? 0ull : (~0ull << (vb + 1))));
    const unsigned long long c = j < W ? (x & gt) : 0ull;
    if (K == 3) {
      const uint32_t n1 = gid == 0 ? (uint32_t)__popcll(c) : 0u;
      const uint32_t nc = group_sum<kWave>(n1);
      if (gid == 0)
        for (unsigned long long w = c; w; w &= w - 1)
          atomicAdd(&lcd[j * 64 + __builtin_ctzll(w)], 1ull);
      if (lane() == 0 && nc) {
        atomicAdd(&lcd[u], (unsigned long long)nc);
        atomicAdd(&lcd[v], (unsigned long long)nc);
        tot += nc;
      }
    } else {
      // members of C in ascending order, kGroups at a time
      uint64_t bal = __ballot(gid == 0 && c != 0);
      unsigned long long w = 0;
      uint32_t wl = 0;
      uint32_t e2 = 0;  // this group's sum of |N(y) & C| (in its leader)
      for (;;) {
        uint32_t ys[kGroups];
        uint32_t got = 0;
#pragma unroll
        for (uint32_t g = 0; g < kGroups; ++g) {
          while (!w && bal) {
            wl = (uint32_t)__builtin_ctzll(bal);
            bal &= bal - 1;
            w = readlane64(c, wl);
          }
          ys[g] = 0;
          if (w) {
            ys[g] = wl * 64 + (uint32_t)__builtin_ctzll(w);
            w &= w - 1;
            ++got;
          }
        }
        if (!got) break;
        uint32_t y = ys[0];
#pragma unroll
        for (uint32_t g = 1; g < kGroups; ++g)
          if (gid == g) y = ys[g];
        const bool valid = gid < got;
        const unsigned long long r = bits[(size_t)y * W + jj] & c;
        const uint32_t py = group_sum<G>(valid ? (uint32_t)__popcll(r) : 0u);
        if (j == 0 && valid && py) {
          atomicAdd(&lcd[y], (unsigned long long)py);
          e2 += py;
        }
      }
      const uint32_t es = group_sum<kWave>(e2);
      if (lane() == 0 && es) {
        const unsigned long long e = es / 2;
        atomicAdd(&lcd[u], e);
        atomicAdd(&lcd[v], e);
        tot += e;
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x)
    if (lcd[i]) atomicAdd(&cd[i], lcd[i]);
  if (lane() == 0 && tot) atomicAdd(total, tot);
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

hipError_t bits_clique_count(uint32_t k, uint32_t ns, const uint32_t *off,
                             const uint32_t *adj, uint32_t arcs, unsigned long long *bits,
                             uint32_t *src, unsigned long long *cd,
                             unsigned long long *total, hipStream_t s) {
  if ((k != 3 && k != 4) || ns == 0 || ns > kBitsCountMax) return hipErrorInvalidValue;
  const uint32_t W = (ns + 63) / 64;
  hipLaunchKernelGGL(k_bitset_rows, dim3(grid_for((size_t)ns * kWave, 256, 8192)), dim3(256),
                     0, s, ns, off, adj, W, bits);
  hipLaunchKernelGGL(k_arc_src, dim3(grid_for((size_t)ns * kWave, 256, 8192)), dim3(256), 0, s,
                     ns, off, src);
  hipError_t e = hipMemsetAsync(cd, 0, (size_t)ns * sizeof(unsigned long long), s);
  if (e == hipSuccess) e = hipMemsetAsync(total, 0, sizeof(unsigned long long), s);
  if (e != hipSuccess) return e;
  if (arcs == 0) return hipGetLastError();
  const size_t lds = (size_t)ns * sizeof(unsigned long long);
  const unsigned grid = grid_for((size_t)arcs * kWave, 1024, 512);
#define CDS_BITS_COUNT(KK, GG)                                                              \
  hipLaunchKernelGGL((k_bits_count<KK, GG>), dim3(grid), dim3(1024), lds, s, ns, W, src, adj, \
                     arcs, bits, cd, total)
  if (k == 3) {
    CDS_BITS_COUNT(3, 64);
  } else if (W <= 8) {
    CDS_BITS_COUNT(4, 8);
  } else if (W <= 16) {
    CDS_BITS_COUNT(4, 16);
  } else if (W <= 32) {
    CDS_BITS_COUNT(4, 32);
  } else {
    CDS_BITS_COUNT(4, 64);
  }
#undef CDS_BITS_COUNT
  return hipGetLastError();
}

hipError_t prune_select(uint32_t ns, const uint64_t *cd, uint64_t tau, const uint32_t *off,
                        uint8_t *state, uint32_t *D, uint32_t *nd,
                        unsigned long long *stats, hipStream_t s) {
  hipLaunchKernelGGL(k_prune_select, dim3(grid_for(ns, 256, 4096)), dim3(256), 0, s, ns, cd,
                     tau, off, state, D, nd, stats);
  return hipGetLastError();
}

hipError_t prune_unlink(uint32_t k, uint32_t nd, const uint32_t *D, uint32_t max_deg,
                        const uint32_t *off, const uint32_t *adj, const uint8_t *state,
                        unsigned long long *cd, unsigned long long *removed, uint32_t *deg,
                        uint32_t *poff, DeviceArena &a) {
  if ((k != 3 && k != 4) || max_deg > kUnlinkMaxDegree) return hipErrorInvalidValue;
  if (nd == 0) return hipSuccess;
  hipStream_t s = a.stream();
  hipLaunchKernelGGL(k_prune_drop_degrees, dim3(grid_for(nd, 256, 1024)), dim3(256), 0, s, nd,
                     D, off, deg);
  hipError_t e = scan_u32_32(deg, poff, nd, a);
  if (e != hipSuccess) return e;
  const uint32_t cap = max_deg < 1 ? 1 : max_deg;
  const size_t lds = (size_t)4 * cap * sizeof(uint32_t);
  const unsigned grid = grid_for((size_t)nd * cap * kWave, 256, 8192);
  if (k == 3)
    hipLaunchKernelGGL(k_prune_unlink<3>, dim3(grid), dim3(256), lds, s, nd, D, poff, off,
                       adj, state, cd, removed, cap);
  else
    hipLaunchKernelGGL(k_prune_unlink<4>, dim3(grid), dim3(256), lds, s, nd, D, poff, off,
                       adj, state, cd, removed, cap);
  return hipGetLastError();
}

hipError_t prune_finish(uint32_t nd, const uint32_t *D, const uint32_t *sel, uint8_t *state,
                        uint8_t *keep, hipStream_t s) {
  if (nd == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prune_finish, dim3(grid_for(nd, 256, 1024)), dim3(256), 0, s, nd, D,
                     sel, state, keep);
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
