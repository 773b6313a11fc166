// Densest-core extraction, edge pruning and connected components.
//
// Reference GPU path: generateDensestCore / generateNeighborDensestCore
// (cudaCode/src/helpers.cu:821-896), pruneEdges + pruned CSR (:898-1041),
// componentDecomposek label propagation (:937-978) and the per-component
// clique bucketing getConnectedComponentStatus / rearrangeCliqueData
// (:1043-1096).  Oracle: locateDensestCore, pruneInvalidEdges,
// connectedComponentDecompose (cpuCode/c++/src/helpers.cpp:710-788,
// :958-1116).
//
// The core's cliques are the listed cliques whose kill level (the minimum
// core number of their members) is >= k', so no re-listing is needed.
// Components of the pruned graph are the classes of "shares a core clique";
// they come from a lock-free union-find that always hooks the larger root
// under the smaller, so every root is its component's smallest vertex id,
// which is the oracle's component order.
#include "dev.h"
#include "kernels.h"

namespace cds {
namespace {

// Core vertices: every vertex (n of them) with core >= kp, or the n listed
// in cv when cv is given (all of them core vertices).
__device__ __forceinline__ bool core_vertex(const uint32_t *cv, const uint32_t *core,
                                            uint32_t kp, uint32_t i, uint32_t *v) {
  if (cv) {
    *v = cv[i];
    return true;
  }
  *v = i;
  return core[i] >= kp;
}

// A core clique by index: the i-th of the list cc when given (T entries),
// else clique i when its kill level is >= kp, or clique i when kill is
// null too (every clique is in the core).
__device__ __forceinline__ bool core_clique(const uint32_t *cc, const uint32_t *kill,
                                            uint32_t kp, uint64_t i, uint64_t *c) {
  if (cc) {
    *c = cc[i];
    return true;
  }
  *c = i;
  return !kill || kill[i] >= kp;
}

__global__ __launch_bounds__(256) void k_core_edges(
    uint32_t n, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ adj, const uint32_t *__restrict__ core,
    uint32_t kp, unsigned long long *out, const uint32_t *__restrict__ cv) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  unsigned long long c = 0;
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       i < n; i += waves) {
    uint32_t v;
    if (!core_vertex(cv, core, kp, i, &v)) continue;
    for (uint32_t j = off[v] + lane(); j < off[v + 1]; j += kWave)
      c += core[adj[j]] >= kp;
  }
  c = wave_sum64(c);
  if (lane() == 0 && c) atomicAdd(out, c);
}

// Counts the cliques with kill level >= kp (with kill null: whose members
// all have core >= kp, the same cliques) and, when cc is given, lists their
// ids there (the order is not fixed).  A wave takes 64 * kCoreSpan
// consecutive cliques per pass and reserves its slots with one atomic.
constexpr uint32_t kCoreSpan = 16;

__device__ __forceinline__ bool in_core(uint64_t i, uint32_t k, const uint32_t *cl,
                                        const uint32_t *core, const uint32_t *kill,
                                        uint32_t kp) {
  if (kill) return kill[i] >= kp;
  bool all = true;
  for (uint32_t j = 0; j < k; ++j) all &= core[cl[i * k + j]] >= kp;
  return all;
}

__global__ void k_core_cliques(uint64_t T, uint32_t k, const uint32_t *__restrict__ cl,
                               const uint32_t *__restrict__ core,
                               const uint32_t *__restrict__ kill,
                               uint32_t kp, uint32_t *__restrict__ cc,
                               unsigned long long *out) {
  const uint64_t span = (uint64_t)kWave * kCoreSpan;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
  for (uint64_t b = ((uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * span;
       b < T; b += waves * span) {
    uint64_t m[kCoreSpan];
    uint32_t tot = 0;
#pragma unroll
    for (uint32_t r = 0; r < kCoreSpan; ++r) {
      const uint64_t i = b + (uint64_t)r * kWave + lane();
      m[r] = __ballot(i < T && in_core(i, k, cl, core, kill, kp));
      tot += (uint32_t)__popcll(m[r]);
    }
    if (tot == 0) continue;  // wave-uniform
    unsigned long long base = 0;
    if (lane() == 0) base = atomicAdd(out, (unsigned long long)tot);
    base = __shfl(base, 0, kWave);
    if (!cc) continue;
#pragma unroll
    for (uint32_t r = 0; r < kCoreSpan; ++r) {
      if ((m[r] >> lane()) & 1ull)
        cc[base + __popcll(m[r] & lanes_below())] = (uint32_t)(b + (uint64_t)r * kWave + lane());
      base += __popcll(m[r]);
    }
  }
}

// The cliques whose first listed member (the lowest in the degree order)
// is core vertex cv[i]: one contiguous range of the clique array, from the
// DAG offsets (k = 2) or the scanned listing items (k >= 3).  rchunks[i]:
// the range in spans of 64 * kCoreSpan cliques.
__global__ void k_core_ranges(uint32_t cn, const uint32_t *__restrict__ cv, uint32_t k,
                              const uint32_t *__restrict__ doff,
                              const uint64_t *__restrict__ item_off,
                              const uint64_t *__restrict__ item_base,
                              uint64_t *__restrict__ rstart, uint64_t *__restrict__ rend,
                              uint32_t *__restrict__ rchunks) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cn;
       i += gridDim.x * blockDim.x) {
    const uint32_t u = cv[i];
    uint64_t b, e;
    if (k == 2) {
      b = doff[u];
      e = doff[u + 1];
    } else {
      b = item_base[item_off[u]];
      e = item_base[item_off[u + 1]];
    }
    rstart[i] = b;
    rend[i] = e;
    rchunks[i] = (uint32_t)((e - b + kWave * kCoreSpan - 1) / (kWave * kCoreSpan));
  }
}

// Lists the core's cliques from the core vertices' ranges only: a core
// clique's first member is a core vertex.  One span per wave and pass.
__global__ void k_core_cliques_ranged(uint32_t cn, const uint64_t *__restrict__ choff,
                                      const uint64_t *__restrict__ rstart,
                                      const uint64_t *__restrict__ rend, uint32_t k,
                                      const uint32_t *__restrict__ cl,
                                      const uint32_t *__restrict__ core, uint32_t kp,
                                      uint32_t *__restrict__ cc, unsigned long long *out) {
  const uint64_t total = choff[cn];
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
  for (uint64_t ch = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       ch < total; ch += waves) {
    const uint32_t i = wave_last_le(choff, cn, ch);
    const uint64_t b = rstart[i] + (ch - choff[i]) * (kWave * kCoreSpan);
    const uint64_t e = min(rend[i], b + (uint64_t)kWave * kCoreSpan);
    uint64_t m[kCoreSpan];
    uint32_t tot = 0;
#pragma unroll
    for (uint32_t r = 0; r < kCoreSpan; ++r) {
      const uint64_t c = b + (uint64_t)r * kWave + lane();
      m[r] = __ballot(c < e && in_core(c, k, cl, core, nullptr, kp));
      tot += (uint32_t)__popcll(m[r]);
    }
    if (tot == 0) continue;  // wave-uniform
    unsigned long long base = 0;
    if (lane() == 0) base = atomicAdd(out, (unsigned long long)tot);
    base = __shfl(base, 0, kWave);
#pragma unroll
    for (uint32_t r = 0; r < kCoreSpan; ++r) {
      if ((m[r] >> lane()) & 1ull)
        cc[base + __popcll(m[r] & lanes_below())] = (uint32_t)(b + (uint64_t)r * kWave + lane());
      base += __popcll(m[r]);
    }
  }
}

__global__ void k_uf_init(uint32_t n, uint32_t *__restrict__ p) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x)
    p[v] = v;
}

__device__ __forceinline__ uint32_t uf_load(uint32_t *p, uint32_t x) {
  return __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving (benign races: every write stores an
// ancestor of the slot's current value).
__device__ uint32_t uf_find(uint32_t *p, uint32_t x) {
  for (;;) {
    const uint32_t y = uf_load(p, x);
    if (y == x) return x;
    const uint32_t z = uf_load(p, y);
    if (z != y)
      __hip_atomic_store(&p[x], z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = y;
  }
}

__device__ void uf_unite(uint32_t *p, uint32_t a, uint32_t b) {
  for (;;) {
    a = uf_find(p, a);
    b = uf_find(p, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    // hook the larger root a under the smaller root b
    if (atomicCAS(&p[a], a, b) == a) return;
  }
}

// Unions the endpoints of every edge marked as covered by a core clique
// (the pruned graph): one wave per core vertex, each edge once (a < b).
__global__ __launch_bounds__(256) void k_uf_union_edges(
    uint32_t n, const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
    const uint8_t *__restrict__ mark, const uint32_t *__restrict__ core,
    uint32_t kp, uint32_t *__restrict__ p, const uint32_t *__restrict__ cv) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       i < n; i += waves) {
    uint32_t a;
    if (!core_vertex(cv, core, kp, i, &a)) continue;
    for (uint32_t j = off[a] + lane(); j < off[a + 1]; j += kWave) {
      const uint32_t b = adj[j];
      if (b > a && mark[j]) uf_unite(p, a, b);
    }
  }
}

__global__ void k_uf_flatten(uint32_t n, uint32_t *__restrict__ p) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x)
    p[v] = uf_find(p, v);
}

__global__ void k_prune_mark(uint64_t T, uint32_t k,
                             const uint32_t *__restrict__ cl,
                             const uint32_t *__restrict__ kill, uint32_t kp,
                             const uint32_t *__restrict__ off,
                             const uint32_t *__restrict__ adj,
                             uint8_t *__restrict__ mark,
                             const uint32_t *__restrict__ cc) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < T;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t c;
    if (!core_clique(cc, kill, kp, i, &c)) continue;
    for (uint32_t i = 0; i < k; ++i) {
      const uint32_t a = cl[c * k + i];
      const uint32_t ab = off[a], ad = off[a + 1] - ab;
      for (uint32_t j = 0; j < k; ++j) {
        if (i == j) continue;
        const int64_t pos = find_sorted64(adj + ab, ad, cl[c * k + j]);
        if (pos >= 0) mark[ab + pos] = 1;
      }
    }
  }
}

// Dense prune for a small densest core: cidx numbers the core vertices,
// cmat[i * cn + j] holds the CSR position of the edge between core
// vertices i and j (or kNoEdge), so every ordered member pair of a core
// clique is one lookup in a matrix that stays in cache instead of a binary
// search through a hub's adjacency list.
constexpr uint32_t kNoEdge = 0xFFFFFFFFu;

// Numbers the core vertices: cidx[v] = slot (kNoEdge off the core) and
// cv[slot] = v, for slots below cap (the host checks the count).
__global__ void k_core_list(uint32_t n, const uint32_t *__restrict__ core,
                            uint32_t kp, uint32_t *__restrict__ cidx,
                            uint32_t *__restrict__ cv, uint32_t cap,
                            uint32_t *__restrict__ count) {
  for (uint32_t base = blockIdx.x * blockDim.x; base < n;
       base += gridDim.x * blockDim.x) {
    const uint32_t v = base + threadIdx.x;
    const bool in = v < n && core[v] >= kp;
    const uint32_t slot = aggregated_reserve(count, 0, in);
    if (v < n) cidx[v] = in ? slot : kNoEdge;
    if (in && slot < cap) cv[slot] = v;
  }
}

__global__ void k_fill_edges(uint64_t n, uint32_t *__restrict__ a) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = kNoEdge;
}

// Row iu of the matrix from the adjacency of core vertex cv[iu].
__global__ void k_core_edge_matrix(const uint32_t *__restrict__ cv,
                                   const uint32_t *__restrict__ off,
                                   const uint32_t *__restrict__ adj,
                                   const uint32_t *__restrict__ cidx, uint32_t cn,
                                   uint32_t *__restrict__ cmat) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t iu = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; iu < cn;
       iu += waves) {
    const uint32_t u = cv[iu];
    for (uint32_t p = off[u] + lane(); p < off[u + 1]; p += kWave) {
      const uint32_t iv = cidx[adj[p]];
      if (iv < cn) cmat[(uint64_t)iu * cn + iv] = p;
    }
  }
}

__global__ void k_prune_mark_dense(uint64_t T, uint32_t k,
                                   const uint32_t *__restrict__ cl,
                                   const uint32_t *__restrict__ cc,
                                   const uint32_t *__restrict__ cidx, uint32_t cn,
                                   const uint32_t *__restrict__ cmat,
                                   uint8_t *__restrict__ mark) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < T;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t c = cc ? cc[i] : i;
    for (uint32_t i = 0; i < k; ++i) {
      const uint32_t ci = cidx[cl[c * k + i]];
      if (ci >= cn) continue;  // cannot happen for a core clique
      for (uint32_t j = 0; j < k; ++j) {
        const uint32_t cj = cidx[cl[c * k + j]];
        if (i == j || cj >= cn) continue;
        const uint32_t pos = cmat[(uint64_t)ci * cn + cj];
        // a core edge is in many core cliques: store its mark once (a
        // cached read instead of a partial-line write per clique pair;
        // threads racing on the same mark all store 1)
        if (pos != kNoEdge && !mark[pos]) mark[pos] = 1;
      }
    }
  }
}

// A small core: the core's adjacency as bitset rows (row i: the core
// neighbours of cv[i], by their cidx numbers, W words), then a directed
// core edge (a, b) lies in a core clique iff k = 2; (k = 3) a and b have a
// common core neighbour: rows a and b intersect; (k = 4) their common core
// neighbourhood X holds an edge: some c in X has a neighbour in X.
__global__ __launch_bounds__(256) void k_core_rows(const uint32_t *__restrict__ cv, uint32_t cn,
                                                   uint32_t W, const uint32_t *__restrict__ off,
                                                   const uint32_t *__restrict__ adj,
                                                   const uint32_t *__restrict__ cidx,
                                                   unsigned long long *__restrict__ bits) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t ia = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; ia < cn;
       ia += waves) {
    const uint32_t a = cv[ia];
    for (uint32_t p = off[a] + lane(); p < off[a + 1]; p += kWave) {
      const uint32_t ib = cidx[adj[p]];
      if (ib < cn) atomicOr(&bits[(size_t)ia * W + (ib >> 6)], 1ull << (ib & 63));
    }
  }
}

__global__ __launch_bounds__(256) void k_prune_mark_bits(
    uint32_t k, const uint32_t *__restrict__ cv, uint32_t cn, uint32_t W,
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
    const uint32_t *__restrict__ cidx, const unsigned long long *__restrict__ bits,
    uint8_t *__restrict__ mark) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t ia = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; ia < cn;
       ia += waves) {
    const uint32_t a = cv[ia];
    const unsigned long long *ra = bits + (size_t)ia * W;
    for (uint32_t p = off[a] + lane(); p < off[a + 1]; p += kWave) {
      const uint32_t ib = cidx[adj[p]];
      if (ib >= cn) continue;
      bool in = k == 2;
      const unsigned long long *rb = bits + (size_t)ib * W;
      if (k == 3) {
        for (uint32_t w = 0; w < W && !in; ++w) in = (ra[w] & rb[w]) != 0;
      } else if (k == 4) {
        for (uint32_t w = 0; w < W && !in; ++w) {
          unsigned long long x = ra[w] & rb[w];
          while (x && !in) {
            const uint32_t c = w * 64 + (uint32_t)(__ffsll((long long)x) - 1);
            x &= x - 1;
            const unsigned long long *rc = bits + (size_t)c * W;
            for (uint32_t w2 = 0; w2 < W && !in; ++w2) in = (rc[w2] & ra[w2] & rb[w2]) != 0;
          }
        }
      }
      if (in) mark[p] = 1;
    }
  }
}

__global__ void k_count_nonzero_u8(size_t n, const uint8_t *__restrict__ a,
                                   unsigned long long *out) {
  unsigned long long c = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    c += a[i] != 0;
  c = wave_sum64(c);
  if (lane() == 0 && c) atomicAdd(out, c);
}

// Cliques of each component (wave-aggregated counters).
// With at most kCompLds components the per-component counts of a
// workgroup are summed in LDS, so a component holding most core cliques
// sees one global atomic per workgroup (count) or per tile (slots).
constexpr uint32_t kCompLds = 1024;
constexpr int kCompThreads = 256;
constexpr int kCompTile = 8;  // cliques per thread and tile

__global__ __launch_bounds__(kCompThreads) void k_comp_count2(
    uint64_t T, const uint32_t *__restrict__ cl, uint32_t k,
    const uint32_t *__restrict__ kill, uint32_t kp,
    const uint32_t *__restrict__ comp_of, uint32_t ncomp,
    uint32_t *__restrict__ ccount, const uint32_t *__restrict__ cc) {
  __shared__ uint32_t cnt[kCompLds];
  const bool lds = ncomp <= kCompLds;
  if (lds) {
    for (uint32_t i = threadIdx.x; i < ncomp; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
  }
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < T;
       base += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t c = 0;
    const bool in = base + threadIdx.x < T && core_clique(cc, kill, kp, base + threadIdx.x, &c);
    const uint32_t comp = in ? comp_of[cl[c * k]] : 0;
    (void)aggregated_reserve(lds ? cnt : ccount, comp, in);
  }
  if (lds) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ncomp; i += blockDim.x)
      if (cnt[i]) atomicAdd(&ccount[i], cnt[i]);
  }
}

// Places every core clique in its component's range of the flow arrays:
// fmem (flow-vertex ids of the members) and cslot[c] = its flow index.
// Three member ids moved with one 12-byte load or store.
struct __attribute__((packed, aligned(4))) Words3 {
  uint32_t w[3];
};

// Clique c's members in flow ids at record f (stride ms words): one vector
// load and one vector store per clique for k = 3 and 4 (clique lists and
// records start 256-byte aligned; a k = 4 record stride is 4 or 8 words).
__device__ __forceinline__ void comp_place(uint64_t c, uint64_t f, uint32_t k,
                                           uint32_t ms,
                                           const uint32_t *__restrict__ cl,
                                           const uint32_t *__restrict__ fv_of,
                                           uint32_t *__restrict__ fmem,
                                           uint32_t *__restrict__ cslot) {
  if (cslot) cslot[c] = (uint32_t)f;
  if (k == 4 && (ms & 3) == 0) {
    const uint4 m = *reinterpret_cast<const uint4 *>(cl + c * 4);
    *reinterpret_cast<uint4 *>(fmem + f * ms) =
        make_uint4(fv_of[m.x], fv_of[m.y], fv_of[m.z], fv_of[m.w]);
  } else if (k == 3) {
    const Words3 m = *reinterpret_cast<const Words3 *>(cl + c * 3);
    *reinterpret_cast<Words3 *>(fmem + f * ms) =
        Words3{{fv_of[m.w[0]], fv_of[m.w[1]], fv_of[m.w[2]]}};
  } else {
    for (uint32_t j = 0; j < k; ++j) fmem[f * ms + j] = fv_of[cl[c * k + j]];
  }
}

__global__ __launch_bounds__(kCompThreads) void k_comp_scatter2(
    uint64_t T, uint32_t k, const uint32_t *__restrict__ cl,
    const uint32_t *__restrict__ kill, uint32_t kp,
    const uint32_t *__restrict__ comp_of, const uint32_t *__restrict__ fv_of,
    uint32_t ncomp, const uint64_t *__restrict__ coff,
    uint32_t *__restrict__ ccur, uint32_t *__restrict__ fmem,
    uint32_t *__restrict__ cslot, uint32_t ms, const uint32_t *__restrict__ cc) {
  __shared__ uint32_t cnt[kCompLds], cbase[kCompLds];
  if (ncomp > kCompLds) {
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < T;
         base += (uint64_t)gridDim.x * blockDim.x) {
      uint64_t c = 0;
      const bool in = base + threadIdx.x < T && core_clique(cc, kill, kp, base + threadIdx.x, &c);
      const uint32_t comp = in ? comp_of[cl[c * k]] : 0;
      const uint32_t slot = aggregated_reserve(ccur, comp, in);
      if (in) comp_place(c, coff[comp] + slot, k, ms, cl, fv_of, fmem, cslot);
    }
    return;
  }
  for (uint32_t i = threadIdx.x; i < ncomp; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  // tiles of blockDim * kCompTile cliques: slots within the tile from LDS
  // counters, then one global reservation per component present
  const uint64_t tile = (uint64_t)blockDim.x * kCompTile;
  for (uint64_t t0 = (uint64_t)blockIdx.x * tile; t0 < T;
       t0 += (uint64_t)gridDim.x * tile) {
    uint32_t comp[kCompTile], local[kCompTile];
    uint64_t cid[kCompTile];
    bool in[kCompTile];
#pragma unroll
    for (int it = 0; it < kCompTile; ++it) {
      const uint64_t i = t0 + (uint64_t)it * blockDim.x + threadIdx.x;
      cid[it] = 0;
      in[it] = i < T && core_clique(cc, kill, kp, i, &cid[it]);
      comp[it] = in[it] ? comp_of[cl[cid[it] * k]] : 0;
      local[it] = aggregated_reserve(cnt, comp[it], in[it]);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ncomp; i += blockDim.x) {
      const uint32_t n = cnt[i];
      if (n) {
        cbase[i] = atomicAdd(&ccur[i], n);
        cnt[i] = 0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kCompTile; ++it) {
      if (!in[it]) continue;
      comp_place(cid[it], coff[comp[it]] + cbase[comp[it]] + local[it], k, ms, cl, fv_of,
                 fmem, cslot);
    }
    __syncthreads();  // cbase is rewritten by the next tile
  }
}

// Flow-network incidence of the core, from its clique bucket fmem (T'
// cliques of k flow-vertex ids each): every (clique slot s, position j) pair
// belongs to vertex fmem[s*k+j], whose list holds s*k+j.  A workgroup takes
// a tile of kFlowTile cliques and aggregates its pairs in an LDS hash keyed
// by vertex (a core is a few hub-like vertices in most cliques), so each
// distinct vertex of a tile costs one global atomic: in the count pass its
// tile total, in the fill pass the reservation of a run of its list.  The
// order of a vertex's list depends on the order the reservations land in;
// the max flow and its minimal min-cut side do not.
constexpr uint32_t kFlowThreads = 1024;
constexpr uint32_t kFlowTile = 2048;                 // cliques per tile
constexpr uint32_t kFlowPairs = kFlowTile * 4;       // k <= 4
constexpr uint32_t kFlowHash = 2 * kFlowPairs;       // load <= 1/2
constexpr uint32_t kFlowPerThread = kFlowPairs / kFlowThreads;

// Inserts the tile's pairs into the hash; slot[] / rank[] per pair (rank:
// position among the tile's pairs of the same vertex).
__device__ __forceinline__ void flow_tile_hash(const uint32_t *__restrict__ fmem,
                                               uint32_t k, uint32_t ms,
                                               uint64_t p0, uint32_t pairs,
                                               uint32_t *hkey, uint32_t *hcnt,
                                               uint32_t *slot, uint32_t *rank) {
#pragma unroll
  for (uint32_t r = 0; r < kFlowPerThread; ++r) {
    const uint32_t p = r * kFlowThreads + threadIdx.x;
    if (p >= pairs) break;
    const uint32_t x = member_at(fmem, p0 + p, k, ms);
    uint32_t h = (x * 2654435761u) & (kFlowHash - 1);
    for (;;) {  // at most pairs < kFlowHash keys: a free slot exists
      const uint32_t prev = atomicCAS(&hkey[h], kAlive, x);
      if (prev == kAlive || prev == x) break;
      h = (h + 1) & (kFlowHash - 1);
    }
    slot[r] = h;
    rank[r] = atomicAdd(&hcnt[h], 1u);
  }
}

__global__ __launch_bounds__(kFlowThreads) void k_flow_deg(
    uint64_t Tc, uint32_t k, const uint32_t *__restrict__ fmem,
    uint32_t *__restrict__ fcd, uint32_t ms) {
  __shared__ uint32_t hkey[kFlowHash];
  __shared__ uint32_t hcnt[kFlowHash];
  const uint64_t tiles = (Tc + kFlowTile - 1) / kFlowTile;
  for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    for (uint32_t i = threadIdx.x; i < kFlowHash; i += kFlowThreads) {
      hkey[i] = kAlive;
      hcnt[i] = 0;
    }
    __syncthreads();
    const uint64_t c0 = tile * kFlowTile;
    const uint32_t pairs = (uint32_t)min<uint64_t>(kFlowTile, Tc - c0) * k;
    uint32_t slot[kFlowPerThread], rank[kFlowPerThread];
    flow_tile_hash(fmem, k, ms, c0 * k, pairs, hkey, hcnt, slot, rank);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kFlowHash; i += kFlowThreads)
      if (hcnt[i]) atomicAdd(&fcd[hkey[i]], hcnt[i]);
    __syncthreads();
  }
}

__global__ __launch_bounds__(kFlowThreads) void k_flow_fill(
    uint64_t Tc, uint32_t k, const uint32_t *__restrict__ fmem,
    const uint64_t *__restrict__ fioff, uint32_t *__restrict__ fcur,
    uint32_t *__restrict__ finc, uint32_t div, uint32_t ms) {
  __shared__ uint32_t hkey[kFlowHash];
  __shared__ uint32_t hcnt[kFlowHash];
  const uint64_t tiles = (Tc + kFlowTile - 1) / kFlowTile;
  for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    for (uint32_t i = threadIdx.x; i < kFlowHash; i += kFlowThreads) {
      hkey[i] = kAlive;
      hcnt[i] = 0;
    }
    __syncthreads();
    const uint64_t c0 = tile * kFlowTile;
    const uint32_t pairs = (uint32_t)min<uint64_t>(kFlowTile, Tc - c0) * k;
    uint32_t slot[kFlowPerThread], rank[kFlowPerThread];
    flow_tile_hash(fmem, k, ms, c0 * k, pairs, hkey, hcnt, slot, rank);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kFlowHash; i += kFlowThreads)
      if (hcnt[i]) hcnt[i] = atomicAdd(&fcur[hkey[i]], hcnt[i]);  // -> run base
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kFlowPerThread; ++r) {
      const uint32_t p = r * kFlowThreads + threadIdx.x;
      if (p >= pairs) break;
      const uint32_t x = hkey[slot[r]];
      finc[fioff[x] + hcnt[slot[r]] + rank[r]] = (uint32_t)((c0 * k + p) / div);
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t core_list(uint32_t n, const uint32_t *core, uint32_t kp, uint32_t *cidx,
                     uint32_t *cv, uint32_t cap, uint32_t *count, hipStream_t s) {
  hipError_t e = hipMemsetAsync(count, 0, sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_core_list, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, core,
                     kp, cidx, cv, cap, count);
  return hipGetLastError();
}

hipError_t core_edge_count(uint32_t n, const uint32_t *off,
                           const uint32_t *adj, const uint32_t *core,
                           uint32_t kp, unsigned long long *out,
                           hipStream_t s, const uint32_t *cv) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(*out), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_core_edges, dim3(grid_for((size_t)n * kWave, 256, 16384)),
                     dim3(256), 0, s, n, off, adj, core, kp, out, cv);
  return hipGetLastError();
}

hipError_t core_cliques(uint64_t T, uint32_t k, const uint32_t *cliques, const uint32_t *core,
                        const uint32_t *kill, uint32_t kp, uint32_t *cc,
                        unsigned long long *out, hipStream_t s) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(*out), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_core_cliques, dim3(grid_for((T + kCoreSpan - 1) / kCoreSpan, 256, 8192)),
                     dim3(256), 0, s, T, k, cliques, core, kill, kp, cc, out);
  return hipGetLastError();
}

hipError_t uf_init(uint32_t n, uint32_t *parent, hipStream_t s) {
  hipLaunchKernelGGL(k_uf_init, dim3(grid_for(n, 256, 65535)), dim3(256), 0, s,
                     n, parent);
  return hipGetLastError();
}

hipError_t core_cliques_ranged(uint32_t cn, const uint32_t *cv, uint32_t k,
                               const uint32_t *doff, const uint64_t *item_off,
                               const uint64_t *item_base, const uint32_t *cliques,
                               const uint32_t *core, uint32_t kp, uint32_t *cc,
                               unsigned long long *out, DeviceArena &a) {
  hipStream_t s = a.stream();
  hipError_t e = hipMemsetAsync(out, 0, sizeof(*out), s);
  if (e != hipSuccess || cn == 0) return e;
  void *rs = nullptr, *re = nullptr, *rc = nullptr, *co = nullptr;
  auto done = [&](hipError_t r) {
    for (void *p : {rs, re, rc, co})
      if (p) a.release(p);
    return r;
  };
  if ((e = a.alloc(&rs, (size_t)cn * 8)) != hipSuccess) return done(e);
  if ((e = a.alloc(&re, (size_t)cn * 8)) != hipSuccess) return done(e);
  if ((e = a.alloc(&rc, (size_t)cn * 4)) != hipSuccess) return done(e);
  if ((e = a.alloc(&co, ((size_t)cn + 1) * 8)) != hipSuccess) return done(e);
  auto *rstart = static_cast<uint64_t *>(rs), *rend = static_cast<uint64_t *>(re);
  auto *rchunks = static_cast<uint32_t *>(rc);
  auto *choff = static_cast<uint64_t *>(co);
  hipLaunchKernelGGL(k_core_ranges, dim3(grid_for(cn, 256, 1024)), dim3(256), 0, s, cn, cv, k,
                     doff, item_off, item_base, rstart, rend, rchunks);
  if ((e = hipGetLastError()) != hipSuccess) return done(e);
  if ((e = scan_u32(rchunks, choff, cn, a)) != hipSuccess) return done(e);
  hipLaunchKernelGGL(k_core_cliques_ranged, dim3(4096), dim3(256), 0, s, cn, choff, rstart,
                     rend, k, cliques, core, kp, cc, out);
  return done(hipGetLastError());
}

hipError_t uf_union_edges(uint32_t n, const uint32_t *off, const uint32_t *adj,
                          const uint8_t *mark, const uint32_t *core,
                          uint32_t kp, uint32_t *parent, hipStream_t s,
                          const uint32_t *cv) {
  hipLaunchKernelGGL(k_uf_union_edges,
                     dim3(grid_for((size_t)n * kWave, 256, 16384)), dim3(256),
                     0, s, n, off, adj, mark, core, kp, parent, cv);
  return hipGetLastError();
}

hipError_t uf_flatten(uint32_t n, uint32_t *parent, hipStream_t s) {
  hipLaunchKernelGGL(k_uf_flatten, dim3(grid_for(n, 256, 65535)), dim3(256), 0,
                     s, n, parent);
  return hipGetLastError();
}

hipError_t prune_mark_dense(uint64_t Tc, uint32_t k, const uint32_t *cliques,
                            const uint32_t *cc, const uint32_t *cv, uint32_t cn,
                            const uint32_t *cidx, const uint32_t *off, const uint32_t *adj,
                            uint32_t *cmat, uint8_t *mark, hipStream_t s) {
  const uint64_t cells = (uint64_t)cn * cn;
  hipLaunchKernelGGL(k_fill_edges, dim3(grid_for(cells, 256, 16384)), dim3(256), 0, s,
                     cells, cmat);
  hipLaunchKernelGGL(k_core_edge_matrix, dim3(grid_for((size_t)cn * kWave, 256, 8192)),
                     dim3(256), 0, s, cv, off, adj, cidx, cn, cmat);
  if (Tc)
    hipLaunchKernelGGL(k_prune_mark_dense, dim3(grid_for(Tc, 256, 65535)), dim3(256), 0, s,
                       Tc, k, cliques, cc, cidx, cn, cmat, mark);
  return hipGetLastError();
}

hipError_t prune_mark_bits(uint32_t k, const uint32_t *cv, uint32_t cn, const uint32_t *cidx,
                           const uint32_t *off, const uint32_t *adj,
                           unsigned long long *bits, uint8_t *mark, hipStream_t s) {
  if (k < 2 || k > 4) return hipErrorInvalidValue;
  const uint32_t W = (cn + 63) / 64;
  hipError_t e = hipMemsetAsync(bits, 0, (size_t)cn * W * 8, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_core_rows, dim3(grid_for((size_t)cn * kWave, 256, 8192)), dim3(256), 0,
                     s, cv, cn, W, off, adj, cidx, bits);
  hipLaunchKernelGGL(k_prune_mark_bits, dim3(grid_for((size_t)cn * kWave, 256, 8192)),
                     dim3(256), 0, s, k, cv, cn, W, off, adj, cidx, bits, mark);
  return hipGetLastError();
}

hipError_t prune_mark(uint64_t T, uint32_t k, const uint32_t *cliques,
                      const uint32_t *kill, uint32_t kp, const uint32_t *off,
                      const uint32_t *adj, uint8_t *mark, hipStream_t s,
                      const uint32_t *cc) {
  if (T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prune_mark, dim3(grid_for(T, 256, 65535)), dim3(256), 0,
                     s, T, k, cliques, kill, kp, off, adj, mark, cc);
  return hipGetLastError();
}

__global__ void k_gather_u32(uint32_t m, const uint32_t *__restrict__ idx,
                             const uint32_t *__restrict__ src, uint32_t *__restrict__ dst) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

__global__ void k_scatter_u32x2(uint32_t m, const uint32_t *__restrict__ idx,
                                const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                                uint32_t *__restrict__ da, uint32_t *__restrict__ db) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    da[idx[i]] = a[i];
    db[idx[i]] = b[i];
  }
}

__global__ void k_scatter_u32(uint32_t m, const uint32_t *__restrict__ idx,
                              const uint32_t *__restrict__ a, uint32_t *__restrict__ dst) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
    dst[idx[i]] = a[i];
}

hipError_t scatter_u32(uint32_t m, const uint32_t *idx, const uint32_t *a, uint32_t *dst,
                       hipStream_t s) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_u32, dim3(grid_for(m, 256, 4096)), dim3(256), 0, s, m, idx, a,
                     dst);
  return hipGetLastError();
}

hipError_t gather_u32(uint32_t m, const uint32_t *idx, const uint32_t *src, uint32_t *dst,
                      hipStream_t s) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(m, 256, 4096)), dim3(256), 0, s, m, idx, src,
                     dst);
  return hipGetLastError();
}

hipError_t scatter_u32x2(uint32_t m, const uint32_t *idx, const uint32_t *a, const uint32_t *b,
                         uint32_t *da, uint32_t *db, hipStream_t s) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_u32x2, dim3(grid_for(m, 256, 4096)), dim3(256), 0, s, m, idx,
                     a, b, da, db);
  return hipGetLastError();
}

hipError_t count_nonzero_u8(size_t n, const uint8_t *a,
                            unsigned long long *out, hipStream_t s) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(*out), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_count_nonzero_u8, dim3(grid_for(n, 256, 8192)),
                     dim3(256), 0, s, n, a, out);
  return hipGetLastError();
}

hipError_t comp_cliques_build(uint64_t T, uint32_t k, const uint32_t *cliques,
                              const uint32_t *kill, uint32_t kp,
                              const uint32_t *comp_of, const uint32_t *fv_of,
                              uint32_t ncomp, const uint64_t *coff,
                              uint32_t *ccur, uint32_t *fmem, uint32_t *cslot,
                              hipStream_t s, uint32_t ms, const uint32_t *cc) {
  if (ms == 0) ms = k;
  if (T == 0) return hipSuccess;
  const unsigned grid = ncomp > kCompLds
                            ? grid_for(T, kCompThreads, 65535)
                            : grid_for(T, kCompThreads * kCompTile, 4096);
  hipLaunchKernelGGL(k_comp_scatter2, dim3(grid), dim3(kCompThreads), 0, s, T, k,
                     cliques, kill, kp, comp_of, fv_of, ncomp, coff, ccur, fmem,
                     cslot, ms, cc);
  return hipGetLastError();
}

__global__ void k_core_flow_degrees(uint32_t cn, const uint32_t *__restrict__ cv,
                                    const uint64_t *__restrict__ cd,
                                    const uint32_t *__restrict__ fv_of,
                                    uint32_t *__restrict__ fcd) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cn;
       i += gridDim.x * blockDim.x) {
    const uint32_t v = cv[i];
    fcd[fv_of[v]] = (uint32_t)cd[v];
  }
}

hipError_t core_flow_degrees(uint32_t cn, const uint32_t *cv, const uint64_t *cd,
                             const uint32_t *fv_of, uint32_t *fcd, hipStream_t s) {
  if (cn == 0) return hipSuccess;
  hipLaunchKernelGGL(k_core_flow_degrees, dim3(grid_for(cn, 256, 4096)), dim3(256), 0, s, cn,
                     cv, cd, fv_of, fcd);
  return hipGetLastError();
}

hipError_t comp_clique_totals(uint64_t T, uint32_t k, const uint32_t *cliques,
                              const uint32_t *kill, uint32_t kp,
                              const uint32_t *comp_of, uint32_t ncomp,
                              uint32_t *ccount, hipStream_t s, const uint32_t *cc) {
  if (T == 0) return hipSuccess;
  const unsigned grid = grid_for(T, kCompThreads, ncomp > kCompLds ? 65535 : 4096);
  hipLaunchKernelGGL(k_comp_count2, dim3(grid), dim3(kCompThreads), 0, s, T,
                     cliques, k, kill, kp, comp_of, ncomp, ccount, cc);
  return hipGetLastError();
}

// Cores of at most kFlowDirect vertices (the usual case: a densest core is
// small): counters indexed by vertex in LDS, one contiguous clique range
// per workgroup, one global atomic per (workgroup, vertex).  The fill pass
// counts its range, reserves one run per vertex, then sweeps the range
// again placing each pair at its run's LDS cursor.
constexpr uint32_t kFlowDirect = 32768;

__device__ __forceinline__ void flow_range(uint64_t Tc, uint64_t *lo, uint64_t *hi) {
  const uint64_t per = (Tc + gridDim.x - 1) / gridDim.x;
  *lo = min<uint64_t>(Tc, (uint64_t)blockIdx.x * per);
  *hi = min<uint64_t>(Tc, *lo + per);
}

__global__ __launch_bounds__(kFlowThreads) void k_flow_deg_direct(
    uint64_t Tc, uint32_t k, const uint32_t *__restrict__ fmem, uint32_t nv,
    uint32_t *__restrict__ fcd, uint32_t ms) {
  __shared__ uint32_t cnt[kFlowDirect];
  for (uint32_t i = threadIdx.x; i < nv; i += kFlowThreads) cnt[i] = 0;
  __syncthreads();
  uint64_t lo, hi;
  flow_range(Tc, &lo, &hi);
  for (uint64_t p = lo * k + threadIdx.x; p < hi * k; p += kFlowThreads)
    atomicAdd(&cnt[member_at(fmem, p, k, ms)], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nv; i += kFlowThreads)
    if (cnt[i]) atomicAdd(&fcd[i], cnt[i]);
}

__global__ __launch_bounds__(kFlowThreads) void k_flow_fill_direct(
    uint64_t Tc, uint32_t k, const uint32_t *__restrict__ fmem, uint32_t nv,
    const uint64_t *__restrict__ fioff, uint32_t *__restrict__ fcur,
    uint32_t *__restrict__ finc, uint32_t div, uint32_t ms) {
  __shared__ uint32_t cnt[kFlowDirect];
  for (uint32_t i = threadIdx.x; i < nv; i += kFlowThreads) cnt[i] = 0;
  __syncthreads();
  uint64_t lo, hi;
  flow_range(Tc, &lo, &hi);
  for (uint64_t p = lo * k + threadIdx.x; p < hi * k; p += kFlowThreads)
    atomicAdd(&cnt[member_at(fmem, p, k, ms)], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nv; i += kFlowThreads)
    if (cnt[i]) cnt[i] = atomicAdd(&fcur[i], cnt[i]);  // -> this range's run
  __syncthreads();
  for (uint64_t p = lo * k + threadIdx.x; p < hi * k; p += kFlowThreads) {
    const uint32_t x = member_at(fmem, p, k, ms);
    finc[fioff[x] + atomicAdd(&cnt[x], 1u)] = (uint32_t)(p / div);
  }
}

// The same two kernels by clique for k = 2, 3, 4: one vector load of a
// record's K members (no slot division), counters in dynamic LDS sized to
// the core, so small cores leave room for several workgroups per CU.
// Largest core the by-clique kernels take (its counters in 64 KB of LDS).
constexpr uint32_t kFlowCliqueLds = 16384;

template <uint32_t K>
__device__ __forceinline__ void record_members(const uint32_t *__restrict__ fmem, uint64_t c,
                                               uint32_t ms, uint32_t (&m)[K]) {
  const uint32_t *r = fmem + c * ms;
  if (K == 4 && (ms & 3) == 0) {
    const uint4 v = *reinterpret_cast<const uint4 *>(r);
    m[0] = v.x;
    m[1] = v.y;
    m[2] = v.z;
    m[3] = v.w;
  } else if (K == 3) {
    const Words3 v = *reinterpret_cast<const Words3 *>(r);
    m[0] = v.w[0];
    m[1] = v.w[1];
    m[2] = v.w[2];
  } else if (K == 2 && (ms & 1) == 0) {
    const uint2 v = *reinterpret_cast<const uint2 *>(r);
    m[0] = v.x;
    m[1] = v.y;
  } else {
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) m[j] = r[j];
  }
}

template <uint32_t K>
__global__ __launch_bounds__(kFlowThreads) void k_flow_deg_cliques(
    uint64_t Tc, const uint32_t *__restrict__ fmem, uint32_t nv, uint32_t *__restrict__ fcd,
    uint32_t ms) {
  extern __shared__ uint32_t lcnt[];
  for (uint32_t i = threadIdx.x; i < nv; i += kFlowThreads) lcnt[i] = 0;
  __syncthreads();
  uint64_t lo, hi;
  flow_range(Tc, &lo, &hi);
  for (uint64_t c = lo + threadIdx.x; c < hi; c += kFlowThreads) {
    uint32_t m[K];
    record_members<K>(fmem, c, ms, m);
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) atomicAdd(&lcnt[m[j]], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nv; i += kFlowThreads)
    if (lcnt[i]) atomicAdd(&fcd[i], lcnt[i]);
}

template <uint32_t K>
__global__ __launch_bounds__(kFlowThreads) void k_flow_fill_cliques(
    uint64_t Tc, const uint32_t *__restrict__ fmem, uint32_t nv,
    const uint64_t *__restrict__ fioff, uint32_t *__restrict__ fcur,
    uint32_t *__restrict__ finc, bool by_clique, uint32_t ms) {
  extern __shared__ uint32_t lcnt[];
  for (uint32_t i = threadIdx.x; i < nv; i += kFlowThreads) lcnt[i] = 0;
  __syncthreads();
  uint64_t lo, hi;
  flow_range(Tc, &lo, &hi);
  for (uint64_t c = lo + threadIdx.x; c < hi; c += kFlowThreads) {
    uint32_t m[K];
    record_members<K>(fmem, c, ms, m);
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) atomicAdd(&lcnt[m[j]], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nv; i += kFlowThreads)
    if (lcnt[i]) lcnt[i] = atomicAdd(&fcur[i], lcnt[i]);  // -> this range's run
  __syncthreads();
  for (uint64_t c = lo + threadIdx.x; c < hi; c += kFlowThreads) {
    uint32_t m[K];
    record_members<K>(fmem, c, ms, m);
#pragma unroll
    for (uint32_t j = 0; j < K; ++j)
      finc[fioff[m[j]] + atomicAdd(&lcnt[m[j]], 1u)] =
          by_clique ? (uint32_t)c : (uint32_t)(c * K + j);
  }
}

hipError_t flow_incidence_degrees(uint64_t Tc, uint32_t k, const uint32_t *fmem,
                                  uint32_t nv, uint32_t *fcd, hipStream_t s, uint32_t ms) {
  if (k > 4) return hipErrorInvalidValue;
  if (ms == 0) ms = k;
  hipError_t e = hipMemsetAsync(fcd, 0, (size_t)(nv ? nv : 1) * sizeof(uint32_t), s);
  if (e != hipSuccess || Tc == 0) return e;
  const size_t lds = (size_t)nv * sizeof(uint32_t);
  const unsigned dgrid = grid_for(Tc, 16 * kFlowTile, 1024);
  if (nv <= kFlowCliqueLds && k == 4)
    hipLaunchKernelGGL(k_flow_deg_cliques<4>, dim3(dgrid), dim3(kFlowThreads), lds, s, Tc,
                       fmem, nv, fcd, ms);
  else if (nv <= kFlowCliqueLds && k == 3)
    hipLaunchKernelGGL(k_flow_deg_cliques<3>, dim3(dgrid), dim3(kFlowThreads), lds, s, Tc,
                       fmem, nv, fcd, ms);
  else if (nv <= kFlowCliqueLds && k == 2)
    hipLaunchKernelGGL(k_flow_deg_cliques<2>, dim3(dgrid), dim3(kFlowThreads), lds, s, Tc,
                       fmem, nv, fcd, ms);
  else if (nv <= kFlowDirect)
    hipLaunchKernelGGL(k_flow_deg_direct, dim3(dgrid), dim3(kFlowThreads), 0, s, Tc, k, fmem,
                       nv, fcd, ms);
  else
    hipLaunchKernelGGL(k_flow_deg, dim3(grid_for(Tc, kFlowTile, 4096)), dim3(kFlowThreads),
                       0, s, Tc, k, fmem, fcd, ms);
  return hipGetLastError();
}

hipError_t flow_incidence_fill(uint64_t Tc, uint32_t k, const uint32_t *fmem,
                               const uint64_t *fioff, uint32_t nv, uint32_t *fcur,
                               uint32_t *finc, bool by_clique, hipStream_t s,
                               uint32_t ms) {
  if (k > 4 || k < 1) return hipErrorInvalidValue;
  if (ms == 0) ms = k;
  const uint32_t div = by_clique ? k : 1;
  hipError_t e = hipMemsetAsync(fcur, 0, (size_t)(nv ? nv : 1) * sizeof(uint32_t), s);
  if (e != hipSuccess || Tc == 0) return e;
  const size_t lds = (size_t)nv * sizeof(uint32_t);
  const unsigned dgrid = grid_for(Tc, 16 * kFlowTile, 1024);
  if (nv <= kFlowCliqueLds && k == 4)
    hipLaunchKernelGGL(k_flow_fill_cliques<4>, dim3(dgrid), dim3(kFlowThreads), lds, s, Tc,
                       fmem, nv, fioff, fcur, finc, by_clique, ms);
  else if (nv <= kFlowCliqueLds && k == 3)
    hipLaunchKernelGGL(k_flow_fill_cliques<3>, dim3(dgrid), dim3(kFlowThreads), lds, s, Tc,
                       fmem, nv, fioff, fcur, finc, by_clique, ms);
  else if (nv <= kFlowCliqueLds && k == 2)
    hipLaunchKernelGGL(k_flow_fill_cliques<2>, dim3(dgrid), dim3(kFlowThreads), lds, s, Tc,
                       fmem, nv, fioff, fcur, finc, by_clique, ms);
  else if (nv <= kFlowDirect)
    hipLaunchKernelGGL(k_flow_fill_direct, dim3(dgrid), dim3(kFlowThreads), 0, s, Tc, k,
                       fmem, nv, fioff, fcur, finc, div, ms);
  else
    hipLaunchKernelGGL(k_flow_fill, dim3(grid_for(Tc, kFlowTile, 4096)), dim3(kFlowThreads),
                       0, s, Tc, k, fmem, fioff, fcur, finc, div, ms);
  return hipGetLastError();
}

}  // namespace cds
