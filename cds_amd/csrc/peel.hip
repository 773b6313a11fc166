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
#include <algorithm>
#include <cstdlib>

#include "dev.h"
#include "kernels.h"

namespace cds {
namespace {

// Incidence entries are k words: the clique id, then its other k-1
// members, so the peel streams the members instead of gathering them.
// Slots are reserved with wave-aggregated cursor increments: consecutive
// cliques share their lowest vertex (they come from the same listing item)
// and hubs recur, so most lanes of a wave share one atomic.
// Incidence build.  A workgroup takes a tile of kIncPairs / k cliques and
// aggregates its (vertex, clique) pairs in an LDS hash table keyed by
// vertex: every pair gets its rank among the tile's pairs of that vertex
// from an LDS atomic, then one global atomicAdd per distinct vertex of the
// tile reserves a run of the vertex's incidence list.  Hub vertices, which
// appear in most cliques of a power-law graph, thus cost one global atomic
// per tile instead of one per clique.
// Tiles of 8192 (vertex, clique) pairs, 8192/k cliques (128 KB of LDS, one
// workgroup per CU): larger tiles share more reservations; on R-MAT scale
// 20 k=3 the build took 71 ms with 512-clique tiles, 41 ms with 1024 and
// 29 ms with 2048.
constexpr uint32_t kIncThreads = 1024;
constexpr uint32_t kIncPairs = 8192;                     // pairs per tile: 8192/k cliques
constexpr uint32_t kIncHash = 2 * kIncPairs;             // load <= 1/2
constexpr uint32_t kIncPerThread = kIncPairs / kIncThreads;

// An entry of k words stored with one instruction (k = 3: 12 bytes at
// 4-byte alignment), so a scattered entry costs one memory request.
struct __attribute__((packed, aligned(4))) Entry3 {
  uint32_t w[3];
};
struct __attribute__((packed, aligned(4))) Entry4 {
  uint32_t w[4];
};

// With small (T * k < 2^32) the reservation pass turns each vertex's slot
// count into the absolute position of its run, so the write pass reads no
// ioff word per pair.
__global__ __launch_bounds__(kIncThreads) void k_build_inc(
    uint64_t T, uint32_t k, const uint32_t *__restrict__ cl,
    const uint64_t *__restrict__ ioff, uint32_t *__restrict__ icur,
    uint32_t *__restrict__ inc, uint32_t small) {
  __shared__ uint32_t hkey[kIncHash];
  __shared__ uint32_t hcnt[kIncHash];
  // occupied slots of the tile, so the reservation pass and the reset visit
  // the tile's distinct vertices only (a tile holds at most kIncPairs)
  __shared__ uint16_t used[kIncPairs];
  __shared__ uint32_t nused;
  static_assert(kIncHash <= 65536, "slot indices are 16-bit");
  const uint32_t tile_c = kIncPairs / k;  // cliques per tile: kIncPairs pairs
  const uint64_t tiles = (T + tile_c - 1) / tile_c;
  for (uint32_t i = threadIdx.x; i < kIncHash; i += kIncThreads) {
    hkey[i] = kAlive;
    hcnt[i] = 0;
  }
  if (threadIdx.x == 0) nused = 0;
  __syncthreads();
  for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const uint64_t c0 = tile * tile_c;
    const uint32_t pairs = (uint32_t)min<uint64_t>(tile_c, T - c0) * k;
    uint32_t slot[kIncPerThread], rank[kIncPerThread];
#pragma unroll
    for (uint32_t r = 0; r < kIncPerThread; ++r) {
      const uint32_t p = r * kIncThreads + threadIdx.x;
      if (p >= pairs) break;
      const uint32_t x = cl[c0 * k + p];
      uint32_t h = (x * 2654435761u) & (kIncHash - 1);
      for (;;) {
        const uint32_t prev = atomicCAS(&hkey[h], kAlive, x);
        if (prev == kAlive) {
          used[atomicAdd(&nused, 1u)] = (uint16_t)h;
          break;
        }
        if (prev == x) break;
        h = (h + 1) & (kIncHash - 1);
      }
      slot[r] = h;
      rank[r] = atomicAdd(&hcnt[h], 1u);
    }
    __syncthreads();
    const uint32_t nu = nused;
    for (uint32_t i = threadIdx.x; i < nu; i += kIncThreads) {
      const uint32_t h = used[i], x = hkey[h];
      const uint32_t r0 = atomicAdd(&icur[x], hcnt[h]);  // -> base
      hcnt[h] = small ? (uint32_t)(ioff[x] + r0) : r0;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kIncPerThread; ++r) {
      const uint32_t p = r * kIncThreads + threadIdx.x;
      if (p >= pairs) break;
      const uint32_t q = p / k, j = p - q * k;
      const uint64_t c = c0 + q;
      const uint64_t at = small ? (uint64_t)hcnt[slot[r]] + rank[r]
                                : ioff[hkey[slot[r]]] + hcnt[slot[r]] + rank[r];
      uint32_t *e = inc + at * k;
      if (k == 3) {
        const uint32_t a = cl[c * 3 + (j == 0 ? 1 : 0)], b = cl[c * 3 + (j == 2 ? 1 : 2)];
        *reinterpret_cast<Entry3 *>(e) = Entry3{{(uint32_t)c, a, b}};
      } else if (k == 4) {
        Entry4 v;
        v.w[0] = (uint32_t)c;
        for (uint32_t t = 0, o = 1; t < 4; ++t)
          if (t != j) v.w[o++] = cl[c * 4 + t];
        *reinterpret_cast<Entry4 *>(e) = v;
      } else {
        e[0] = (uint32_t)c;
        for (uint32_t t = 0, o = 1; t < k; ++t)
          if (t != j) e[o++] = cl[c * k + t];
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nu; i += kIncThreads) {
      const uint32_t h = used[i];
      hkey[h] = kAlive;
      hcnt[h] = 0;
    }
    if (threadIdx.x == 0) nused = 0;
    __syncthreads();
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
      const uint32_t *en = inc + q * k;
      const uint32_t c = en[0];
      if (kill[c] != kAlive) continue;
      if (atomicCAS(&kill[c], kAlive, level) != kAlive) continue;
      ++nk;
      for (uint32_t t = 1; t < k; ++t) {
        const uint32_t u = en[t];
        if (core[u] != kAlive) continue;
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

// ---- persistent peel: every level in one cooperative launch ----
//
// All workgroups are co-resident (hipLaunchCooperativeKernel), so a grid
// barrier is a counter plus a generation word.  Per level: (1) min degree
// over the alive list, (2) the frontier (alive, degree <= level) and the
// compacted alive list for the next level, (3) sub-rounds of frontier
// processing until no vertex drops to the level.  Frontier work is split
// into chunks of kPeelChunk incidence entries; one packed 64-bit atomic per
// appended vertex (vertex count << 40 | chunk count) gives every vertex its
// slot and its first chunk, so chunk i maps to its vertex by a binary search
// and a hub's incidence list is spread over many waves.
// Cliques are claimed without per-clique state.  Every vertex gets a stamp,
// the sub-round whose frontier it joins (stamps grow over the launch).  An
// entry (v; c, other members) processed in sub-round r is dead if another
// member has a stamp below r (c was claimed then), and is claimed by v iff
// no other member with stamp r has a smaller id, so exactly one member of
// each frontier claims each live clique.  Stamps written during sub-round r
// (r + 1, for vertices joining the next frontier) compare like unstamped
// ones, so the race with them decides nothing.  The stamp reads are k-1
// random 4-byte loads into an n-word array that stays in cache; the CAS on
// a T-word kill array they replace cost one HBM line per entry (measured 57
// bytes of traffic per entry at R-MAT scale 20, k = 3; a one-bit-per-clique
// set with atomicOr was slower still, consecutive clique ids of a hub
// serialising on one word).  Kill words are derived after the launch
// (k_kill_from_core).
constexpr uint32_t kPeelChunk = 64;     // one wave pass per chunk
// Two member words read with one 8-byte load at 4-byte alignment.
struct __attribute__((packed, aligned(4))) Members2 {
  uint32_t x, y;
  __device__ operator uint2() const { return make_uint2(x, y); }
};
constexpr uint32_t kPeelBatch = 2;      // chunks per wave pass with loads in flight
constexpr int kPeelThreads = 1024;
constexpr uint64_t kChunkMask = (1ull << 40) - 1;
// alive vertices aggregated in LDS (128 KB; the persistent peel runs one
// workgroup per CU): 8192 -> 32768 cut the R-MAT scale 20 k=3 peel from 127
// to 107 ms, the coarse steps reaching the aggregated regime sooner
constexpr uint32_t kPeelLdsCap = 32768;

// While many vertices are alive, degree decrements are aggregated per
// workgroup and sub-round in an LDS hash table over the same 128 KB: keys
// (vertex + 1, 0 = empty) in the first half, counts in the second, linear
// probing over kPeelHashProbe slots.  Decrements of a power-law graph
// concentrate on hub vertices, which recur across a workgroup's chunks, so
// most reach HBM as one atomic per (hub, workgroup, sub-round).  A
// decrement that finds no slot goes to the global degree directly.  Every
// atomic on deg[u] applies its count at once, so exactly one of them takes
// it from above the level to at most the level, and that one appends u.
constexpr uint32_t kPeelHashSlots = kPeelLdsCap / 2;
constexpr uint32_t kPeelHashProbe = 8;

__device__ __forceinline__ bool hash_add(uint32_t *table, uint32_t u) {
  const uint32_t key = u + 1;
  uint32_t h = (key * 2654435761u) >> 18;  // 14 bits: kPeelHashSlots
  for (uint32_t p = 0; p < kPeelHashProbe; ++p) {
    uint32_t cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == 0) cur = atomicCAS(&table[h], 0u, key);
    if (cur == 0 || cur == key) {
      atomicAdd(&table[kPeelHashSlots + h], 1u);
      return true;
    }
    h = (h + 1) & (kPeelHashSlots - 1);
  }
  return false;
}

__device__ __forceinline__ void grid_sync(PeelCtl *ctl) {
  const unsigned long long t0 =
      (blockIdx.x == 0 && threadIdx.x == 0) ? wall_clock64() : 0;
  grid_barrier(&ctl->bar);
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->t_bar += wall_clock64() - t0;
}

__device__ __forceinline__ uint32_t ld(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t chunks_of(uint64_t len) {
  return (len + kPeelChunk - 1) / kPeelChunk;
}

// Appends v (with its chunk count) to frontier buffer b.
__device__ __forceinline__ void frontier_push(PeelCtl *ctl, int b, uint32_t *F,
                                              uint64_t *FC, uint32_t v,
                                              uint64_t nch) {
  const unsigned long long old =
      atomicAdd(&ctl->fpack[b], (1ull << 40) + (unsigned long long)nch);
  const uint32_t slot = (uint32_t)(old >> 40);
  F[slot] = v;
  FC[slot] = old & kChunkMask;
}

// A wave's walk over its contiguous chunk range: the current vertex slot lo
// and vertex v, v's incidence range [vb, ve), v's first chunk vfc, the next
// vertex's first chunk nfc, and q0, an entry of the range.
struct PeelWalk {
  uint32_t lo, v;
  uint64_t vb, ve, vfc, nfc, q0;
};

// kPeelBatch chunks [ci, ci + kPeelBatch) of the wave's range (below c1):
// the entries' K-1 other members, their stamps and (AGG) alive-list
// positions are all loaded before any is used; every load is unconditional
// (a lane with no entry re-reads entry q0) and branch-free, so they are in
// flight together.  Then the claims and the degree decrements.
template <uint32_t K, bool AGG>
__device__ __forceinline__ void peel_batch(
    uint64_t ci, uint64_t c1, PeelWalk &wk, uint32_t fcount, const uint32_t *Fc,
    const uint64_t *FCc, const uint64_t *__restrict__ ioff, const uint32_t *__restrict__ inc,
    uint32_t *stamp, const uint32_t *pos, uint32_t *lds_cnt, uint32_t hagg, uint32_t *deg,
    uint32_t *core, uint32_t sr, uint32_t level, PeelCtl *ctl, int nb, uint32_t *Fn,
    uint64_t *FCn, uint32_t &nk) {
  constexpr uint32_t KM = K - 1;
  uint32_t bv[kPeelBatch];
  uint64_t bq[kPeelBatch];
  bool bok[kPeelBatch];
#pragma unroll
  for (uint32_t s = 0; s < kPeelBatch; ++s) {
    const uint64_t cs = ci + s;
    bok[s] = false;
    bv[s] = 0;
    bq[s] = wk.q0;
    if (cs < c1) {
      while (cs >= wk.nfc) {  // next vertex (empty lists own no chunk)
        ++wk.lo;
        wk.v = Fc[wk.lo];
        wk.vb = ioff[wk.v];
        wk.ve = ioff[wk.v + 1];
        wk.vfc = wk.nfc;
        wk.nfc = wk.lo + 1 < fcount ? FCc[wk.lo + 1] : ~0ull;
      }
      const uint64_t q = wk.vb + (cs - wk.vfc) * kPeelChunk + lane();
      bok[s] = q < wk.ve;
      bv[s] = wk.v;
      bq[s] = bok[s] ? q : wk.q0;
    }
  }
  uint32_t mu[kPeelBatch][KM], msu[kPeelBatch][KM], mpos[kPeelBatch][KM];
#pragma unroll
  for (uint32_t s = 0; s < kPeelBatch; ++s) {
    if (K == 3) {  // both members in one 8-byte load
      const uint2 m2 = *reinterpret_cast<const Members2 *>(inc + bq[s] * 3 + 1);
      mu[s][0] = m2.x;
      mu[s][KM - 1] = m2.y;
    } else {
#pragma unroll
      for (uint32_t t = 0; t < KM; ++t) mu[s][t] = inc[bq[s] * K + t + 1];
    }
  }
#pragma unroll
  for (uint32_t s = 0; s < kPeelBatch; ++s)
#pragma unroll
    for (uint32_t t = 0; t < KM; ++t) {
      msu[s][t] = ld(&stamp[mu[s][t]]);
      mpos[s][t] = AGG ? pos[mu[s][t]] : 0u;
    }
#pragma unroll
  for (uint32_t s = 0; s < kPeelBatch; ++s) {
    if (!bok[s]) continue;
    bool claim = true;
#pragma unroll
    for (uint32_t t = 0; t < KM; ++t) {
      const uint32_t su = msu[s][t];
      claim = claim && su >= sr && (su != sr || mu[s][t] > bv[s]);
    }
    if (!claim) continue;
    ++nk;
#pragma unroll
    for (uint32_t t = 0; t < KM; ++t) {
      // members of this frontier (stamp sr) are removed and parked ones
      // keep their core value in deg: neither is decremented
      const uint32_t su = msu[s][t], u = mu[s][t];
      if (su == sr || su == kParked) continue;
      if (AGG) {
        atomicAdd(&lds_cnt[mpos[s][t]], 1u);
        continue;
      }
      if (hagg && hash_add(lds_cnt, u)) continue;
      const uint32_t old = atomicSub(&deg[u], 1u);
      if (old == level + 1) {
        core[u] = level;
        stamp[u] = sr + 1;
        frontier_push(ctl, nb, Fn, FCn, u, chunks_of(ioff[u + 1] - ioff[u]));
      }
    }
  }
}

__global__ __launch_bounds__(kPeelThreads) void k_peel_persistent(
    uint32_t n, uint32_t k, const uint64_t *__restrict__ ioff,
    const uint32_t *__restrict__ inc, const uint32_t *__restrict__ cl,
    uint32_t *deg, uint32_t *core, uint32_t *stamp, uint32_t sbase, uint32_t *A0,
    uint32_t *A1,
    uint32_t *F0, uint32_t *F1, uint32_t *F2, uint64_t *FC0, uint64_t *FC1,
    uint64_t *FC2, PeelCtl *ctl, uint32_t *lv_level,
    unsigned long long *lv_removed, unsigned long long *lv_killed,
    uint32_t *pos, uint32_t *sdeg, uint32_t *lv_min, uint32_t cshift,
    uint32_t max_level, uint32_t hagg) {
  // Once few vertices remain alive, degree decrements are summed per
  // workgroup in LDS (indexed by position in the alive list) and applied
  // with one atomic per vertex and sub-round; the vertex whose aggregated
  // decrement crosses level+1 -> level joins the next sub-frontier.
  __shared__ uint32_t lds_cnt[kPeelLdsCap];
  // Per-wave partials for workgroup-level aggregation: counters every
  // workgroup updates (level minimum, killed cliques, frontier and alive
  // list reservations) see one atomic per workgroup instead of per wave.
  constexpr int kPeelWaves = kPeelThreads / kWave;
  __shared__ uint32_t sh_min[kPeelWaves], sh_an[kPeelWaves], sh_fn[kPeelWaves];
  __shared__ unsigned long long sh_kill[kPeelWaves], sh_fc[kPeelWaves];
  __shared__ unsigned long long sh_fbase;
  __shared__ uint32_t sh_abase;
  const uint32_t wid = threadIdx.x / kWave, nwaves = blockDim.x / kWave;
  for (uint32_t i = threadIdx.x; i < kPeelLdsCap; i += blockDim.x) lds_cnt[i] = 0;
  __syncthreads();
  uint32_t *A[2] = {A0, A1};
  uint32_t *F[3] = {F0, F1, F2};
  uint64_t *FC[3] = {FC0, FC1, FC2};
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t gthreads = gridDim.x * blockDim.x;
  const uint32_t gwave = gtid / kWave, gwaves = gthreads / kWave;
  const bool leader = gtid == 0;
  int ab = 0, gp = 0, fb = 0;
  uint32_t lvl_idx = 0;
  uint32_t sr = sbase;  // current sub-round, identical in every thread
  for (;;) {
    // (1) minimum degree among alive vertices
    unsigned long long ts = leader ? wall_clock64() : 0;
    const uint32_t na = ld(&ctl->alive_n[ab]);
    uint32_t m = kAlive;
    for (uint32_t i = gtid; i < na; i += gthreads) {
      const uint32_t v = A[ab][i];
      if (ld(&core[v]) == kAlive) {
        // degree at the start of the level; kept for the level that
        // removes v (its clique count in the core C_level)
        const uint32_t d = ld(&deg[v]);
        sdeg[v] = d;
        m = min(m, d);
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = min(m, (uint32_t)__shfl_xor(m, off, kWave));
    if (lane() == 0) sh_min[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t bm = kAlive;
      for (uint32_t w = 0; w < nwaves; ++w) bm = min(bm, sh_min[w]);
      if (bm != kAlive) atomicMin(&ctl->gmin[gp], bm);
    }
    if (leader) {
      ctl->alive_n[1 - ab] = 0;
      ctl->t_scan += wall_clock64() - ts;
    }
    grid_sync(ctl);
    // The step removes every alive vertex whose degree drops to `level`:
    // the minimum degree (one exact level), or with cshift > 0 the range
    // [min, min + min >> cshift] at once (a coarse step: its start and end
    // are exact clique cores, the core numbers inside it are not).
    const uint32_t mlev = ld(&ctl->gmin[gp]);
    if (mlev == kAlive) break;
    const uint32_t level =
        (cshift && mlev < max_level) ? mlev + min(mlev >> cshift, max_level - mlev) : mlev;
    if (leader) {
      ctl->gmin[1 - gp] = kAlive;
      ts = wall_clock64();
    }
    // (2) frontier + compacted alive list
    for (uint32_t base = blockIdx.x * blockDim.x; base < na; base += gthreads) {
      const uint32_t i = base + threadIdx.x;
      uint32_t v = 0;
      bool alive = false, take = false;
      if (i < na) {
        v = A[ab][i];
        alive = ld(&core[v]) == kAlive;
        take = alive && ld(&deg[v]) <= level;
      }
      const bool keep = alive && !take;
      // inclusive wave scans: frontier count and chunks, alive count
      unsigned long long fch = take ? chunks_of(ioff[v + 1] - ioff[v]) : 0;
      uint32_t fcnt = take ? 1u : 0u, acnt = keep ? 1u : 0u;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const unsigned long long t = __shfl_up(fch, off, kWave);
        const uint32_t tf = __shfl_up(fcnt, off, kWave);
        const uint32_t ta = __shfl_up(acnt, off, kWave);
        if ((int)lane() >= off) {
          fch += t;
          fcnt += tf;
          acnt += ta;
        }
      }
      if (lane() == kWave - 1) {
        sh_fc[wid] = fch;
        sh_fn[wid] = fcnt;
        sh_an[wid] = acnt;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned long long rc = 0;
        uint32_t rf = 0, ra = 0;
        for (uint32_t w = 0; w < nwaves; ++w) {
          const unsigned long long c = sh_fc[w];
          const uint32_t f = sh_fn[w], a = sh_an[w];
          sh_fc[w] = rc;
          sh_fn[w] = rf;
          sh_an[w] = ra;
          rc += c;
          rf += f;
          ra += a;
        }
        sh_fbase = rf ? atomicAdd(&ctl->fpack[fb], ((unsigned long long)rf << 40) + rc) : 0;
        sh_abase = ra ? atomicAdd(&ctl->alive_n[1 - ab], ra) : 0;
      }
      __syncthreads();
      if (take) {
        core[v] = level;
        stamp[v] = sr;
        const uint32_t slot = (uint32_t)(sh_fbase >> 40) + sh_fn[wid] + fcnt - 1;
        F[fb][slot] = v;
        FC[fb][slot] = (sh_fbase & kChunkMask) + sh_fc[wid] + fch -
                       chunks_of(ioff[v + 1] - ioff[v]);
      }
      if (keep) {
        const uint32_t slot = sh_abase + sh_an[wid] + acnt - 1;
        A[1 - ab][slot] = v;
        pos[v] = slot;
      }
    }
    if (leader) ctl->t_scan += wall_clock64() - ts;
    grid_sync(ctl);
    const uint32_t na_next = ld(&ctl->alive_n[1 - ab]);
    const bool agg = na_next <= kPeelLdsCap;
    // (3) sub-rounds
    unsigned long long removed = 0;
    for (;;) {
      const unsigned long long pack = __hip_atomic_load(
          &ctl->fpack[fb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t fcount = (uint32_t)(pack >> 40);
      const uint64_t nchunks = pack & kChunkMask;
      if (fcount == 0) break;
      removed += fcount;
      const int nb = (fb + 1) % 3, rb = (fb + 2) % 3;
      if (leader) ctl->fpack[rb] = 0;  // used as "next" one sub-round later
      const uint32_t *Fc = F[fb];
      const uint64_t *FCc = FC[fb];
      uint32_t nk = 0;
      unsigned long long tp = leader ? wall_clock64() : 0;
      // Each wave takes a contiguous range of the sub-round's chunks: one
      // search finds the first chunk's vertex, later chunks walk forward
      // (a vertex change is rare next to the chunks of the large lists), and
      // kPeelBatch chunks per pass have their loads in flight together.
      const uint64_t per = (nchunks + gwaves - 1) / gwaves;
      const uint64_t c0 = min(nchunks, (uint64_t)gwave * per);
      const uint64_t c1 = min(nchunks, c0 + per);
      if (c0 < c1) {
        uint32_t lo = wave_last_le(FCc, fcount, c0);  // c0's vertex slot
        uint32_t v = Fc[lo];
        uint64_t vb = ioff[v], ve = ioff[v + 1], vfc = FCc[lo];
        const uint64_t q0 = vb + (c0 - vfc) * kPeelChunk;  // an entry of the range
        uint64_t nfc = lo + 1 < fcount ? FCc[lo + 1] : ~0ull;
        PeelWalk wk{lo, v, vb, ve, vfc, nfc, q0};
        for (uint64_t ci = c0; ci < c1; ci += kPeelBatch) {
#define CDS_PEEL_BATCH(K)                                                                  \
  (agg ? peel_batch<K, true>(ci, c1, wk, fcount, Fc, FCc, ioff, inc, stamp, pos, lds_cnt, \
                             hagg, deg, core, sr, level, ctl, nb, F[nb], FC[nb], nk)       \
       : peel_batch<K, false>(ci, c1, wk, fcount, Fc, FCc, ioff, inc, stamp, pos, lds_cnt,\
                              hagg, deg, core, sr, level, ctl, nb, F[nb], FC[nb], nk))
          if (k == 3)
            CDS_PEEL_BATCH(3);
          else if (k == 2)
            CDS_PEEL_BATCH(2);
          else
            CDS_PEEL_BATCH(4);
#undef CDS_PEEL_BATCH
        }
      }
      if (leader) {
        const unsigned long long now = wall_clock64();
        ctl->t_proc += now - tp;
        tp = now;
      }
      if (hagg && !agg) {
        // apply the workgroup's aggregated decrements, one atomic per vertex
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < kPeelHashSlots; i += blockDim.x) {
          const uint32_t key = lds_cnt[i];
          if (!key) continue;
          const uint32_t cnt = lds_cnt[kPeelHashSlots + i];
          lds_cnt[i] = 0;
          lds_cnt[kPeelHashSlots + i] = 0;
          const uint32_t u = key - 1;
          const uint32_t old = atomicSub(&deg[u], cnt);
          if (old > level && old - cnt <= level) {
            core[u] = level;
            stamp[u] = sr + 1;
            frontier_push(ctl, nb, F[nb], FC[nb], u,
                          chunks_of(ioff[u + 1] - ioff[u]));
          }
        }
        __syncthreads();
        if (leader) ctl->t_flush += wall_clock64() - tp;
      }
      if (agg) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < na_next; i += blockDim.x) {
          const uint32_t cnt = lds_cnt[i];
          if (!cnt) continue;
          lds_cnt[i] = 0;
          const uint32_t u = A[1 - ab][i];
          const uint32_t old = atomicSub(&deg[u], cnt);
          if (old > level && old - cnt <= level) {
            core[u] = level;
            stamp[u] = sr + 1;
            frontier_push(ctl, nb, F[nb], FC[nb], u,
                          chunks_of(ioff[u + 1] - ioff[u]));
          }
        }
        __syncthreads();
        if (leader) ctl->t_flush += wall_clock64() - tp;
      }
      nk = wave_sum(nk);
      if (lane() == 0) sh_kill[wid] = nk;
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned long long bk = 0;
        for (uint32_t w = 0; w < nwaves; ++w) bk += sh_kill[w];
        if (bk) atomicAdd(&lv_killed[lvl_idx], bk);
      }
      grid_sync(ctl);
      fb = nb;
      ++sr;
      if (leader) ctl->rounds++;
    }
    // The last sub-round's barrier already ordered every update of this
    // level; the next min scan touches only buffers no block still reads.
    if (leader) {
      lv_level[lvl_idx] = level;
      lv_min[lvl_idx] = mlev;
      lv_removed[lvl_idx] = removed;
      ctl->removed += removed;
    }
    ++lvl_idx;
    ab = 1 - ab;
    gp = 1 - gp;
  }
  if (leader) ctl->nlevels = lvl_idx;
}

// ---- exact re-peel of one coarse step (refine_prepare / restore) ----

__global__ void k_refine_vertices(uint32_t n, uint32_t level,
                                  uint32_t *__restrict__ core,
                                  uint32_t *__restrict__ deg,
                                  const uint32_t *__restrict__ sdeg,
                                  uint32_t *__restrict__ alive,
                                  uint32_t *__restrict__ count) {
  for (uint32_t base = blockIdx.x * blockDim.x; base < n;
       base += gridDim.x * blockDim.x) {
    const uint32_t v = base + threadIdx.x;
    const uint32_t c = v < n ? core[v] : 0;
    const bool mine = v < n && c == level;
    if (v < n && c > level) {  // a later step: parked, value kept in deg
      deg[v] = c;
      core[v] = kParked;
    }
    if (mine) {
      core[v] = kAlive;
      deg[v] = sdeg[v];  // its degree when the step began
    }
    const uint32_t slot = aggregated_reserve(count, 0, mine);
    if (mine) alive[slot] = v;
  }
}

__global__ void k_refine_cliques(uint64_t T, uint32_t level,
                                 uint32_t *__restrict__ kill) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
       c += (uint64_t)gridDim.x * blockDim.x)
    if (kill[c] == level) kill[c] = kAlive;
}

__global__ void k_refine_restore(uint32_t n, uint32_t *__restrict__ core,
                                 const uint32_t *__restrict__ deg) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x)
    if (core[v] == kParked) core[v] = deg[v];
}

// Stamps before a launch: vertices the launch may meet alive compare as
// never removed: kAlive, or kParked for those parked in a later coarse step
// (they take part in claims but get no decrements: their deg holds their
// core value); vertices removed by earlier launches keep stamps below the
// launch's first sub-round.
__global__ void k_stamp_reset(uint32_t n, const uint32_t *__restrict__ core,
                              uint32_t *__restrict__ stamp, uint32_t first) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x) {
    const uint32_t c = core[v];
    if (c == kAlive || c == kParked) stamp[v] = c;
    else if (first) stamp[v] = 0;
  }
}

// After the persistent peel launches: every clique was claimed at the level
// its first member was removed, which is the smallest core value among its
// members.  Every kill word is written (no parked vertex may remain).
__global__ void k_kill_from_core(uint64_t T, uint32_t k,
                                 const uint32_t *__restrict__ cl,
                                 const uint32_t *__restrict__ core,
                                 uint32_t *__restrict__ kill) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
       c += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t m = kAlive;
    for (uint32_t j = 0; j < k; ++j) m = min(m, core[cl[c * k + j]]);
    kill[c] = m;
  }
}

__global__ void k_iota(uint32_t n, uint32_t *a) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    a[i] = i;
}

__global__ void k_peel_ctl_init(PeelCtl *ctl, uint32_t alive, unsigned long long limit) {
  unsigned *bw = reinterpret_cast<unsigned *>(&ctl->bar);
  for (size_t i = 0; i < sizeof(GridBarrier) / sizeof(unsigned); ++i) bw[i] = 0;
  ctl->bar.limit = limit;
  ctl->gmin[0] = kAlive;
  ctl->gmin[1] = kAlive;
  ctl->alive_n[0] = alive;
  ctl->alive_n[1] = 0;
  ctl->fpack[0] = ctl->fpack[1] = ctl->fpack[2] = 0;
  ctl->nlevels = 0;
  ctl->status = 0;
  ctl->removed = 0;
  ctl->rounds = 0;
  ctl->t_bar = ctl->t_scan = ctl->t_proc = ctl->t_flush = 0;
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
    uint32_t k, uint32_t level, const uint32_t *__restrict__ count,
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
      d += kill[inc[q * k]] == level;
    d = wave_sum(d);
    if (lane() == 0) {
      pdeg[i] = d;
      palive[i] = 1;
    }
  }
}

// Phase degrees from the degrees recorded at the start of the level.
__global__ void k_phase_deg_snap(const uint32_t *__restrict__ count,
                                 const uint32_t *__restrict__ pv,
                                 const uint32_t *__restrict__ sdeg,
                                 uint32_t *__restrict__ pdeg,
                                 uint8_t *__restrict__ palive) {
  const uint32_t M = *count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < M;
       i += gridDim.x * blockDim.x) {
    pdeg[i] = sdeg[pv[i]];
    palive[i] = 1;
  }
}

// Copies each phase vertex's phase-clique incidence entries (kill ==
// level) to pinc at pioff[i], in incidence order (ballot compaction).
__global__ __launch_bounds__(256) void k_phase_compact(
    uint32_t k, uint32_t level, const uint32_t *__restrict__ count,
    const uint32_t *__restrict__ pv, const uint64_t *__restrict__ ioff,
    const uint32_t *__restrict__ inc, const uint32_t *__restrict__ kill,
    const uint64_t *__restrict__ pioff, uint32_t *__restrict__ pinc) {
  const uint32_t M = *count;
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       i < M; i += waves) {
    const uint32_t v = pv[i];
    uint64_t out = pioff[i];
    const uint64_t e = ioff[v + 1];
    for (uint64_t q0 = ioff[v]; q0 < e; q0 += kWave) {
      const uint64_t q = q0 + lane();
      const bool take = q < e && kill[inc[q * k]] == level;
      unsigned cnt;
      const unsigned slot = wave_slot(take, &cnt);
      if (take)
        for (uint32_t t = 0; t < k; ++t) pinc[(out + slot) * k + t] = inc[q * k + t];
      out += cnt;
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
// density of the intermediate states (after 1..M-1 removals).  The phase's
// vertices are pv[base .. base+M); pdeg, palive and pioff are indexed by
// the same positions and posmap maps a vertex to its position.
__device__ void replay_one(uint32_t k, uint32_t level, uint32_t base,
                           uint32_t M, const uint32_t *__restrict__ pv,
                           uint32_t *__restrict__ pdeg,
                           const uint32_t *__restrict__ posmap,
                           uint8_t *__restrict__ palive,
                           const uint32_t *__restrict__ core,
                           const uint64_t *__restrict__ pioff,
                           const uint32_t *__restrict__ pinc,
                           uint32_t *__restrict__ rdead, uint64_t T_start,
                           uint64_t N_start, ReplayOut *out) {
  __shared__ uint64_t s_wmin[kReplayThreads / kWave];
  __shared__ uint64_t s_key;
  const uint32_t tid = threadIdx.x, wid = tid / kWave;
  uint64_t R = 0, best_num = 0, best_den = 1;
  uint32_t found = 0, steps = 0;
  for (uint32_t j = 0; j < M; ++j) {
    uint64_t key = ~0ull;
    for (uint32_t i = base + tid; i < base + M; i += kReplayThreads)
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
    // v's phase cliques (compacted beforehand), those not killed yet
    const uint32_t iv = posmap[v];
    const uint64_t e = pioff[iv + 1];
    for (uint64_t q = pioff[iv] + tid; q < e; q += kReplayThreads) {
      const uint32_t *en = pinc + q * k;
      const uint32_t c = en[0];
      if (rdead[c] == level + 1) continue;
      rdead[c] = level + 1;
      for (uint32_t t = 1; t < k; ++t) {
        const uint32_t u = en[t];
        if (core[u] == level) atomicSub(&pdeg[posmap[u]], 1u);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    out->best_num = best_num;
    out->best_den = best_den;
    out->steps = steps;
    out->found = found;
    out->stopped = 0;
  }
}

__global__ __launch_bounds__(kReplayThreads) void k_replay(
    uint32_t k, uint32_t level, const uint32_t *__restrict__ count,
    const uint32_t *__restrict__ pv, uint32_t *__restrict__ pdeg,
    const uint32_t *__restrict__ posmap, uint8_t *__restrict__ palive,
    const uint32_t *__restrict__ core, const uint64_t *__restrict__ pioff,
    const uint32_t *__restrict__ pinc, uint32_t *__restrict__ rdead,
    uint64_t T_start, uint64_t N_start, ReplayOut *out) {
  replay_one(k, level, 0, *count, pv, pdeg, posmap, palive, core, pioff, pinc,
             rdead, T_start, N_start, out);
}

// Many independent phases (of different graphs, or different levels of one
// graph: their vertex and clique sets are disjoint), one workgroup each.
__global__ __launch_bounds__(kReplayThreads) void k_replay_multi(
    uint32_t k, const ReplayPhase *__restrict__ phases,
    const uint32_t *__restrict__ pv, uint32_t *__restrict__ pdeg,
    const uint32_t *__restrict__ posmap, uint8_t *__restrict__ palive,
    const uint32_t *__restrict__ core, const uint64_t *__restrict__ pioff,
    const uint32_t *__restrict__ pinc, uint32_t *__restrict__ rdead,
    ReplayOut *out) {
  const ReplayPhase ph = phases[blockIdx.x];
  replay_one(k, ph.level, ph.base, ph.M, pv, pdeg, posmap, palive, core, pioff,
             pinc, rdead, ph.T_start, ph.N_start, out + blockIdx.x);
}

// Phase set-up for k_replay_multi: pdeg from the level-start degrees,
// palive = 1, then the phase cliques of every position are compacted.
__global__ void k_phase_init_multi(const ReplayPhase *__restrict__ phases,
                                   const uint32_t *__restrict__ pv,
                                   const uint32_t *__restrict__ sdeg,
                                   uint32_t *__restrict__ pdeg,
                                   uint8_t *__restrict__ palive) {
  const ReplayPhase ph = phases[blockIdx.x];
  for (uint32_t i = ph.base + threadIdx.x; i < ph.base + ph.M; i += blockDim.x) {
    pdeg[i] = sdeg[pv[i]];
    palive[i] = 1;
  }
}

__global__ __launch_bounds__(256) void k_phase_compact_multi(
    uint32_t k, const ReplayPhase *__restrict__ phases,
    const uint32_t *__restrict__ pv, const uint64_t *__restrict__ ioff,
    const uint32_t *__restrict__ inc, const uint32_t *__restrict__ kill,
    const uint64_t *__restrict__ pioff, uint32_t *__restrict__ pinc) {
  const ReplayPhase ph = phases[blockIdx.x];
  const uint32_t waves = blockDim.x / kWave;
  for (uint32_t i = ph.base + threadIdx.x / kWave; i < ph.base + ph.M; i += waves) {
    const uint32_t v = pv[i];
    uint64_t out = pioff[i];
    const uint64_t e = ioff[v + 1];
    for (uint64_t q0 = ioff[v]; q0 < e; q0 += kWave) {
      const uint64_t q = q0 + lane();
      const bool take = q < e && kill[inc[q * k]] == ph.level;
      unsigned cnt;
      const unsigned slot = wave_slot(take, &cnt);
      if (take)
        for (uint32_t t = 0; t < k; ++t) pinc[(out + slot) * k + t] = inc[q * k + t];
      out += cnt;
    }
  }
}

// Whole-GPU replay for phases with few vertices (M <= kReplayLdsCap) and
// many cliques per removal.  Every workgroup computes the same argmin from
// pdeg/palive, takes a contiguous slice of the removed vertex's incidence
// list, counts the decrements of the other phase members in LDS
// and publishes them as one row of `partial` (G x M); after a barrier one
// wave per phase vertex sums its column and updates pdeg.  Two grid
// barriers per removal; the removal order and the recorded densities are
// those of k_replay.
constexpr uint32_t kReplayLdsCap = kReplayGridMaxVertices;
constexpr uint64_t kReplayEntriesPerBlock = 256;

__global__ __launch_bounds__(kReplayThreads) void k_replay_grid(
    uint32_t k, uint32_t level, uint32_t M, const uint32_t *__restrict__ pv,
    uint32_t *pdeg, const uint32_t *__restrict__ posmap, uint8_t *palive,
    const uint32_t *__restrict__ core, const uint32_t *__restrict__ kill,
    const uint64_t *__restrict__ ioff, const uint32_t *__restrict__ inc,
    uint32_t *rdead, uint32_t *partial, GridBarrier *bar, uint64_t T_start,
    uint64_t N_start, ReplayOut *out, uint64_t kp, uint64_t nstar) {
  __shared__ uint32_t dec[kReplayLdsCap];
  __shared__ uint64_t s_wmin[kReplayThreads / kWave];
  __shared__ uint64_t s_key;
  const uint32_t tid = threadIdx.x, wid = tid / kWave;
  const uint32_t G = gridDim.x;
  const bool boss = blockIdx.x == 0 && tid == 0;
  for (uint32_t i = tid; i < M; i += kReplayThreads) dec[i] = 0;
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
    const uint32_t iv = posmap[v];
    if (boss) {
      R += kk >> 32;
      ++steps;
      if (j + 1 < M) {
        const uint64_t num = T_start - R, den = N_start - (j + 1);
        if (!found || frac_gt(num, den, best_num, best_den)) {
          best_num = num;
          best_den = den;
          found = 1;
        }
        // no later row can exceed kp: end after this removal
        if (nstar && (den <= nstar || num <= kp * nstar)) out->stopped = 1;
      }
    }
    // v's whole incidence list, one contiguous slice per workgroup; the
    // phase cliques are those killed at this level
    const uint64_t e0 = ioff[v], e1 = ioff[v + 1];
    const uint64_t per = (e1 - e0 + G - 1) / G;
    const uint64_t b0 = e0 + per * blockIdx.x;
    const uint64_t b1 = min(e1, b0 + per);
    for (uint64_t q = b0 + tid; q < b1; q += kReplayThreads) {
      const uint32_t *en = inc + q * k;
      const uint32_t c = en[0];
      if (kill[c] != level || rdead[c] == level + 1) continue;
      rdead[c] = level + 1;
      for (uint32_t t = 1; t < k; ++t) {
        const uint32_t u = en[t];
        if (core[u] == level) atomicAdd(&dec[posmap[u]], 1u);
      }
    }
    __syncthreads();
    uint32_t *row = partial + (size_t)blockIdx.x * M;
    for (uint32_t i = tid; i < M; i += kReplayThreads) {
      row[i] = dec[i];
      dec[i] = 0;
    }
    grid_barrier(bar);
    const uint32_t waves = G * (kReplayThreads / kWave);
    for (uint32_t i = blockIdx.x * (kReplayThreads / kWave) + wid; i < M; i += waves) {
      uint32_t sum = 0;
      for (uint32_t b = lane(); b < G; b += kWave) sum += partial[(size_t)b * M + i];
      sum = wave_sum(sum);
      if (lane() == 0 && sum) pdeg[i] -= sum;
    }
    if (boss) palive[iv] = 0;
    grid_barrier(bar);
    if (__hip_atomic_load(&out->stopped, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      break;  // the whole grid reads the same flag after the barrier
  }
  if (boss) {
    out->best_num = best_num;
    out->best_den = best_den;
    out->steps = steps;
    out->found = found;
  }
}

__global__ void k_zero_barrier(GridBarrier *b, unsigned long long limit) {
  unsigned *w = reinterpret_cast<unsigned *>(b);
  for (size_t i = threadIdx.x; i < sizeof(GridBarrier) / sizeof(unsigned); i += blockDim.x)
    w[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) b->limit = limit;
}

// ---- replay on adjacency bitsets (k = 2, 3; few vertices alive) ----
//
// Positions 0..N-1 are the N vertices alive at the level's start (core >=
// level: the phase vertices and those of higher levels); row i of `bits`
// holds position i's neighbours among them (W = ceil(N/64) words).  When
// phase vertex v goes, a phase vertex u adjacent to v loses the alive
// cliques holding both: 1 for k = 2, and for k = 3 the triangles {u, v, w}
// with w alive, popcount(N(u) & N(v) & alive).  One workgroup does every
// removal with LDS barriers only: no incidence list is read and no grid
// barrier is crossed.
constexpr uint32_t kReplayBitsWords = kReplayBitsMaxVertices / 64;

__global__ void k_collect_alive(uint32_t n, uint32_t level,
                                const uint32_t *__restrict__ core,
                                uint32_t *__restrict__ av,
                                uint32_t *__restrict__ posmap, uint32_t *count) {
  for (uint32_t base = blockIdx.x * blockDim.x; base < n;
       base += gridDim.x * blockDim.x) {
    const uint32_t v = base + threadIdx.x;
    const bool take = v < n && core[v] >= level;
    unsigned cnt;
    const unsigned slot = wave_slot(take, &cnt);
    uint32_t start = 0;
    if (lane() == 0 && cnt) start = atomicAdd(count, cnt);
    start = __shfl(start, 0, kWave);
    if (take) {
      av[start + slot] = v;
      posmap[v] = start + slot;
    }
  }
}

// Rows from the degree-ordered DAG: every arc v -> u between alive
// vertices sets bit u in v's row and bit v in u's row (hubs have short
// out-lists, so the dense top of the graph costs little).  One wave per
// position; positions at or beyond N (an alive count other than the host's)
// are left out, and the replay then refuses to run.
__global__ __launch_bounds__(256) void k_replay_rows(
    uint32_t N, uint32_t W, uint32_t level, const uint32_t *__restrict__ count,
    const uint32_t *__restrict__ av, const uint32_t *__restrict__ posmap,
    const uint32_t *__restrict__ core, const uint32_t *__restrict__ doff,
    const uint32_t *__restrict__ dadj, unsigned long long *__restrict__ bits) {
  const uint32_t rows = min(N, *count);
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < rows;
       i += waves) {
    const uint32_t v = av[i];
    for (uint32_t p = doff[v] + lane(); p < doff[v + 1]; p += kWave) {
      const uint32_t u = dadj[p];
      if (core[u] < level) continue;
      const uint32_t q = posmap[u];
      if (q >= rows) continue;
      atomicOr(&bits[(size_t)i * W + (q >> 6)], 1ull << (q & 63));
      atomicOr(&bits[(size_t)q * W + (i >> 6)], 1ull << (i & 63));
    }
  }
}

// Degree updates of one removal: list entry t (a phase position u adjacent
// to v) loses popcount(row u & nbr), summed over a group of G >= W lanes
// (one word each); kU entries per lane are loaded before any is summed.
template <uint32_t G, uint32_t WPL>
__device__ __forceinline__ void bits_dec(uint32_t nu, uint32_t W, const uint32_t *ul,
                                         const unsigned long long *__restrict__ bits,
                                         const unsigned long long *nbr, uint32_t *deg) {
  constexpr uint32_t kWaves = kReplayThreads / kWave;
  constexpr uint32_t kU = 8;
  constexpr uint32_t gpw = kWave / G, chunk = kU * gpw;
  const uint32_t wid = threadIdx.x / kWave, ln = lane();
  const uint32_t gl = ln % G, grp = ln / G;
  // this lane's words of a row: gl, gl + G, ... (WPL of them, those past
  // the row re-read its last word and are masked off)
  unsigned long long mine[WPL];
  uint32_t gw[WPL];
#pragma unroll
  for (uint32_t q = 0; q < WPL; ++q) {
    const uint32_t w = gl + q * G;
    mine[q] = w < W ? nbr[w] : 0ull;
    gw[q] = min(w, W - 1);
  }
  for (uint32_t base = wid * chunk; base < nu; base += kWaves * chunk) {
    // every load unconditional (entries past the list re-read the last one),
    // so all kU * WPL are in flight before the first is used
    unsigned long long x[kU][WPL];
#pragma unroll
    for (uint32_t r = 0; r < kU; ++r) {
      const unsigned long long *row =
          bits + (size_t)ul[min(base + r * gpw + grp, nu - 1)] * W;
#pragma unroll
      for (uint32_t q = 0; q < WPL; ++q) x[r][q] = row[gw[q]];
    }
#pragma unroll
    for (uint32_t r = 0; r < kU; ++r) {
      const uint32_t t = base + r * gpw + grp;
      uint32_t c = 0;
#pragma unroll
      for (uint32_t q = 0; q < WPL; ++q) c += (uint32_t)__popcll(x[r][q] & mine[q]);
      c = group_sum<G>(c);
      if (gl == 0 && t < nu) atomicSub(&deg[ul[t]], c);  // no return: no wait
    }
  }
}

// The replay itself: the removal order, recorded densities and early end of
// k_replay_grid, from the level-start degrees sdeg.  Per removal, three
// workgroup barriers: (1) each wave's smallest (degree, id) key with its
// position; (2) every wave reduces the keys, loads v's row and lists v's
// alive phase neighbours for its share of the words; (3) the degree updates.
// Wave reductions and scans go through DPP lane exchanges.
template <uint32_t K>
__global__ __launch_bounds__(kReplayThreads) void k_replay_bits(
    uint32_t level, uint32_t N, uint32_t M, const uint32_t *__restrict__ count,
    const uint32_t *__restrict__ av, const uint32_t *__restrict__ core,
    const uint32_t *__restrict__ sdeg, const unsigned long long *__restrict__ bits,
    uint64_t T_start, ReplayOut *out, uint64_t kp, uint64_t nstar, uint32_t *__restrict__ seq) {
  constexpr uint32_t kWaves = kReplayThreads / kWave;
  static_assert(kWaves == 16, "the key reduction spans one row of 16 lanes");
  __shared__ uint32_t deg[kReplayBitsMaxVertices];
  __shared__ uint32_t vid[kReplayBitsMaxVertices];
  __shared__ uint32_t ul[kReplayBitsMaxVertices];
  __shared__ unsigned long long alive[kReplayBitsWords], phase[kReplayBitsWords],
      nbr[kReplayBitsWords];
  __shared__ uint64_t s_wmin[kWaves];
  __shared__ uint32_t s_wpos[kWaves];
  __shared__ uint32_t s_stop;
  const uint32_t tid = threadIdx.x, wid = tid / kWave, ln = lane();
  if (*count != N || N > kReplayBitsMaxVertices) {
    if (tid == 0) {  // not the alive set the host sized: 0 steps, an error
      out->steps = 0;
      out->found = 0;
      out->stopped = 0;
    }
    return;
  }
  const uint32_t W = (N + 63) / 64;
  for (uint32_t i = tid; i < kReplayBitsMaxVertices; i += kReplayThreads) {
    ul[i] = 0;  // every entry a valid position, listed or not
    if (i < N) {
      const uint32_t v = av[i];
      vid[i] = v;
      deg[i] = core[v] == level ? sdeg[v] : 0;
    }
  }
  for (uint32_t b = wid * kWave; b < W * kWave; b += kReplayThreads) {
    const uint32_t i = b + ln;
    const bool in = i < N;
    const bool ph = in && core[av[i]] == level;
    const unsigned long long ma = __ballot(in), mp = __ballot(ph);
    if (ln == 0) {
      alive[b / kWave] = ma;
      phase[b / kWave] = mp;
    }
  }
  if (tid == 0) s_stop = 0;
  // lanes per row for the degree updates: groups of 8 (DPP-only sums) while
  // a lane's share of a row stays at most 3 words
  const uint32_t G = W <= 24 ? 8 : W <= 48 ? 16 : 32;
  uint64_t R = 0, best_num = 0, best_den = 1;
  uint32_t found = 0, steps = 0, stopped = 0;
  __syncthreads();
  for (uint32_t j = 0; j < M; ++j) {
    // (1) this wave's smallest key among the alive phase positions
    uint64_t key = ~0ull;
    uint32_t pos = 0;
    for (uint32_t w = wid; w < W; w += kWaves) {
      const unsigned long long live = alive[w] & phase[w];
      if ((live >> ln) & 1ull) {
        const uint32_t i = w * kWave + ln;
        const uint64_t kk = ((uint64_t)deg[i] << 32) | vid[i];
        if (kk < key) {
          key = kk;
          pos = i;
        }
      }
    }
    wave_argmin64(key, pos);
    if (ln == 0) {
      s_wmin[wid] = key;
      s_wpos[wid] = pos;
    }
    __syncthreads();
    // (2) every wave: the removed position iv and v's alive neighbours
    uint64_t mk = ln < kWaves ? s_wmin[ln] : ~0ull;
    uint32_t mp = ln < kWaves ? s_wpos[ln] : 0;
    row_argmin64(mk, mp);
    const uint64_t m = readlane64(mk, 0);
    const uint32_t iv = (uint32_t)__builtin_amdgcn_readlane((int)mp, 0);
    if (m == ~0ull) break;  // inconsistent counts; reported via steps
    unsigned long long a = 0, cm = 0;
    if (ln < W) {
      a = bits[(size_t)iv * W + ln] & alive[ln];
      if (ln == iv / kWave) a &= ~(1ull << (iv % kWave));
      cm = a & phase[ln];
    }
    const uint32_t c = (uint32_t)__popcll(cm);
    const uint32_t incl = wave_incl_scan(c);
    const uint32_t excl = incl - c;
    const uint32_t nu = min((uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1), N);
    for (uint32_t w = wid; w < W; w += kWaves) {
      const unsigned long long cw = readlane64(cm, w), aw = readlane64(a, w);
      const uint32_t ew = (uint32_t)__builtin_amdgcn_readlane((int)excl, (int)w);
      if ((cw >> ln) & 1ull) {
        const uint32_t at = ew + (uint32_t)__popcll(cw & ((1ull << ln) - 1));
        if (at < kReplayBitsMaxVertices) ul[at] = w * kWave + ln;
      }
      if (ln == 0) nbr[w] = aw;
    }
    if (tid == 0) {
      alive[iv / kWave] &= ~(1ull << (iv % kWave));
      R += m >> 32;
      if (seq) {  // the removal sequence: vertex, then its clique count
        seq[2 * j] = (uint32_t)m;
        seq[2 * j + 1] = (uint32_t)(m >> 32);
      }
      ++steps;
      if (j + 1 < M) {
        const uint64_t num = T_start - R, den = (uint64_t)N - (j + 1);
        if (!found || frac_gt(num, den, best_num, best_den)) {
          best_num = num;
          best_den = den;
          found = 1;
        }
        // no later row can exceed kp: end after this removal
        if (nstar && (den <= nstar || num <= kp * nstar)) s_stop = stopped = 1;
      }
    }
    __syncthreads();
    // (3) degree updates of the listed neighbours
    if (K == 2) {
      for (uint32_t t = tid; t < nu; t += kReplayThreads) deg[ul[t]] -= 1;
    } else if (G == 8) {
      bits_dec<8, 3>(nu, W, ul, bits, nbr, deg);
    } else if (G == 16) {
      bits_dec<16, 3>(nu, W, ul, bits, nbr, deg);
    } else {
      bits_dec<32, 2>(nu, W, ul, bits, nbr, deg);
    }
    __syncthreads();
    if (s_stop) break;
  }
  if (tid == 0) {
    out->best_num = best_num;
    out->best_den = best_den;
    out->steps = steps;
    out->found = found;
    out->stopped = stopped;
  }
}

}  // namespace

hipError_t build_incidence(uint64_t T, uint32_t k, const uint32_t *cliques,
                           const uint64_t *ioff, uint32_t *icur,
                           uint32_t *inc, hipStream_t s) {
  if (k > 4) return hipErrorInvalidValue;
  if (k < 2) return hipErrorInvalidValue;
  const uint32_t small = (unsigned __int128)T * k < ((unsigned __int128)1 << 32) ? 1u : 0u;
  hipLaunchKernelGGL(k_build_inc, dim3(grid_for(T, kIncPairs / k, 16384)),
                     dim3(kIncThreads), 0, s, T, k, cliques, ioff, icur, inc, small);
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

hipError_t refine_prepare(uint32_t n, uint64_t T, uint32_t level,
                          uint32_t *core, uint32_t *deg, const uint32_t *sdeg,
                          uint32_t *kill, uint32_t *alive, uint32_t *count,
                          hipStream_t s) {
  if (level >= kParked) return hipErrorInvalidValue;
  hipError_t e = hipMemsetAsync(count, 0, 4, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_refine_vertices, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s,
                     n, level, core, deg, sdeg, alive, count);
  if (kill)
    hipLaunchKernelGGL(k_refine_cliques, dim3(grid_for(T, 256, 16384)), dim3(256), 0, s,
                       T, level, kill);
  return hipGetLastError();
}

hipError_t refine_restore(uint32_t n, uint32_t *core, const uint32_t *deg,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_refine_restore, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s,
                     n, core, deg);
  return hipGetLastError();
}

hipError_t peel_persistent(uint32_t n, uint32_t k, uint64_t T, const uint64_t *ioff,
                           const uint32_t *inc, const uint32_t *cliques,
                           uint32_t *deg, uint32_t *core, uint32_t *kill,
                           uint32_t *stamp, uint32_t sbase, uint32_t *scratch32,
                           uint64_t *scratch64,
                           PeelCtl *ctl, uint32_t *lv_level, uint32_t *lv_min,
                           unsigned long long *lv_removed,
                           unsigned long long *lv_killed, uint32_t *sdeg,
                           const uint32_t *alive0, uint32_t na0, uint32_t cshift,
                           uint32_t max_level, unsigned *grid, hipStream_t s) {
  if (n >= (1u << 24) || *grid < 1 || cshift > 31) return hipErrorInvalidValue;
  if (max_level >= kParked) max_level = kParked - 1;
  const uint32_t na = alive0 ? na0 : n;
  hipError_t e;
  if ((e = hipMemsetAsync(lv_level, 0, (size_t)n * 4, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(lv_min, 0, (size_t)n * 4, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(lv_removed, 0, (size_t)n * 8, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(lv_killed, 0, (size_t)n * 8, s)) != hipSuccess) return e;
  uint32_t *A0 = scratch32, *A1 = A0 + n, *F0 = A1 + n, *F1 = F0 + n,
           *F2 = F1 + n, *pos = F2 + n;
  uint64_t *FC0 = scratch64, *FC1 = FC0 + n, *FC2 = FC1 + n;
  if (alive0) {
    if (na && (e = hipMemcpyAsync(A0, alive0, (size_t)na * 4, hipMemcpyDeviceToDevice,
                                  s)) != hipSuccess)
      return e;
  } else {
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, A0);
  }
  hipLaunchKernelGGL(k_peel_ctl_init, dim3(1), dim3(1), 0, s, ctl, na,
                     barrier_limit_ticks());
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (sbase < 1 || sbase > kParked / 2) return hipErrorInvalidValue;
  // LDS hash aggregation of decrements (CDS_PEEL_HASH=0 turns it off)
  static const uint32_t hagg = [] {
    const char *v = getenv("CDS_PEEL_HASH");
    return (uint32_t)(v ? atoi(v) != 0 : 1);
  }();
  hipLaunchKernelGGL(k_stamp_reset, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n,
                     core, stamp, (uint32_t)(sbase == 1));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  void *args[] = {&n,  &k,  (void *)&ioff, (void *)&inc, (void *)&cliques,
                  &deg, &core, &stamp, &sbase, &A0, &A1, &F0, &F1, &F2, &FC0, &FC1,
                  &FC2, &ctl, &lv_level, &lv_removed, &lv_killed, &pos,
                  &sdeg, &lv_min, &cshift, &max_level, (void *)&hagg};
  return launch_grid_shrinking(reinterpret_cast<const void *>(k_peel_persistent), grid,
                               kPeelThreads, args, 0, s);
}

hipError_t peel_derive_kill(uint64_t T, uint32_t k, const uint32_t *cliques,
                            const uint32_t *core, uint32_t *kill, hipStream_t s) {
  if (!T) return hipSuccess;
  hipLaunchKernelGGL(k_kill_from_core, dim3(grid_for(T, 256, 16384)), dim3(256), 0, s,
                     T, k, cliques, core, kill);
  return hipGetLastError();
}

__global__ void k_posmap_local(uint32_t M, const uint32_t *__restrict__ pv,
                               uint32_t *__restrict__ posmap) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < M;
       i += gridDim.x * blockDim.x)
    posmap[pv[i]] = i;
}

hipError_t replay_phase_list(uint32_t k, uint32_t level, const uint32_t *pv, uint32_t M,
                             const uint32_t *core, const uint32_t *kill,
                             const uint64_t *ioff, const uint32_t *inc,
                             uint64_t T_start, uint64_t N_start, uint32_t *pdeg,
                             uint32_t *posmap, uint8_t *palive, uint32_t *rdead,
                             uint32_t *pcount, ReplayOut *out, uint32_t *partial,
                             GridBarrier *bar, unsigned *grid, const uint32_t *sdeg,
                             uint64_t kp, uint64_t nstar, hipStream_t s) {
  if (*grid < 1 || M == 0 || M > kReplayLdsCap) return hipErrorInvalidValue;
  hipError_t e = hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(pcount), M, 1, s);
  if (e != hipSuccess) return e;
  if ((e = hipMemsetAsync(out, 0, sizeof(ReplayOut), s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_posmap_local, dim3(grid_for(M, 256, 4096)), dim3(256), 0, s, M, pv,
                     posmap);
  hipLaunchKernelGGL(k_phase_deg_snap, dim3(grid_for(M, 256, 4096)), dim3(256), 0, s,
                     pcount, pv, sdeg, pdeg, palive);
  hipLaunchKernelGGL(k_zero_barrier, dim3(1), dim3(256), 0, s, bar, barrier_limit_ticks());
  if ((e = hipGetLastError()) != hipSuccess) return e;
  uint32_t *pvm = const_cast<uint32_t *>(pv);  // the kernel reads it only
  void *args[] = {&k, &level, &M, &pvm, &pdeg, (void *)&posmap, &palive,
                  (void *)&core, (void *)&kill, (void *)&ioff, (void *)&inc,
                  &rdead, &partial, &bar, &T_start, &N_start, &out, &kp, &nstar};
  return launch_grid_shrinking(reinterpret_cast<const void *>(k_replay_grid), grid,
                               kReplayThreads, args, 0, s);
}

hipError_t replay_phase_bits(uint32_t k, uint32_t level, uint32_t n, const uint32_t *doff,
                             const uint32_t *dadj, const uint32_t *core, const uint32_t *sdeg,
                             uint64_t T_start, uint32_t N, uint32_t M, uint32_t *av,
                             uint32_t *posmap, uint32_t *pcount, unsigned long long *bits,
                             ReplayOut *out, hipEvent_t prepared, uint64_t kp, uint64_t nstar,
                             hipStream_t s, uint32_t *seq) {
  if ((k != 2 && k != 3) || N == 0 || N > kReplayBitsMaxVertices || M == 0 || M > N)
    return hipErrorInvalidValue;
  const uint32_t W = (N + 63) / 64;
  hipError_t e = hipMemsetAsync(pcount, 0, sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  if ((e = hipMemsetAsync(out, 0, sizeof(ReplayOut), s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(bits, 0, (size_t)N * W * 8, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_collect_alive, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n,
                     level, core, av, posmap, pcount);
  hipLaunchKernelGGL(k_replay_rows, dim3(grid_for((size_t)N * kWave, 256, 8192)), dim3(256),
                     0, s, N, W, level, pcount, av, posmap, core, doff, dadj, bits);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (prepared) (void)hipEventRecord(prepared, s);
  if (k == 2)
    hipLaunchKernelGGL(k_replay_bits<2>, dim3(1), dim3(kReplayThreads), 0, s, level, N, M,
                       pcount, av, core, sdeg, bits, T_start, out, kp, nstar, seq);
  else
    hipLaunchKernelGGL(k_replay_bits<3>, dim3(1), dim3(kReplayThreads), 0, s, level, N, M,
                       pcount, av, core, sdeg, bits, T_start, out, kp, nstar, seq);
  return hipGetLastError();
}

hipError_t replay_phase(uint32_t k, uint32_t level, uint32_t n,
                        const uint32_t *core, const uint32_t *kill,
                        const uint64_t *ioff, const uint32_t *inc,
                        uint64_t T_start, uint64_t N_start, uint32_t M,
                        uint32_t *pv, uint32_t *pdeg, uint32_t *posmap,
                        uint8_t *palive, uint32_t *rdead, uint32_t *pcount,
                        uint64_t *pioff, uint32_t *pinc, ReplayOut *out,
                        uint32_t *partial, GridBarrier *bar, unsigned *grid,
                        const uint32_t *sdeg, DeviceArena &arena,
                        hipEvent_t prepared, uint64_t kp, uint64_t nstar,
                        hipStream_t s) {
  if (*grid && M > kReplayLdsCap) return hipErrorInvalidValue;
  hipError_t e = hipMemsetAsync(pcount, 0, sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  if ((e = hipMemsetAsync(out, 0, sizeof(ReplayOut), s)) != hipSuccess) return e;
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_collect_level, dim3(grid_for(n, 256, 8192)), dim3(256),
                     0, s, n, level, core, pv, posmap, pcount);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const unsigned g = grid_for((size_t)M * kWave, 256, 8192);
  if (*grid && sdeg)
    hipLaunchKernelGGL(k_phase_deg_snap, dim3(grid_for(M, 256, 4096)), dim3(256),
                       0, s, pcount, pv, sdeg, pdeg, palive);
  else
    hipLaunchKernelGGL(k_phase_deg, dim3(g), dim3(256), 0, s, k, level, pcount,
                       pv, ioff, inc, kill, pdeg, palive);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (*grid) {
    hipLaunchKernelGGL(k_zero_barrier, dim3(1), dim3(256), 0, s, bar, barrier_limit_ticks());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (prepared) (void)hipEventRecord(prepared, s);
    void *args[] = {&k, &level, &M, &pv, &pdeg, (void *)&posmap, &palive,
                    (void *)&core, (void *)&kill, (void *)&ioff, (void *)&inc,
                    &rdead, &partial, &bar, &T_start, &N_start, &out, &kp, &nstar};
    return launch_grid_shrinking(reinterpret_cast<const void *>(k_replay_grid), grid,
                                 kReplayThreads, args, 0, s);
  }
  if ((e = scan_u32(pdeg, pioff, M, arena)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_phase_compact, dim3(g), dim3(256), 0, s, k, level,
                     pcount, pv, ioff, inc, kill, pioff, pinc);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (prepared) (void)hipEventRecord(prepared, s);
  hipLaunchKernelGGL(k_replay, dim3(1), dim3(kReplayThreads), 0, s, k, level,
                     pcount, pv, pdeg, posmap, palive, core, pioff, pinc,
                     rdead, T_start, N_start, out);
  return hipGetLastError();
}

hipError_t replay_phases_multi(uint32_t k, uint32_t nphases,
                               const ReplayPhase *phases, uint32_t npos,
                               const uint32_t *pv, const uint32_t *sdeg,
                               uint32_t *pdeg, const uint32_t *posmap,
                               uint8_t *palive, const uint32_t *core,
                               const uint32_t *kill, const uint64_t *ioff,
                               const uint32_t *inc, uint64_t *pioff,
                               uint32_t *pinc, uint32_t *rdead, ReplayOut *out,
                               DeviceArena &arena, hipStream_t s) {
  if (nphases == 0) return hipSuccess;
  hipError_t e = hipMemsetAsync(pdeg, 0, (size_t)npos * sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_phase_init_multi, dim3(nphases), dim3(256), 0, s, phases,
                     pv, sdeg, pdeg, palive);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = scan_u32(pdeg, pioff, npos, arena)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_phase_compact_multi, dim3(nphases), dim3(256), 0, s, k,
                     phases, pv, ioff, inc, kill, pioff, pinc);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_replay_multi, dim3(nphases), dim3(kReplayThreads), 0, s, k,
                     phases, pv, pdeg, posmap, palive, core, pioff, pinc, rdead,
                     out);
  return hipGetLastError();
}

// Co-resident workgroups of a cooperative kernel: 0 if it cannot run.
static unsigned coop_cus(const void *fn, int threads, int device) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) !=
          hipSuccess ||
      per_cu < 1)
    return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) !=
      hipSuccess)
    return 0;
  return (unsigned)cus;
}

unsigned peel_grid_blocks(uint32_t n, uint64_t T, uint32_t k, int device,
                          unsigned force) {
  const unsigned cus = coop_cus(reinterpret_cast<const void *>(k_peel_persistent),
                                kPeelThreads, device);
  if (cus == 0) return 0;
  // 1024-thread workgroups, up to one per CU.  Every grid barrier costs a
  // release and an acquire fence per workgroup, but once claims and
  // decrements are aggregated the sub-rounds' chunk processing dominates
  // and scales with the CUs (R-MAT scale 20, k = 3: 60 ms with 128
  // workgroups, 42 with 192, 34 with 256; barriers 1.1 -> 1.6 ms).
  // Smaller graphs get one workgroup per 32768 vertices or 2^22 incidence
  // entries, so most run on one workgroup as a plain launch.
  const uint64_t cap = cus;
  uint64_t want = std::max<uint64_t>((n + 32767) / 32768, (k * T + (1u << 22) - 1) >> 22);
  want = std::min(std::max<uint64_t>(want, 1), cap);
  // an explicit width (tuning, CDS_PEEL_BLOCKS) may take every CU: one
  // 1024-thread workgroup per CU is co-resident
  if (force >= 1) want = std::min<uint64_t>(force, cus);
  if (const char *gb = getenv("CDS_PEEL_BLOCKS")) {  // diagnostics: wins
    const long v = atol(gb);
    if (v >= 1 && (uint64_t)v <= cus) want = (uint64_t)v;
  }
  return (unsigned)want;
}

uint64_t replay_nstar(uint32_t k, uint64_t kp) {
  if (k < 2) return 0;
  const double lim = (double)k * (double)kp;
  // C(N-1, k-1) for N = k, k+1, ...: C(m, r) grows as m increases
  double c = 1;  // C(k-1, k-1)
  for (uint64_t N = k; N < (1ull << 32); ++N) {
    if (c > lim) return N;
    const double m = (double)(N - 1);  // C(m+1, r) = C(m, r) (m+1) / (m+1-r)
    c = c * (m + 1) / (m + 1 - (double)(k - 1));
  }
  return 0;
}

unsigned replay_grid_blocks(int device, uint64_t entries_per_step, unsigned force) {
  const unsigned cus = coop_cus(reinterpret_cast<const void *>(k_replay_grid),
                                kReplayThreads, device);
  const unsigned cap = cus > 1 ? cus / 2 : cus;
  if (cap == 0) return 0;
  // Measured on R-MAT scale 20 k=3 and scale 17 k=4: the replay time falls
  // with every doubling of workgroups up to the cap (the removals' slices
  // and the column sums are throughput-bound), so the whole cap is used
  // unless a removal touches fewer than kReplayEntriesPerBlock entries per
  // workgroup.
  uint64_t want = (entries_per_step + kReplayEntriesPerBlock - 1) / kReplayEntriesPerBlock;
  if (force >= 1) want = force;
  if (const char *e = getenv("CDS_REPLAY_BLOCKS")) want = strtoull(e, nullptr, 10);
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(want, cap));
}

}  // namespace cds
