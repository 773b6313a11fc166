// (c) Exact max flow of one component's vertex/clique network.
//
// Reference GPU path: createFlowNetwork / preFlow / pushRelabel /
// globalRelabel (cudaCode/src/helpers.cu:1173-1492), double capacities and
// a cooperative grid.  Oracle: createFlownetwork + edmondsKarp
// (cpuCode/c++/src/helpers.cpp:1302-1405), float capacities.
//
// Network (oracle orientation): s->v cap cd_v, v->t cap fl(alpha*k),
// v->clique cap 1, clique->v cap k-1.  Every capacity is multiplied by
// U = 2^sc so fl(alpha*k) is an integer; flows are int64 and exact.
//
// The kernel solves the REVERSED network (source t, sink s) because the
// oracle's answer set is the vertices reachable from s in the residual of
// a maximum flow, i.e. the minimal min-cut side, and in the reversed
// network that set is exactly "can reach the sink" after a maximum
// PREflow, which a backward BFS from the sink (the global relabel)
// computes.  The arcs t->v start saturated (excess fl(alpha*k)*U at every
// vertex) and are never used again.
//
// Representation is implicit: per (clique, member slot) one int64 x = net
// flow vertex->clique in the reversed network, so residual(v->c) =
// (k-1)U - x and residual(c->v) = U + x; per vertex rvs = residual(v->sink).
// A dense core of a scale-20 R-MAT graph has ~1e3 vertices and ~1e8
// cliques, so vertex nodes have incidence lists of 1e5..1e7 arcs and
// clique nodes have k; small components run the same kernel on one
// workgroup.
//
// One cooperative launch.  A round is four phases separated by grid
// barriers:
//   1. clique nodes (one thread each) push excess to lower members and
//      relabel if excess remains;
//   2. vertex nodes push to the sink, and every still-active vertex appends
//      its incidence list, cut into chunks, to the round's work list;
//   3. each wave takes chunks: an admissible arc (lower clique) takes
//      min(residual, what is left of the vertex's excess) through one
//      64-bit atomic on the vertex's remainder, so a vertex discharges over
//      all its admissible arcs in one round; non-admissible residual arcs
//      feed an atomic min for the relabel;
//   4. vertices keep the remainder as excess and relabel to min + 1 if any
//      is left.
// Pushes only go from a higher to a lower node and relabels only raise
// heights to one above the lowest residual neighbour, so labels stay valid.
// A global relabel (BFS from the sink over reversed residual arcs,
// alternating vertex and clique levels) runs every relabel_rounds() rounds and
// whenever no node is active; the solve ends only when that BFS finds no
// active node, which certifies the preflow maximum.
#include <cstring>

#include "dev.h"
#include "kernels.h"

namespace cds {
namespace {

constexpr uint32_t kGridChunk = kFlowChunkArcs;
constexpr uint32_t kGridLdsVertices = 4096;  // LDS excess accumulator, 32 KB
// Rounds between global relabels: a relabel reads the whole network, so
// large networks relabel rarely (measured on R-MAT cores of 2.3e8 / 7.2e8
// slots: 24 -> 64 rounds cut the flow from 73 / 220 ms to 56 / 135 ms) while
// small ones, whose rounds cost barriers rather than bandwidth, keep
// relabelling often (64 slowed the 256-graph batched sample by 15 %).
__device__ __forceinline__ uint32_t relabel_rounds(uint64_t slots) {
  return slots >= (1ull << 22) ? 64u : 24u;
}
// Current-arc windows of the vertex discharge, in chunks.  After a global
// relabel a vertex scans its whole list; afterwards the window starts at
// its current arc and is sized from the chunks the last discharge used
// (twice that, at least kArcWindow0), doubling while the excess outlasts
// it; after a relabel it restarts at chunk 0 with kArcWindowRelabel.
constexpr uint32_t kArcWindow0 = 4;
constexpr uint32_t kArcWindowRelabel = 64;
constexpr uint32_t kArcWindowMax = 1u << 24;  // more than any list has
constexpr uint64_t kArcFull = (uint64_t)kArcWindowMax << 32;  // chunk 0, whole list
// After a global relabel in mid-solve the excess left is small: windows
// restart at chunk 0 with 256 chunks instead of the whole list.
constexpr uint64_t kArcAfterRelabel = (uint64_t)256 << 32;
constexpr int kChunkSteps = kGridChunk / kWave;  // wave passes per chunk
constexpr int kMaxFlowWaves = 16;                // 1024-thread workgroups

// List appends are aggregated so a counter that millions of nodes append
// to sees few atomics.  wave_append: the items one wave gathered over up to
// N passes (one per lane and pass), one atomic.  All lanes call it.
template <int N>
__device__ __forceinline__ void wave_append(uint32_t *n, uint32_t *out,
                                            const bool (&f)[N],
                                            const uint32_t (&v)[N]) {
  uint32_t mine[N], total = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint64_t b = __ballot(f[i]);
    mine[i] = total + __popcll(b & lanes_below());
    total += __popcll(b);
  }
  if (total == 0) return;
  uint32_t base = 0;
  if (lane() == 0) base = atomicAdd(n, total);
  base = __shfl(base, 0, kWave);
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (f[i]) out[base + mine[i]] = v[i];
}

// block_append: the same over a whole workgroup, one atomic per call;
// app = kMaxFlowWaves + 1 words of LDS.  Every thread of the workgroup
// calls it (it synchronises the workgroup).
template <int N>
__device__ __forceinline__ void block_append(uint32_t *app, uint32_t *n,
                                             uint32_t *out, const bool (&f)[N],
                                             const uint32_t (&v)[N]) {
  uint32_t mine[N], total = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint64_t b = __ballot(f[i]);
    mine[i] = total + __popcll(b & lanes_below());
    total += __popcll(b);
  }
  const uint32_t w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  if (lane() == 0) app[w] = total;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (uint32_t i = 0; i < nw; ++i) {
      const uint32_t t = app[i];
      app[i] = run;
      run += t;
    }
    app[kMaxFlowWaves] = run ? atomicAdd(n, run) : 0;
  }
  __syncthreads();
  const uint32_t base = app[kMaxFlowWaves] + app[w];
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (f[i]) out[base + mine[i]] = v[i];
  __syncthreads();  // app is reused by the next call
}
constexpr uint64_t kMask40 = (1ull << 40) - 1;

template <class T> __device__ __forceinline__ T ald(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Adds d to a clique's excess; returns the value before.
__device__ __forceinline__ long long exc_add(long long *p, long long d) {
  return (long long)atomicAdd((unsigned long long *)p, (unsigned long long)d);
}
__device__ __forceinline__ long long exc_add(int *p, long long d) {
  return (long long)(int)atomicAdd((unsigned int *)p, (unsigned int)(int)d);
}

// XT: the per-(clique, member) residual x and the clique excess (int64, or
// int32 for networks whose capacities fit 31 bits).  kByClique: the
// vertex incidence holds clique ids (from cbase) instead of clique-member
// slots (from sbase), for networks with 2^32 slots or more; the member's
// slot is found among the clique's k members.
template <class XT, bool kByClique_>
struct G {
  static constexpr bool kByClique = kByClique_;
  using X = XT;
  uint32_t k, V, T;
  uint32_t kmag, ksh;  // div_k's magic and shift for k
  uint32_t ms;         // words per clique in mem
  int INF;
  uint32_t vbase;
  uint64_t sbase, cbase;
  long long U;
  const uint32_t *mem;
  const uint64_t *ioff;
  const uint32_t *inc;
  XT *x, *exc;
  long long *rvs, *exv, *rem;
  int *hv, *hc, *hmin;
  // Double buffers are separate members selected with ?: (an array member
  // indexed at run time would live in scratch memory).
  uint32_t *list0, *list1;      // vertex work lists (active vertices / BFS)
  uint64_t *lchunk0, *lchunk1;  // their chunk starts
  uint32_t *cq0, *cq1;          // clique BFS frontiers
  uint64_t *arc;                // per vertex: current chunk | window << 32
  uint32_t *astop;              // per vertex: first unsaturated window chunk
  uint32_t *app;                // LDS: block_append's wave offsets
  GridFlowCtl *ctl;
  uint32_t gtid, gthreads, gwave, gwaves;
  bool leader;
  uint32_t nblocks, bid;  // workgroups on this problem, this one's index
};

// s / k for 2 <= k <= 6 as a multiply-high and a shift: magic m and shift
// sh with floor(s * m / 2^(32 + sh)) = floor(s / k) for every 32-bit s
// (checked exhaustively for k = 2..6).
__device__ __forceinline__ void div_magic(uint32_t k, uint32_t *m, uint32_t *sh) {
  *m = k == 2 ? 0x80000000u : k == 4 ? 0x40000000u : k == 5 ? 0xCCCCCCCDu : 0xAAAAAAABu;
  *sh = k == 3 ? 1u : (k == 5 || k == 6) ? 2u : 0u;
}
template <class Gx>
__device__ __forceinline__ uint32_t div_k(const Gx &g, uint32_t s) {
  return __umulhi(s, g.kmag) >> g.ksh;
}

// Residual word of member j of clique c, the clique's height and excess.
template <class Gx>
__device__ __forceinline__ auto &xcj(const Gx &g, uint64_t c, uint32_t j) {
  return g.x[c * g.k + j];
}
template <class Gx>
__device__ __forceinline__ int &hcl(const Gx &g, uint64_t c) {
  return g.hc[c];
}
template <class Gx>
__device__ __forceinline__ auto &excl(const Gx &g, uint64_t c) {
  return g.exc[c];
}
// Residual word by its index a (as arc_at returns it).
template <class Gx>
__device__ __forceinline__ auto &xa(const Gx &g, uint64_t a) {
  return g.x[a];
}

// Barrier over the workgroups working on one problem.
template <class Gx>
__device__ __forceinline__ void gsync(const Gx &g) {
  group_barrier(&g.ctl->bar, g.nblocks, g.bid);
}

template <class Gx>
__device__ __forceinline__ uint32_t *glist(const Gx &g, int b) {
  return b ? g.list1 : g.list0;
}
template <class Gx>
__device__ __forceinline__ uint64_t *glchunk(const Gx &g, int b) {
  return b ? g.lchunk1 : g.lchunk0;
}
template <class Gx>
__device__ __forceinline__ uint32_t *gcq(const Gx &g, int b) {
  return b ? g.cq1 : g.cq0;
}

template <class Gx>
__device__ __forceinline__ uint64_t nchunks(const Gx &g, uint32_t v) {
  return (g.ioff[v + 1] - g.ioff[v] + kGridChunk - 1) / kGridChunk;
}

template <class Gx>
__device__ __forceinline__ void list_push(const Gx &g, int b, uint32_t v) {
  const unsigned long long old =
      atomicAdd(&g.ctl->lpack[b], (1ull << 40) + (unsigned long long)nchunks(g, v));
  glist(g, b)[old >> 40] = v;
  glchunk(g, b)[old >> 40] = old & kMask40;
}

// Appends v with `chunks` work chunks to vertex list b.
template <class Gx>
__device__ __forceinline__ void list_push_n(const Gx &g, int b, uint32_t v,
                                            uint32_t chunks) {
  const unsigned long long old =
      atomicAdd(&g.ctl->lpack[b], (1ull << 40) + (unsigned long long)chunks);
  glist(g, b)[old >> 40] = v;
  glchunk(g, b)[old >> 40] = old & kMask40;
}

// First chunk and chunk count a vertex with excess scans this round: its
// current-arc window, or the whole list once the current arc has passed
// the end (a full scan, after which it relabels).
template <class Gx>
__device__ __forceinline__ void discharge_range(const Gx &g, uint32_t v,
                                                uint32_t *first, uint32_t *count) {
  const uint32_t nch = (uint32_t)nchunks(g, v);
  const uint64_t a = g.arc[v];
  const uint32_t cur = (uint32_t)a, win = (uint32_t)(a >> 32);
  if (cur >= nch) {
    *first = 0;
    *count = nch;
  } else {
    *first = cur;
    *count = min(win, nch - cur);
  }
}

// Maps work chunk ci of list b (count entries) to (vertex, incidence range);
// `windowed`: the vertex's chunks start at its discharge range.  Called by
// whole waves with a wave-uniform ci.
template <class Gx>
__device__ __forceinline__ void chunk_of(const Gx &g, int b, uint32_t count,
                                         uint64_t ci, uint32_t *v, uint64_t *lo,
                                         uint64_t *hi, bool windowed = false,
                                         uint32_t *vchunk = nullptr) {
  const uint32_t a = wave_last_le(glchunk(g, b), count, ci);
  *v = glist(g, b)[a];
  uint32_t first = 0, n_;
  if (windowed) discharge_range(g, *v, &first, &n_);
  const uint32_t j = first + (uint32_t)(ci - glchunk(g, b)[a]);
  if (vchunk) *vchunk = j;
  *lo = g.ioff[*v] + (uint64_t)j * kGridChunk;
  const uint64_t e = g.ioff[*v + 1];
  *hi = *lo + kGridChunk < e ? *lo + kGridChunk : e;
}

// Clique of a residual word index below 2^32 (slot incidence).
template <class Gx>
__device__ __forceinline__ uint32_t clique_of_arc(const Gx &g, uint32_t a) {
  static_assert(!Gx::kByClique, "slot incidence only");
  return div_k(g, a);
}

// Incidence entry p of vertex v: the index of its residual word (for xa)
// and its clique (local ids).
template <class Gx>
__device__ __forceinline__ uint64_t arc_at(const Gx &g, uint64_t p, uint32_t v,
                                           uint32_t *c) {
  if constexpr (Gx::kByClique) {
    *c = (uint32_t)((uint64_t)g.inc[p] - g.cbase);
    const uint64_t mbase = (uint64_t)*c * g.ms;
    const uint32_t gv = v + g.vbase;
    uint32_t j = 0;
    while (j + 1 < g.k && g.mem[mbase + j] != gv) ++j;  // v is one of the k
    return (uint64_t)*c * g.k + j;
  } else {
    const uint32_t s = (uint32_t)((uint64_t)g.inc[p] - g.sbase);
    *c = div_k(g, s);
    return s;
  }
}

// Exact BFS distances to the sink over reversed residual arcs; returns the
// number of active nodes (excess > 0 and finite height).  Uses list buffer
// 0/1 for vertex frontiers and cq[0]/cq[1] for clique frontiers.
template <class Gx>
__device__ __forceinline__ unsigned long long global_relabel(const Gx &g, uint64_t arc_reset) {
  GridFlowCtl *ctl = g.ctl;
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) g.hv[v] = g.INF;
  for (uint32_t c = g.gtid; c < g.T; c += g.gthreads) hcl(g, c) = g.INF;
  if (g.leader) {
    ctl->lpack[0] = 0;
    ctl->lpack[1] = 0;
    ctl->cqn[0] = 0;
    ctl->cqn[1] = 0;
    ctl->active = 0;
  }
  gsync(g);
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads)
    if (g.rvs[v] > 0) {
      g.hv[v] = 1;
      list_push(g, 0, v);
    }
  gsync(g);
  int vb = 0, cb = 0, d = 1;
  for (;;) {
    // vertices at distance d -> cliques at d+1 (arc clique->v residual U+x)
    const unsigned long long pack = ald(&ctl->lpack[vb]);
    const uint32_t nv = (uint32_t)(pack >> 40);
    if (nv == 0) break;
    const uint64_t nch = pack & kMask40;
    if (g.leader) {
      ctl->cqn[cb] = 0;
      ctl->lpack[1 - vb] = 0;
    }
    gsync(g);
    // Direction choice: when the frontier's incidence is large against the
    // clique count, every unlabelled clique checks its own members instead
    // (a sequential pass over the cliques, no atomics); otherwise the
    // frontier vertices scan their incidence.
    if (nch * kGridChunk >= (uint64_t)g.T / 4) {
      constexpr int kPull = 1;
      for (uint64_t t0 = (uint64_t)g.bid * blockDim.x * kPull; t0 < g.T;
           t0 += (uint64_t)g.nblocks * blockDim.x * kPull) {
        bool f[kPull];
        uint32_t cs[kPull];
#pragma unroll
        for (int it = 0; it < kPull; ++it) {
          const uint64_t c = t0 + (uint64_t)it * blockDim.x + threadIdx.x;
          f[it] = false;
          cs[it] = (uint32_t)c;
          if (c >= g.T || hcl(g, c) != g.INF) continue;
          const uint64_t base = c * g.ms;
          for (uint32_t j = 0; j < g.k; ++j)
            if (g.hv[g.mem[base + j] - g.vbase] == d && g.U + xcj(g, c, j) > 0) {
              f[it] = true;
              break;
            }
          if (f[it]) hcl(g, c) = d + 1;
        }
        block_append(g.app, &ctl->cqn[cb], gcq(g, cb), f, cs);
      }
    } else
    for (uint64_t ci = g.gwave; ci < nch; ci += g.gwaves) {
      uint32_t v;
      uint64_t lo, hi;
      chunk_of(g, vb, nv, ci, &v, &lo, &hi);
      bool f[kChunkSteps];
      uint32_t cs[kChunkSteps];
#pragma unroll
      for (int it = 0; it < kChunkSteps; ++it) {
        const uint64_t p = lo + (uint64_t)it * kWave + lane();
        f[it] = false;
        cs[it] = 0;
        if (p >= hi) continue;
        uint32_t c;
        const uint64_t a = arc_at(g, p, v, &c);
        // a plain load filters: a stale INF only costs a failed CAS; most
        // cliques met after the first level are labelled already
        if (hcl(g, c) != g.INF) continue;
        if (g.U + xa(g, a) <= 0) continue;
        f[it] = atomicCAS(&hcl(g, c), g.INF, d + 1) == g.INF;
        cs[it] = c;
      }
      wave_append(&ctl->cqn[cb], gcq(g, cb), f, cs);
    }
    gsync(g);
    // cliques at d+1 -> vertices at d+2 (arc v->clique residual (k-1)U-x);
    // a large clique frontier is pulled: each unlabelled vertex scans its
    // own incidence for a clique at d+1 and stops at the first one
    const uint32_t nc = ald(&ctl->cqn[cb]);
    if ((uint64_t)nc * 4 >= g.T) {
      for (uint32_t v = g.gwave; v < g.V; v += g.gwaves) {
        if (g.hv[v] != g.INF) continue;  // wave-uniform
        const uint64_t e = g.ioff[v + 1];
        for (uint64_t p0 = g.ioff[v]; p0 < e; p0 += kWave) {
          const uint64_t p = p0 + lane();
          bool hit = false;
          if (p < e) {
            uint32_t c;
            const uint64_t a = arc_at(g, p, v, &c);
            hit = hcl(g, c) == d + 1 && (long long)(g.k - 1) * g.U - xa(g, a) > 0;
          }
          if (__ballot(hit)) {
            if (lane() == 0) {
              g.hv[v] = d + 2;
              list_push(g, 1 - vb, v);
            }
            break;
          }
        }
      }
    } else
    for (uint32_t i = g.gtid; i < nc; i += g.gthreads) {
      const uint32_t c = gcq(g, cb)[i];
      const uint64_t base = (uint64_t)c * g.ms;
      for (uint32_t j = 0; j < g.k; ++j) {
        if ((long long)(g.k - 1) * g.U - xcj(g, c, j) <= 0) continue;
        const uint32_t v = g.mem[base + j] - g.vbase;
        if (g.hv[v] != g.INF) continue;  // plain filter, as above
        if (atomicCAS(&g.hv[v], g.INF, d + 2) == g.INF) list_push(g, 1 - vb, v);
      }
    }
    gsync(g);
    vb = 1 - vb;
    cb = 1 - cb;
    d += 2;
  }
  // Active nodes; the active cliques also become the clique work list of
  // the rounds that follow (cq0, count cqn[0]; cq1 starts empty).
  if (g.leader) {
    ctl->cqn[0] = 0;
    ctl->cqn[1] = 0;
  }
  gsync(g);
  unsigned long long act = 0;
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) {
    act += g.exv[v] > 0 && g.hv[v] < g.INF;
    g.arc[v] = arc_reset;  // new heights: every arc may be admissible again
  }
  constexpr int kScan = 8;  // cliques per thread and block_append
  for (uint64_t t0 = (uint64_t)g.bid * blockDim.x * kScan; t0 < g.T;
       t0 += (uint64_t)g.nblocks * blockDim.x * kScan) {
    bool on[kScan];
    uint32_t cs[kScan];
#pragma unroll
    for (int it = 0; it < kScan; ++it) {
      const uint64_t c = t0 + (uint64_t)it * blockDim.x + threadIdx.x;
      on[it] = c < g.T && excl(g, c) > 0 && hcl(g, c) < g.INF;
      cs[it] = (uint32_t)c;
      act += on[it];
    }
    block_append(g.app, &ctl->cqn[0], g.cq0, on, cs);
  }
  act = wave_sum64(act);
  if (lane() == 0 && act) atomicAdd(&ctl->active, act);
  gsync(g);
  const unsigned long long total = ald(&ctl->active);
  if (g.leader) {
    ctl->lpack[0] = 0;
    ctl->lpack[1] = 0;
  }
  gsync(g);
  return total;
}

constexpr int kGridFlowThreads = 1024;
constexpr int kFlowBatchThreads = 512;

// Problem fields of G; the caller sets the thread geometry.
template <class Gx>
__device__ __forceinline__ void flow_setup(Gx &g, const FlowProblem &P,
                                           const GridFlowState &S, long long U) {
  g.k = P.k;
  div_magic(P.k, &g.kmag, &g.ksh);
  g.ms = P.mstride ? P.mstride : P.k;
  g.V = P.V;
  g.T = P.T;
  // "unreachable" height: above every finite distance to the sink, which
  // alternates vertex and clique nodes (at most 2V + 1); V + T + 1 while
  // that fits an int comfortably (the labels of a valid labelling never
  // exceed a node's distance, so either bound keeps the flow exact)
  g.INF = (uint64_t)P.V + P.T + 1 < (1u << 30) ? (int)(P.V + P.T) + 1 : (int)(2 * P.V + 3);
  g.vbase = P.vbase;
  g.sbase = P.sbase;
  g.cbase = P.k ? P.sbase / P.k : 0;
  g.U = U;
  g.mem = P.mem;
  g.ioff = P.ioff;
  g.inc = P.inc;
  if constexpr (Gx::kByClique) {
    g.x = S.x32;
    g.exc = S.exc32;
  } else {
    g.x = S.x;
    g.exc = S.exc;
  }
  g.rvs = S.rvs;
  g.exv = S.exv;
  g.rem = S.rem;
  g.hv = S.hv;
  g.hc = S.hc;
  g.hmin = S.hmin;
  g.list0 = S.list0;
  g.list1 = S.list1;
  g.lchunk0 = S.lchunk0;
  g.lchunk1 = S.lchunk1;
  g.cq0 = S.cq0;
  g.cq1 = S.cq1;
  g.arc = S.arc;
  g.astop = S.astop;
  g.ctl = S.ctl;
}

// The max flow of one problem by the threads of g (a whole cooperative
// grid, or one workgroup of a batched launch).  lds_ex (LDS, V entries if
// lds_v) accumulates the excess cliques push to vertices.
// kCacheArcs: the discharge's pass 1 keeps its admissible arcs in
// registers for pass 2 (the whole-grid kernel, one workgroup per CU); the
// batched kernel re-reads them and keeps its registers for two workgroups
// per CU.
template <bool kCacheArcs, class Gx>
__device__ __forceinline__ void flow_run(Gx &g, const FlowProblem &P, const GridFlowState &S,
                         long long cap_t, long long U, uint32_t max_rounds,
                         long long *lds_ex, bool lds_v) {
  GridFlowCtl *ctl = S.ctl;
  const uint64_t slots = (uint64_t)g.T * g.k;

  for (uint64_t s = g.gtid; s < slots; s += g.gthreads) g.x[s] = 0;
  for (uint32_t c = g.gtid; c < g.T; c += g.gthreads) g.exc[c] = 0;
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) {
    g.rvs[v] = (long long)P.cd[v] * U;
    g.exv[v] = cap_t;
    g.arc[v] = kArcFull;
  }
  if (g.leader) {
    ctl->sink = 0;
    ctl->rounds = 0;
    ctl->limit = 0;
    ctl->relabels = 0;
  }
  gsync(g);
  unsigned long long active = global_relabel(g, kArcFull);
  int wb = 0;  // clique work list of this round (global_relabel fills cq0)
  uint32_t rounds = 0, relabels = 1, since = 0;
  unsigned long long sink = 0, wl_visits = 0, ch_visits = 0;
  bool limit = false;
  while (active) {
    if (rounds >= max_rounds) {
      limit = true;
      break;
    }
    const unsigned long long tr0 = (S.trace && g.leader) ? wall_clock64() : 0;
    if (g.leader) {
      ctl->lpack[0] = 0;
      ctl->active = 0;
    }
    // 1. cliques push to lower members, then relabel if excess remains.
    // With few vertices, the excess they receive is summed per workgroup in
    // LDS and flushed with one atomic per vertex.
    if (lds_v) {
      for (uint32_t i = threadIdx.x; i < g.V; i += blockDim.x) lds_ex[i] = 0;
      __syncthreads();
    }
    // Only cliques on the work list (those holding excess) are visited;
    // one that keeps excess moves to the next list, and phase 3 appends a
    // clique when its excess goes from zero to positive, so every clique
    // is on a list at most once.
    unsigned long long act = 0;
    const uint32_t nwl = ald(&ctl->cqn[wb]);
    const uint32_t *wl = gcq(g, wb);
    uint32_t *nx = gcq(g, 1 - wb);
    constexpr int kTile = 4;  // list entries per thread and block_append
    for (uint64_t t0 = (uint64_t)g.bid * blockDim.x * kTile; t0 < nwl;
         t0 += (uint64_t)g.nblocks * blockDim.x * kTile) {
     bool keeps[kTile];  // still holds excess: onto the next list
     uint32_t cs[kTile];
#pragma unroll
     for (int it = 0; it < kTile; ++it) {
      const uint64_t i = t0 + (uint64_t)it * blockDim.x + threadIdx.x;
      bool keep = false;
      const uint32_t c = i < nwl ? wl[i] : 0;
      long long e = i < nwl ? (long long)excl(g, c) : 0;
      const int hcc = i < nwl ? hcl(g, c) : g.INF;
      if (e > 0 && hcc < g.INF) {
        ++act;
        const uint64_t base = (uint64_t)c * g.ms;
        int hmin = g.INF;
        for (uint32_t j = 0; j < g.k && e > 0; ++j) {
          const long long res = U + xcj(g, c, j);
          if (res <= 0) continue;
          const uint32_t v = g.mem[base + j] - g.vbase;
          const int hvv = g.hv[v];  // heights change only in phase 4
          if (hvv < hcc) {
            const long long d = e < res ? e : res;
            xcj(g, c, j) -= d;
            e -= d;
            if (lds_v)
              atomicAdd((unsigned long long *)&lds_ex[v], (unsigned long long)d);
            else
              atomicAdd((unsigned long long *)&g.exv[v], (unsigned long long)d);
            if (d < res) hmin = min(hmin, hvv);
          } else {
            hmin = min(hmin, hvv);
          }
        }
        excl(g, c) = e;
        if (e > 0) {
          // all lower members are saturated; recompute over residual arcs
          hmin = g.INF;
          for (uint32_t j = 0; j < g.k; ++j)
            if (U + xcj(g, c, j) > 0)
              hmin = min(hmin, g.hv[g.mem[base + j] - g.vbase]);
          const int hn = hmin >= g.INF - 1 ? g.INF : hmin + 1;
          hcl(g, c) = hn;
          keep = hn < g.INF;
        }
      }
      keeps[it] = keep;
      cs[it] = c;
     }
     block_append(g.app, &ctl->cqn[1 - wb], nx, keeps, cs);
    }
    if (lds_v) {
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < g.V; i += blockDim.x) {
        const long long a = lds_ex[i];
        if (a) atomicAdd((unsigned long long *)&g.exv[i], (unsigned long long)a);
      }
    }
    gsync(g);
    const unsigned long long tr1 = (S.trace && g.leader) ? wall_clock64() : 0;
    // 2. vertices push to the sink; active vertices post their chunks.
    // Every workgroup has read this round's clique list count: it is
    // cleared for the round after next.
    if (g.leader) ctl->cqn[wb] = 0;
    for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) {
      long long e = g.exv[v];
      if (e <= 0 || g.hv[v] >= g.INF) {
        g.rem[v] = 0;
        continue;
      }
      ++act;
      const long long r = g.rvs[v];
      if (r > 0) {
        const long long d = e < r ? e : r;
        g.rvs[v] = r - d;
        e -= d;
        sink += (unsigned long long)d;
      }
      g.exv[v] = e;
      g.rem[v] = e;
      g.hmin[v] = g.INF;
      if (e > 0) {
        uint32_t first, count;
        discharge_range(g, v, &first, &count);
        g.astop[v] = 0xFFFFFFFFu;
        list_push_n(g, 0, v, count);
      }
    }
    act = wave_sum64(act);
    if (lane() == 0 && act) atomicAdd(&ctl->active, act);
    gsync(g);
    const unsigned long long pack = ald(&ctl->lpack[0]);
    const uint32_t nv = (uint32_t)(pack >> 40);
    const uint64_t nch = pack & kMask40;
    const unsigned long long any = ald(&ctl->active);

    // 3. chunked discharge of active vertices over their incidence lists:
    // pass 1 sums the chunk's admissible residual (and the lowest
    // non-admissible height), one atomic reserves that much of the vertex's
    // remaining excess, pass 2 hands the reservation to the admissible arcs
    // in chunk order (wave prefix sums).  A slot belongs to one chunk and
    // clique heights do not change in this phase, so both passes agree, and
    // with kCacheArcs pass 1 keeps each lane's admissible residuals and
    // slots in registers for pass 2 instead of re-reading them (random
    // lines, long evicted).
    for (uint64_t ci = g.gwave; ci < nch; ci += g.gwaves) {
      uint32_t v;
      uint64_t lo, hi;
      uint32_t vchunk;
      chunk_of(g, 0, nv, ci, &v, &lo, &hi, true, &vchunk);
      const int hvv = g.hv[v];
      int hm = g.INF;
      long long radm = 0;
      typename Gx::X rav[kCacheArcs ? kChunkSteps : 1];
      uint32_t sv[kCacheArcs ? kChunkSteps : 1];
      if constexpr (kCacheArcs) {
#pragma unroll
        for (int it = 0; it < kChunkSteps; ++it) {
          rav[it] = 0;
          sv[it] = 0;
          const uint64_t p = lo + (uint64_t)it * kWave + lane();
          if (p >= hi) continue;
          uint32_t c;
          const uint64_t a = arc_at(g, p, v, &c);
          const long long res = (long long)(g.k - 1) * U - xa(g, a);
          if (res <= 0) continue;
          const int hcc = hcl(g, c);
          if (hcc < hvv) {
            radm += res;
            rav[it] = (typename Gx::X)res;
            sv[it] = (uint32_t)a;
          } else {
            hm = min(hm, hcc);
          }
        }
      } else {
        for (uint64_t p = lo + lane(); p < hi; p += kWave) {
          uint32_t c;
          const uint64_t a = arc_at(g, p, v, &c);
          const long long res = (long long)(g.k - 1) * U - xa(g, a);
          if (res <= 0) continue;
          const int hcc = hcl(g, c);
          if (hcc < hvv)
            radm += res;
          else
            hm = min(hm, hcc);
        }
      }
      radm = (long long)wave_sum64((uint64_t)radm);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) hm = min(hm, __shfl_xor(hm, off, kWave));
      if (lane() == 0 && hm < g.INF) atomicMin(&g.hmin[v], hm);
      if (radm == 0) continue;
      long long avail = 0;
      if (lane() == 0) {
        const long long old = (long long)atomicAdd(
            (unsigned long long *)&g.rem[v], (unsigned long long)(-radm));
        avail = old < 0 ? 0 : (old < radm ? old : radm);
      }
      avail = __shfl(avail, 0, kWave);
      // this chunk keeps admissible residual: the current arc stops here
      if (lane() == 0 && avail < radm) atomicMin(&g.astop[v], vchunk);
      long long before = 0;
      bool fresh[kChunkSteps];  // the clique's excess goes from zero to positive
      uint32_t fc[kChunkSteps];
#pragma unroll
      for (int it = 0; it < kChunkSteps; ++it) {
        fresh[it] = false;
        fc[it] = 0;
        const uint64_t p0 = lo + (uint64_t)it * kWave;
        if (p0 >= hi || before >= avail) continue;  // wave-uniform
        long long ra = 0;
        uint64_t a = 0;
        uint32_t c = 0;
        if constexpr (kCacheArcs) {
          ra = rav[it];
          a = sv[it];
          c = clique_of_arc(g, (uint32_t)a);
        } else {
          const uint64_t p = p0 + lane();
          if (p < hi) {
            a = arc_at(g, p, v, &c);
            const long long res = (long long)(g.k - 1) * U - xa(g, a);
            if (res > 0 && hcl(g, c) < hvv) ra = res;
          }
        }
        long long incl = ra;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
          const long long t = __shfl_up(incl, off, kWave);
          if ((int)lane() >= off) incl += t;
        }
        const long long start = before + incl - ra;
        long long take = avail - start;
        take = take < 0 ? 0 : (take < ra ? take : ra);
        if (take > 0) {
          xa(g, a) += take;
          fc[it] = c;
          fresh[it] = exc_add(&excl(g, c), take) == 0;
        }
        before += __shfl(incl, kWave - 1, kWave);
      }
      wave_append(&ctl->cqn[1 - wb], nx, fresh, fc);
    }
    gsync(g);
    const unsigned long long tr2 = (S.trace && g.leader) ? wall_clock64() : 0;
    // 4. remainders become excess, and the current arc moves.  Arcs before
    // the current arc are saturated or not admissible, and stay so until
    // the vertex relabels: pushes into the vertex come from higher cliques
    // and heights only grow.
    //  - excess left after a scan of the whole list: relabel (hmin covers
    //    every residual arc); the current arc restarts at chunk 0;
    //  - excess left after a window: every admissible arc in it is
    //    saturated, so the current arc moves past it and the window
    //    doubles; past the end of the list, the next round scans the whole
    //    list and relabels;
    //  - excess used up: the current arc moves to the first chunk that kept
    //    admissible residual, and the window becomes twice what was used.
    for (uint32_t i = g.gtid; i < nv; i += g.gthreads) {
      const uint32_t v = g.list0[i];
      long long r = g.rem[v];
      if (r < 0) r = 0;
      g.exv[v] = r;
      uint32_t first, count;
      discharge_range(g, v, &first, &count);
      const uint32_t nch = (uint32_t)nchunks(g, v);
      const uint32_t win = (uint32_t)(g.arc[v] >> 32);
      uint32_t cur, nwin;
      if (r > 0 && first == 0 && count == nch) {
        const int hm = g.hmin[v];
        g.hv[v] = hm >= g.INF - 1 ? g.INF : hm + 1;
        cur = 0;
        nwin = kArcWindowRelabel;
      } else if (r > 0) {
        cur = first + count;  // == nch: the whole list next round
        nwin = min(2 * win, kArcWindowMax);
      } else {
        const uint32_t stop = g.astop[v];
        cur = stop < first + count ? stop : first + count;
        const uint32_t used = cur - first + 1;
        nwin = min(max(2 * used, kArcWindow0), kArcWindowMax);
      }
      g.arc[v] = (uint64_t)cur | (uint64_t)nwin << 32;
    }
    gsync(g);
    const unsigned long long tr3 = (S.trace && g.leader) ? wall_clock64() : 0;
    const bool after_relabel = since == 0;
    ++rounds;
    ++since;
    wb = 1 - wb;
    if (any == 0 || since >= relabel_rounds(slots)) {
      active = global_relabel(g, kArcAfterRelabel);
      wb = 0;
      ++relabels;
      since = 0;
    }
    wl_visits += nwl;
    ch_visits += nch;
    if (S.trace && g.leader && rounds - 1 < S.trace_cap) {
      unsigned long long *t = S.trace + 8ull * (rounds - 1);
      t[0] = nwl;
      t[1] = nv;
      t[2] = nch;
      t[3] = any | (after_relabel ? 1ull << 63 : 0);  // top bit: first after a relabel
      t[4] = tr1 - tr0;  // phase 1 (wall_clock64 ticks)
      t[5] = tr2 - tr1;  // phases 2 and 3
      t[6] = tr3 - tr2;  // phase 4
      t[7] = wall_clock64() - tr3;  // global relabel, if one ran
    }
  }
  if (!limit) global_relabel(g, kArcFull);
  sink = wave_sum64(sink);
  if (lane() == 0 && sink) atomicAdd(&ctl->sink, sink);
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads)
    S.reach[v] = g.hv[v] < g.INF ? 1 : 0;
  if (g.leader) {
    ctl->rounds = rounds;
    ctl->limit = limit ? 1 : 0;
    ctl->relabels = relabels;
    ctl->clique_visits = wl_visits;
    ctl->chunk_visits = ch_visits;
  }
}

// ---------------------------------------------------------------------------
// Implicit clique nodes.  A clique node's excess is always zero and its
// height is derived, not stored: h(c) = 1 + the lowest height among the
// members its residual arcs c->w (U + x > 0) reach, INF if none.  A vertex
// v discharging into c therefore pushes straight through it: flow v->c
// (x[v] += t) and c->w (x[w] -= t) to the lowest members w in one step,
// so there is no clique phase, no clique work list and no clique height or
// excess array.  Each clique is one record of rs words: its k member ids,
// then its k residual words x (int32; a residual stays in [-U, (k-1)U]).
//
// A round: (A) vertices push to the sink and post their discharge windows;
// (B) chunked discharge: pass 1 sums per arc what can go through the
// clique (min of the arc's residual and the residual to the clique's
// lowest members), one atomic reserves that much of the vertex's excess,
// pass 2 pushes it, taking each c->w residual with a compare-and-swap so no
// arc is ever over-used (a push that finds less left returns the rest to
// the vertex); what reaches a member is summed in LDS per workgroup and
// added to its inflow; (D) vertices keep the remainder plus inflow as
// excess, move their current arc, and relabel after a whole-list scan
// that found no admissible arc (a push limited by the clique's lowest
// members leaves v->c residual, the clique one level up: such a vertex
// scans its whole list again instead).
//
// Heights stay a valid labelling.  Pushes go from v to members two levels
// down (h(w) <= h(v) - 2, h(c) = h(w) + 1).  Flow into v through c makes
// v->c residual again with h(c) = h(v) + 1, so a vertex that received
// inflow in a round relabels to at most h(v) + 2.  Vertex heights change
// only in (D) and clique heights never fall (a clique gains a residual arc
// c->u only when u pushes into it from above), so an arc skipped as not
// admissible stays so until the vertex relabels, as in the explicit
// network.  The answer never rests on the labels: the solve ends when an
// exact BFS from the sink finds no vertex with excess that reaches it.

template <bool kByClique_>
struct H {
  static constexpr bool kByClique = kByClique_;
  uint32_t k, V, T, rs;
  uint32_t kmag, ksh;
  int INF;
  uint32_t vbase;
  uint64_t sbase, cbase;
  long long U;
  uint32_t *rec;  // records of this problem's cliques (local clique ids)
  const uint64_t *ioff;
  const uint32_t *inc;
  long long *rvs, *exv, *rem, *inflow;
  int *hv, *hmin, *pend;
  uint32_t *list0, *list1;
  uint64_t *lchunk0, *lchunk1;
  uint64_t *arc;
  uint32_t *astop;
  uint32_t *app;
  int *lds_h;           // LDS copy of hv (V entries) or null
  long long *lds_in;    // LDS inflow accumulator (V entries) or null
  GridFlowCtl *ctl;
  uint32_t gtid, gthreads, gwave, gwaves;
  bool leader;
  uint32_t nblocks, bid;
};

constexpr uint32_t kImpMaxK = 6;
// No relabel pending for a vertex (pend[]); below it, pend holds the
// lowest height of a clique behind a residual arc (the new height is one
// more), or -1 when the vertex found an admissible arc instead.
constexpr int kNoPend = 0x7FFFFFFF;

// Record stride in words for clique size k: 2k words rounded up to 4, 8
// or 16, so a record never crosses a 64-byte line.
__host__ __device__ inline uint32_t imp_stride(uint32_t k) {
  return 2 * k <= 4 ? 4u : 2 * k <= 8 ? 8u : 16u;
}

// Clique (local id) and the member position of incidence entry p of v.
template <class Hx>
__device__ __forceinline__ uint32_t imp_arc(const Hx &g, uint64_t p, uint32_t v,
                                            uint32_t *jv) {
  if constexpr (Hx::kByClique) {
    const uint32_t c = (uint32_t)((uint64_t)g.inc[p] - g.cbase);
    const uint32_t *r = g.rec + (uint64_t)c * g.rs;
    const uint32_t gv = v + g.vbase;
    uint32_t j = 0;
    while (j + 1 < g.k && r[j] != gv) ++j;
    *jv = j;
    return c;
  } else {
    const uint32_t s = (uint32_t)((uint64_t)g.inc[p] - g.sbase);
    const uint32_t c = __umulhi(s, g.kmag) >> g.ksh;
    *jv = s - c * g.k;
    return c;
  }
}

// A record's k member ids (local) and residual words.  The arrays are
// indexed by unrolled constants only (a run-time index would put them in
// scratch memory); pick() selects by a run-time position.
template <class Hx>
__device__ __forceinline__ void imp_load(const Hx &g, uint32_t c, uint32_t (&m)[kImpMaxK],
                                         int (&x)[kImpMaxK]) {
  const uint32_t *r = g.rec + (uint64_t)c * g.rs;
#pragma unroll
  for (uint32_t j = 0; j < kImpMaxK; ++j) {
    m[j] = 0;
    x[j] = 0;
  }
  if (g.rs == 8) {  // k = 3 or 4
    const uint4 a = *reinterpret_cast<const uint4 *>(r);
    const uint4 b = *reinterpret_cast<const uint4 *>(r + 4);
    const bool k3 = g.k == 3;
    m[0] = a.x - g.vbase;
    m[1] = a.y - g.vbase;
    m[2] = a.z - g.vbase;
    m[3] = k3 ? 0u : a.w - g.vbase;
    x[0] = (int)(k3 ? a.w : b.x);
    x[1] = (int)(k3 ? b.x : b.y);
    x[2] = (int)(k3 ? b.y : b.z);
    x[3] = k3 ? 0 : (int)b.w;
  } else if (g.rs == 4) {  // k = 2
    const uint4 a = *reinterpret_cast<const uint4 *>(r);
    m[0] = a.x - g.vbase;
    m[1] = a.y - g.vbase;
    x[0] = (int)a.z;
    x[1] = (int)a.w;
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kImpMaxK; ++j) {
      if (j < g.k) {
        m[j] = r[j] - g.vbase;
        x[j] = (int)r[g.k + j];
      }
    }
  }
}

// Residual word j of clique c, read from the record (a select over the
// loaded words would be turned into a scratch-memory lookup).
template <class Hx>
__device__ __forceinline__ int imp_x(const Hx &g, uint32_t c, uint32_t j) {
  return (int)g.rec[(uint64_t)c * g.rs + g.k + j];
}

template <class Hx>
__device__ __forceinline__ int imp_h(const Hx &g, uint32_t v) {
  return g.lds_h ? g.lds_h[v] : g.hv[v];
}

// Copies hv into the workgroup's LDS copy (if any); every thread calls it.
template <class Hx>
__device__ __forceinline__ void imp_stage_h(const Hx &g) {
  if (!g.lds_h) return;
  for (uint32_t i = threadIdx.x; i < g.V; i += blockDim.x) g.lds_h[i] = g.hv[i];
  __syncthreads();
}

// Exact BFS distances to the sink over the residual network, vertex levels
// only (a clique sits one above its lowest residual member); returns the
// vertices with excess that reach the sink.
template <class Hx>
__device__ __forceinline__ unsigned long long imp_relabel(const Hx &g, uint64_t arc_reset) {
  GridFlowCtl *ctl = g.ctl;
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) {
    g.hv[v] = g.INF;
    g.pend[v] = kNoPend;  // exact heights replace a pending relabel
  }
  if (g.leader) {
    ctl->lpack[0] = 0;
    ctl->lpack[1] = 0;
    ctl->active = 0;
  }
  gsync(g);
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads)
    if (g.rvs[v] > 0) {
      g.hv[v] = 1;
      list_push(g, 0, v);
    }
  gsync(g);
  const long long kU = (long long)(g.k - 1) * g.U;
  const uint64_t arcs = (uint64_t)g.T * g.k;
  int vb = 0, d = 1;
  uint32_t labelled = 0;  // vertices with a height so far (each is listed once)
  for (;;) {
    const unsigned long long pack = ald(&ctl->lpack[vb]);
    const uint32_t nv = (uint32_t)(pack >> 40);
    // clique heights are derived, so once every vertex has its height the
    // next level's scan would label nothing (at the start every vertex
    // with sink residual, usually all of them, is at height 1)
    labelled += nv;
    if (nv == 0 || labelled >= g.V) break;
    const uint64_t nch = pack & kMask40;
    if (g.leader) ctl->lpack[1 - vb] = 0;
    gsync(g);
    imp_stage_h(g);
    if (nch * kGridChunk * 4 >= arcs) {
      // pull: each unlabelled vertex looks for a residual arc to a clique
      // with a residual arc to a vertex at d, and stops at the first
      for (uint32_t v = g.gwave; v < g.V; v += g.gwaves) {
        if (imp_h(g, v) != g.INF) continue;  // wave-uniform
        const uint64_t e = g.ioff[v + 1];
        for (uint64_t p0 = g.ioff[v]; p0 < e; p0 += kWave) {
          const uint64_t p = p0 + lane();
          bool hit = false;
          if (p < e) {
            uint32_t jv;
            const uint32_t c = imp_arc(g, p, v, &jv);
            uint32_t m[kImpMaxK];
            int x[kImpMaxK];
            imp_load(g, c, m, x);
            if (kU - imp_x(g, c, jv) > 0) {
#pragma unroll
              for (uint32_t j = 0; j < kImpMaxK; ++j)
                if (j < g.k && j != jv && g.U + x[j] > 0 && imp_h(g, m[j]) == d) hit = true;
            }
          }
          if (__ballot(hit)) {
            if (lane() == 0) {
              g.hv[v] = d + 2;
              list_push(g, 1 - vb, v);
            }
            break;
          }
        }
      }
    } else {
      // push: the frontier's arcs label the other members behind them
      for (uint64_t ci = g.gwave; ci < nch; ci += g.gwaves) {
        uint32_t w;
        uint64_t lo, hi;
        chunk_of(g, vb, nv, ci, &w, &lo, &hi);
        for (uint64_t p = lo + lane(); p < hi; p += kWave) {
          uint32_t jw;
          const uint32_t c = imp_arc(g, p, w, &jw);
          uint32_t m[kImpMaxK];
          int x[kImpMaxK];
          imp_load(g, c, m, x);
          if (g.U + imp_x(g, c, jw) <= 0) continue;
#pragma unroll
          for (uint32_t j = 0; j < kImpMaxK; ++j) {
            if (j >= g.k || j == jw || kU - x[j] <= 0) continue;
            const uint32_t v = m[j];
            if (g.hv[v] != g.INF) continue;  // plain filter; the CAS decides
            if (atomicCAS(&g.hv[v], g.INF, d + 2) == g.INF) list_push(g, 1 - vb, v);
          }
        }
      }
    }
    gsync(g);
    vb = 1 - vb;
    d += 2;
  }
  unsigned long long act = 0;
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) {
    act += g.exv[v] > 0 && g.hv[v] < g.INF;
    g.arc[v] = arc_reset;
  }
  act = wave_sum64(act);
  if (lane() == 0 && act) atomicAdd(&ctl->active, act);
  gsync(g);
  const unsigned long long total = ald(&ctl->active);
  if (g.leader) {
    ctl->lpack[0] = 0;
    ctl->lpack[1] = 0;
  }
  gsync(g);
  return total;
}

template <class Hx>
__device__ __forceinline__ void imp_setup(Hx &g, const FlowProblem &P, const GridFlowState &S,
                                          long long U) {
  g.k = P.k;
  div_magic(P.k, &g.kmag, &g.ksh);
  g.rs = imp_stride(P.k);
  g.V = P.V;
  g.T = P.T;
  // no path to the sink is longer than 2V + 1 vertex levels apart
  g.INF = (int)(2 * P.V + 3);
  g.vbase = P.vbase;
  g.sbase = P.sbase;
  g.cbase = P.k ? P.sbase / P.k : 0;
  g.U = U;
  g.rec = S.rec32;
  g.ioff = P.ioff;
  g.inc = P.inc;
  g.rvs = S.rvs;
  g.exv = S.exv;
  g.rem = S.rem;
  g.inflow = S.inflow;
  g.hv = S.hv;
  g.hmin = S.hmin;
  g.pend = S.pend;
  g.list0 = S.list0;
  g.list1 = S.list1;
  g.lchunk0 = S.lchunk0;
  g.lchunk1 = S.lchunk1;
  g.arc = S.arc;
  g.astop = S.astop;
  g.ctl = S.ctl;
}

constexpr int kImpSteps = kChunkSteps;  // wave passes per chunk

// One arc of v as the relabel scan and the discharge see it: the clique's
// height derived from its residual members (INF if none), whether v->c is
// residual, and through = what c can pass on to its lowest members.
template <class Hx>
__device__ __forceinline__ void imp_view(const Hx &g, uint64_t p, uint32_t v, long long kU,
                                         long long U, uint32_t *c, uint32_t *jv,
                                         long long *rvc, int *hcc, int *mh,
                                         long long *through) {
  *c = imp_arc(g, p, v, jv);
  uint32_t m[kImpMaxK];
  int x[kImpMaxK];
  imp_load(g, *c, m, x);
  *rvc = kU - imp_x(g, *c, *jv);
  int lo = g.INF;
  long long th = 0;
#pragma unroll
  for (uint32_t j = 0; j < kImpMaxK; ++j) {
    if (j >= g.k) break;
    const long long rc = U + x[j];
    if (rc <= 0) continue;
    const int h = imp_h(g, m[j]);
    if (h < lo) {
      lo = h;
      th = rc;
    } else if (h == lo) {
      th += rc;
    }
  }
  *mh = lo;
  *through = th;
  *hcc = lo >= g.INF - 1 ? g.INF : lo + 1;
}

template <class Hx>
__device__ __forceinline__ void imp_run(Hx &g, const FlowProblem &P, const GridFlowState &S,
                                        long long cap_t, long long U, uint32_t max_rounds) {
  GridFlowCtl *ctl = S.ctl;
  const long long kU = (long long)(g.k - 1) * U;
  // residual words start at zero
  for (uint64_t c = g.gtid; c < g.T; c += g.gthreads) {
    uint32_t *r = g.rec + c * g.rs;
    for (uint32_t j = 0; j < g.k; ++j) r[g.k + j] = 0;
  }
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) {
    g.rvs[v] = (long long)P.cd[v] * U;
    g.exv[v] = cap_t;
    g.inflow[v] = 0;
    g.arc[v] = kArcFull;
  }
  if (g.leader) {
    ctl->sink = 0;
    ctl->rounds = 0;
    ctl->limit = 0;
    ctl->relabels = 0;
    ctl->cqn[0] = 0;
    ctl->cqn[1] = 0;
  }
  gsync(g);
  unsigned long long active = imp_relabel(g, kArcFull);
  uint32_t rounds = 0, relabels = 1, since = 0;
  unsigned long long sink = 0, ch_visits = 0;
  bool limit = false;
  const uint64_t slots = (uint64_t)g.T * g.k;
  // Counters: list 0 (this round's discharge windows) and the round's
  // active-vertex count (cqn[0]) are cleared in (C) and filled in (A); list
  // 1 (relabel candidates) is cleared in (A) and filled in (D); each is
  // read a phase later.
  // The first round, when no flow has moved yet and every vertex has its
  // sink arc (height 1, so every clique is at height 2): the sink pushes,
  // and since no clique arc is admissible, a vertex that keeps excess would
  // scan its whole list twice to find the relabel to 3 (every arc v->c is
  // residual) that is set here directly, as those scans would set it.
  {
    unsigned off1 = 0;
    for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) off1 += g.hv[v] != 1 ? 1u : 0u;
    off1 = wave_sum(off1);
    if (lane() == 0 && off1) atomicAdd(&ctl->cqn[1], off1);
    gsync(g);
  }
  if (active && ald(&ctl->cqn[1]) == 0u && max_rounds > 0) {
    for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) {
      g.inflow[v] = 0;
      long long e = g.exv[v];
      const long long r = g.rvs[v];
      if (e > 0 && r > 0) {
        const long long d = e < r ? e : r;
        g.rvs[v] = r - d;
        e -= d;
        sink += (unsigned long long)d;
      }
      g.exv[v] = e > 0 ? e : 0;
      g.rem[v] = e > 0 ? e : 0;
      g.hmin[v] = 0;
      if (e > 0) {
        g.pend[v] = 2;
        g.arc[v] = (uint64_t)kArcWindowRelabel << 32;
      }
    }
    if (S.trace && g.leader && S.trace_cap > 0)
      for (int j = 0; j < 8; ++j) S.trace[j] = 0;
    gsync(g);
    rounds = 1;
    since = 1;
  }
  while (active) {
    if (rounds >= max_rounds) {
      limit = true;
      break;
    }
    const unsigned long long tr0 = (S.trace && g.leader) ? wall_clock64() : 0;
    if (g.leader) ctl->lpack[1] = 0;
    // (A) pending relabels take effect; sink pushes; active vertices post
    // their windows; inflow restarts
    unsigned long long act = 0;
    for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) {
      const int pd = g.pend[v];
      if (pd != kNoPend) {
        if (pd >= 0) g.hv[v] = pd >= g.INF - 1 ? g.INF : pd + 1;
        g.pend[v] = kNoPend;
      }
      g.inflow[v] = 0;
      long long e = g.exv[v];
      if (e <= 0 || g.hv[v] >= g.INF) {
        g.rem[v] = e > 0 ? e : 0;  // stranded excess is kept as it is
        continue;
      }
      ++act;
      const long long r = g.rvs[v];
      if (r > 0) {
        const long long d = e < r ? e : r;
        g.rvs[v] = r - d;
        e -= d;
        sink += (unsigned long long)d;
      }
      g.exv[v] = e;
      g.rem[v] = e;
      g.hmin[v] = 0;
      if (e > 0) {
        uint32_t first, count;
        discharge_range(g, v, &first, &count);
        g.astop[v] = 0xFFFFFFFFu;
        list_push_n(g, 0, v, count);
      }
    }
    act = wave_sum64(act);
    if (lane() == 0 && act) atomicAdd(&ctl->cqn[0], (unsigned)act);
    gsync(g);
    const unsigned long long tr1 = (S.trace && g.leader) ? wall_clock64() : 0;
    const unsigned long long pack = ald(&ctl->lpack[0]);
    const uint32_t nv = (uint32_t)(pack >> 40);
    const uint64_t nch = pack & kMask40;
    const unsigned long long any = ald(&ctl->cqn[0]);
    imp_stage_h(g);
    if (g.lds_in) {
      for (uint32_t i = threadIdx.x; i < g.V; i += blockDim.x) g.lds_in[i] = 0;
      __syncthreads();
    }
    // (B) chunked discharge through the cliques.  hmin[v] becomes 1 when
    // v has an admissible arc this round (it then does not relabel).
    for (uint64_t ci = g.gwave; ci < nch; ci += g.gwaves) {
      uint32_t v;
      uint64_t lo, hi;
      uint32_t vchunk;
      chunk_of(g, 0, nv, ci, &v, &lo, &hi, true, &vchunk);
      const int hvv = imp_h(g, v);
      long long radm = 0;
      int av[kImpSteps];       // what this lane's arc can pass
      uint32_t cv[kImpSteps];  // its clique
      uint32_t tv[kImpSteps];  // lowest member height << 3 | v's position
#pragma unroll
      for (int it = 0; it < kImpSteps; ++it) {
        av[it] = 0;
        cv[it] = 0;
        tv[it] = 0;
        const uint64_t p = lo + (uint64_t)it * kWave + lane();
        if (p >= hi) continue;
        uint32_t c, jv;
        long long rvc, through;
        int hcc, mh;
        imp_view(g, p, v, kU, U, &c, &jv, &rvc, &hcc, &mh, &through);
        if (rvc <= 0 || hcc >= hvv) continue;  // not a residual / admissible arc
        const long long a = rvc < through ? rvc : through;
        radm += a;
        av[it] = (int)a;
        cv[it] = c;
        tv[it] = (uint32_t)mh << 3 | jv;
      }
      radm = (long long)wave_sum64((uint64_t)radm);
      if (radm == 0) continue;
      long long avail = 0;
      if (lane() == 0) {
        g.hmin[v] = 1;
        const long long old = (long long)atomicAdd(
            (unsigned long long *)&g.rem[v], (unsigned long long)(-radm));
        avail = old < 0 ? 0 : (old < radm ? old : radm);
      }
      avail = __shfl(avail, 0, kWave);
      if (lane() == 0 && avail < radm) atomicMin(&g.astop[v], vchunk);
      long long before = 0, back = 0;
#pragma unroll
      for (int it = 0; it < kImpSteps; ++it) {
        if (before >= avail) break;  // wave-uniform
        const long long ra = av[it];
        long long incl = ra;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
          const long long t = __shfl_up(incl, off, kWave);
          if ((int)lane() >= off) incl += t;
        }
        const long long start = before + incl - ra;
        long long take = avail - start;
        take = take < 0 ? 0 : (take < ra ? take : ra);
        if (take > 0) {
          const uint32_t c = cv[it], jv = tv[it] & 7u;
          const int mh = (int)(tv[it] >> 3);
          uint32_t *r = g.rec + (uint64_t)c * g.rs;
          long long left = take;
          for (uint32_t j = 0; j < g.k && left > 0; ++j) {
            if (j == jv) continue;
            const uint32_t w = r[j] - g.vbase;
            if (imp_h(g, w) != mh) continue;
            int *xp = reinterpret_cast<int *>(r + g.k + j);
            int old = __hip_atomic_load(xp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (;;) {
              const long long room = U + old;
              if (room <= 0) break;
              const int t = (int)(left < room ? left : room);
              const int prev = atomicCAS(xp, old, old - t);
              if (prev == old) {
                left -= t;
                if (g.lds_in)
                  atomicAdd((unsigned long long *)&g.lds_in[w], (unsigned long long)t);
                else
                  atomicAdd((unsigned long long *)&g.inflow[w], (unsigned long long)t);
                break;
              }
              old = prev;
            }
          }
          const long long got = take - left;
          if (got > 0)
            atomicAdd(reinterpret_cast<int *>(r + g.k + jv), (int)got);
          back += left;
        }
        before += __shfl(incl, kWave - 1, kWave);
      }
      // what a push could not pass on stays v's excess; rem may already be
      // below zero (the reservation can overshoot), so it joins the inflow
      back = (long long)wave_sum64((uint64_t)back);
      if (lane() == 0 && back)
        atomicAdd((unsigned long long *)&g.inflow[v], (unsigned long long)back);
    }
    if (g.lds_in) {
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < g.V; i += blockDim.x) {
        const long long a = g.lds_in[i];
        if (a) atomicAdd((unsigned long long *)&g.inflow[i], (unsigned long long)a);
      }
    }
    gsync(g);
    const unsigned long long tr2 = (S.trace && g.leader) ? wall_clock64() : 0;
    // (D) current arcs of the listed vertices (their excess left in rem);
    // a vertex that scanned its whole list, found no admissible arc and
    // kept excess becomes a relabel candidate (list 1); every vertex's
    // excess = remainder + inflow
    for (uint32_t i = g.gtid; i < nv; i += g.gthreads) {
      const uint32_t v = g.list0[i];
      const long long r = g.rem[v];
      uint32_t first, count;
      discharge_range(g, v, &first, &count);
      const uint32_t nchv = (uint32_t)nchunks(g, v);
      const uint32_t win = (uint32_t)(g.arc[v] >> 32);
      uint32_t cur, nwin;
      if (r > 0 && first == 0 && count == nchv && g.hmin[v] == 0) {
        g.pend[v] = g.INF;
        list_push_n(g, 1, v, nchv);
        cur = 0;
        nwin = kArcWindowRelabel;
      } else if (r > 0) {
        cur = first + count;
        nwin = min(2 * win, kArcWindowMax);
      } else {
        const uint32_t stop = g.astop[v];
        cur = stop < first + count ? stop : first + count;
        const uint32_t used = cur - first + 1;
        nwin = min(max(2 * used, kArcWindow0), kArcWindowMax);
      }
      g.arc[v] = (uint64_t)cur | (uint64_t)nwin << 32;
    }
    for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) {
      const long long r = g.rem[v];
      g.exv[v] = (r > 0 ? r : 0) + g.inflow[v];
    }
    gsync(g);
    const unsigned long long tr3 = (S.trace && g.leader) ? wall_clock64() : 0;
    // (C) relabel scan of the candidates over the settled network: no
    // residual word changes in this phase and heights only in (A), so the
    // lowest clique height behind a residual arc is exact; a candidate
    // that now finds an admissible arc (flow came in through a clique)
    // does not relabel.  The new heights apply at the next (A).
    {
      const unsigned long long pk = ald(&ctl->lpack[1]);
      const uint32_t nv1 = (uint32_t)(pk >> 40);
      const uint64_t nch1 = pk & kMask40;
      if (g.leader) {
        ctl->lpack[0] = 0;
        ctl->cqn[0] = 0;
      }
      for (uint64_t ci = g.gwave; ci < nch1; ci += g.gwaves) {
        uint32_t v;
        uint64_t lo, hi;
        chunk_of(g, 1, nv1, ci, &v, &lo, &hi);
        const int hvv = imp_h(g, v);
        int hm = g.INF;
        bool adm = false;
        for (uint64_t p = lo + lane(); p < hi; p += kWave) {
          uint32_t c, jv;
          long long rvc, through;
          int hcc, mh;
          imp_view(g, p, v, kU, U, &c, &jv, &rvc, &hcc, &mh, &through);
          if (rvc <= 0) continue;
          if (hcc < hvv)
            adm = true;
          else
            hm = min(hm, hcc);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) hm = min(hm, __shfl_xor(hm, off, kWave));
        if (__ballot(adm)) hm = -1;
        if (lane() == 0) atomicMin(&g.pend[v], hm);
      }
      ch_visits += nch1;
    }
    gsync(g);
    const bool after_relabel = since == 0;
    ++rounds;
    ++since;
    if (any == 0 || since >= relabel_rounds(slots)) {
      active = imp_relabel(g, kArcAfterRelabel);
      ++relabels;
      since = 0;
    }
    ch_visits += nch;
    if (S.trace && g.leader && rounds - 1 < S.trace_cap) {
      unsigned long long *t = S.trace + 8ull * (rounds - 1);
      t[0] = 0;
      t[1] = nv;
      t[2] = nch;
      t[3] = any | (after_relabel ? 1ull << 63 : 0);
      t[4] = tr1 - tr0;  // (A)
      t[5] = tr2 - tr1;  // (B)
      t[6] = tr3 - tr2;  // (D)
      t[7] = wall_clock64() - tr3;  // (C) and the global relabel
    }
  }
  // without the round limit the loop ended on a relabel that found no
  // active vertex, so hv already holds the final exact heights
  sink = wave_sum64(sink);
  if (lane() == 0 && sink) atomicAdd(&ctl->sink, sink);
  for (uint32_t v = g.gtid; v < g.V; v += g.gthreads) S.reach[v] = g.hv[v] < g.INF ? 1 : 0;
  if (g.leader) {
    ctl->rounds = rounds;
    ctl->limit = limit ? 1 : 0;
    ctl->relabels = relabels;
    ctl->clique_visits = 0;
    ctl->chunk_visits = ch_visits;
  }
}

template <bool kByClique>
__device__ __forceinline__ void imp_grid_body(FlowProblem P, GridFlowState S, long long cap_t,
                                              long long U, uint32_t max_rounds) {
  H<kByClique> g;
  imp_setup(g, P, S, U);
  g.gtid = blockIdx.x * blockDim.x + threadIdx.x;
  g.gthreads = gridDim.x * blockDim.x;
  g.gwave = g.gtid / kWave;
  g.gwaves = g.gthreads / kWave;
  g.leader = g.gtid == 0;
  g.nblocks = gridDim.x;
  g.bid = blockIdx.x;
  __shared__ uint32_t app[kMaxFlowWaves + 1];
  g.app = app;
  extern __shared__ long long lds_imp[];
  const bool lds = g.V <= kGridLdsVertices;
  g.lds_in = lds ? lds_imp : nullptr;
  g.lds_h = lds ? reinterpret_cast<int *>(lds_imp + g.V) : nullptr;
  imp_run(g, P, S, cap_t, U, max_rounds);
}

__global__ __launch_bounds__(kGridFlowThreads) void k_grid_flow_imp(FlowProblem P,
                                                       GridFlowState S,
                                                       long long cap_t, long long U,
                                                       uint32_t max_rounds) {
  imp_grid_body<false>(P, S, cap_t, U, max_rounds);
}

__global__ __launch_bounds__(kGridFlowThreads) void k_grid_flow_imp_wide(FlowProblem P,
                                                            GridFlowState S,
                                                            long long cap_t, long long U,
                                                            uint32_t max_rounds) {
  imp_grid_body<true>(P, S, cap_t, U, max_rounds);
}


// The whole-GPU max flow: Gx = G<int, false, true> (clique records, the
// discharge's admissible arcs kept in registers between its two passes)
// for networks of fewer than 2^32 clique-member slots whose unit fits
// 32-bit residuals, G<long long, false> (int64 arrays) when it does not,
// G<int, true> for 2^32 slots or more (clique-id incidence, 32-bit
// residuals).  The last two re-read the arcs in the second pass: their
// lookups leave no registers for the cache.
template <class Gx, bool kCacheArcs>
__device__ __forceinline__ void grid_flow_body(FlowProblem P, GridFlowState S,
                                               long long cap_t, long long U,
                                               uint32_t max_rounds) {
  Gx g;
  flow_setup(g, P, S, U);
  g.gtid = blockIdx.x * blockDim.x + threadIdx.x;
  g.gthreads = gridDim.x * blockDim.x;
  g.gwave = g.gtid / kWave;
  g.gwaves = g.gthreads / kWave;
  g.leader = g.gtid == 0;
  g.nblocks = gridDim.x;
  g.bid = blockIdx.x;
  __shared__ uint32_t app[kMaxFlowWaves + 1];
  g.app = app;
  extern __shared__ long long lds_ex[];
  if constexpr (Gx::kByClique) {
    // inlined here: as a call, G would live in scratch memory
    [[clang::always_inline]] flow_run<kCacheArcs>(g, P, S, cap_t, U, max_rounds, lds_ex,
                                                  g.V <= kGridLdsVertices);
  } else {
    flow_run<kCacheArcs>(g, P, S, cap_t, U, max_rounds, lds_ex, g.V <= kGridLdsVertices);
  }
}

__global__ __launch_bounds__(kGridFlowThreads) void k_grid_flow(FlowProblem P,
                                                   GridFlowState S,
                                                   long long cap_t, long long U,
                                                   uint32_t max_rounds) {
  grid_flow_body<G<long long, false>, false>(P, S, cap_t, U, max_rounds);
}

// 512-thread workgroups: the slot lookups need more registers than the
// 128 a lane of a 1024-thread workgroup may hold.
constexpr int kGridFlowWideThreads = 512;
__global__ __launch_bounds__(kGridFlowWideThreads) void k_grid_flow_wide(FlowProblem P,
                                                        GridFlowState S,
                                                        long long cap_t, long long U,
                                                        uint32_t max_rounds) {
  grid_flow_body<G<int, true>, false>(P, S, cap_t, U, max_rounds);
}

// Many independent problems in one launch; item i owns workgroups
// [first_block, first_block + nblocks) (items ascending by first_block).
// A problem on several workgroups synchronises them through its own
// control block's barrier, so such launches are cooperative.  Control
// blocks are zeroed by the caller; lds_cap = LDS excess slots per workgroup.
__global__ __launch_bounds__(kFlowBatchThreads) void k_flow_batch(
    const FlowBatchItem *__restrict__ items, uint32_t count, uint32_t max_rounds,
    uint32_t lds_cap) {
  uint32_t lo = 0, hi = count - 1;  // last item with first_block <= blockIdx
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (items[mid].first_block <= blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const FlowBatchItem &it = items[lo];
  G<long long, false> g;
  flow_setup(g, it.P, it.S, it.U);
  g.nblocks = it.nblocks;
  g.bid = blockIdx.x - it.first_block;
  g.gtid = g.bid * blockDim.x + threadIdx.x;
  g.gthreads = g.nblocks * blockDim.x;
  g.gwave = g.gtid / kWave;
  g.gwaves = g.gthreads / kWave;
  g.leader = g.gtid == 0;
  __shared__ uint32_t app[kMaxFlowWaves + 1];
  g.app = app;
  extern __shared__ long long lds_ex[];
  flow_run<false>(g, it.P, it.S, it.capT, it.U, max_rounds, lds_ex, g.V <= lds_cap);
}

// Cliques of each batched problem with every member on its min-cut side.
__global__ void k_flow_batch_count(const FlowBatchItem *__restrict__ items,
                                   unsigned long long *out) {
  const FlowBatchItem &it = items[blockIdx.x];
  unsigned long long c = 0;
  for (uint32_t t = threadIdx.x; t < it.P.T; t += blockDim.x) {
    bool all = true;
    for (uint32_t j = 0; j < it.P.k; ++j)
      all &= it.S.reach[it.P.mem[(uint64_t)t * (it.P.mstride ? it.P.mstride : it.P.k) + j] -
                        it.P.vbase] != 0;
    c += all;
  }
  c = wave_sum64(c);
  __shared__ unsigned long long s_c;
  if (threadIdx.x == 0) s_c = 0;
  __syncthreads();
  if (lane() == 0 && c) atomicAdd(&s_c, c);
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = s_c;
}

__global__ void k_grid_ctl_init(GridFlowCtl *ctl, unsigned long long limit) {
  unsigned *bw = reinterpret_cast<unsigned *>(&ctl->bar);
  for (size_t i = 0; i < sizeof(GridBarrier) / sizeof(unsigned); ++i) bw[i] = 0;
  ctl->bar.limit = limit;
  ctl->lpack[0] = ctl->lpack[1] = 0;
  ctl->cqn[0] = ctl->cqn[1] = 0;
  ctl->active = 0;
  ctl->sink = 0;
}

}  // namespace

hipError_t flow_batch(const FlowBatchItem *items, uint32_t count,
                      uint32_t total_blocks, bool cooperative, uint32_t max_rounds,
                      uint32_t max_v, unsigned long long *reach_cliques,
                      hipStream_t s) {
  if (count == 0) return hipSuccess;
  uint32_t lds_cap = max_v < kGridLdsVertices ? max_v : kGridLdsVertices;
  const size_t lds = (size_t)lds_cap * sizeof(long long);
  hipError_t e;
  if (cooperative) {
    void *args[] = {(void *)&items, &count, &max_rounds, &lds_cap};
    e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_flow_batch),
                                   dim3(total_blocks), dim3(kFlowBatchThreads), args,
                                   lds, s);
  } else {
    hipLaunchKernelGGL(k_flow_batch, dim3(total_blocks), dim3(kFlowBatchThreads), lds,
                       s, items, count, max_rounds, lds_cap);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_flow_batch_count, dim3(count), dim3(256), 0, s, items,
                     reach_cliques);
  return hipGetLastError();
}

uint32_t imp_record_words(uint32_t k) { return imp_stride(k); }

// ---- the batched search's flow step as a HIP graph ----
// A graph of three nodes -- copy the step's items to the device, the
// batched max flow (single-workgroup problems, a plain launch), the clique
// count of each min-cut side -- built on the first step and re-aimed at
// every later step's items, sizes and counters through the executable
// graph's node parameters, so a step is one graph launch instead of three
// stream operations.  The items are staged in pinned host memory the graph
// copies from; the device copy is the graph's own buffer.
struct FlowBatchGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t copy = nullptr, flow = nullptr, count = nullptr;
  FlowBatchItem *host = nullptr, *dev = nullptr;
  size_t cap = 0;  // items the two buffers hold
  unsigned long long launches = 0;
};

namespace {
hipKernelNodeParams flow_node(void **args, uint32_t blocks, uint32_t lds_cap) {
  hipKernelNodeParams kp{};
  kp.func = reinterpret_cast<void *>(k_flow_batch);
  kp.gridDim = dim3(blocks);
  kp.blockDim = dim3(kFlowBatchThreads);
  kp.sharedMemBytes = (unsigned)((size_t)lds_cap * sizeof(long long));
  kp.kernelParams = args;
  kp.extra = nullptr;
  return kp;
}
hipKernelNodeParams count_node(void **args, uint32_t count) {
  hipKernelNodeParams kp{};
  kp.func = reinterpret_cast<void *>(k_flow_batch_count);
  kp.gridDim = dim3(count);
  kp.blockDim = dim3(256);
  kp.sharedMemBytes = 0;
  kp.kernelParams = args;
  kp.extra = nullptr;
  return kp;
}
}  // namespace

void flow_batch_graph_free(FlowBatchGraph *g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->dev) (void)hipFree(g->dev);
  if (g->host) (void)hipHostFree(g->host);
  delete g;
}

unsigned long long flow_batch_graph_launches(const FlowBatchGraph *g) {
  return g ? g->launches : 0;
}

hipError_t flow_batch_graph(FlowBatchGraph **pg, const FlowBatchItem *items, uint32_t count,
                            uint32_t total_blocks, uint32_t max_rounds, uint32_t max_v,
                            unsigned long long *reach_cliques, hipStream_t s) {
  if (count == 0) return hipSuccess;
  if (!*pg) *pg = new FlowBatchGraph();
  FlowBatchGraph *g = *pg;
  hipError_t e = hipSuccess;
  if (count > g->cap) {
    // the previous step's copy has run: every step ends in a readback
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    size_t cap = g->cap ? g->cap : 256;
    while (cap < count) cap *= 2;
    if (g->dev) (void)hipFree(g->dev);
    if (g->host) (void)hipHostFree(g->host);
    g->dev = nullptr;
    g->host = nullptr;
    g->cap = 0;
    if ((e = hipMalloc(&g->dev, cap * sizeof(FlowBatchItem))) != hipSuccess) return e;
    if ((e = hipHostMalloc(&g->host, cap * sizeof(FlowBatchItem), hipHostMallocDefault)) !=
        hipSuccess)
      return e;
    g->cap = cap;
  } else if ((e = hipStreamSynchronize(s)) != hipSuccess) {
    return e;  // the staging buffer may still feed the previous copy
  }
  memcpy(g->host, items, count * sizeof(FlowBatchItem));
  uint32_t lds_cap = max_v < kGridLdsVertices ? max_v : kGridLdsVertices;
  const FlowBatchItem *dev = g->dev;
  void *fargs[] = {(void *)&dev, &count, &max_rounds, &lds_cap};
  void *cargs[] = {(void *)&dev, &reach_cliques};
  const hipKernelNodeParams fp = flow_node(fargs, total_blocks, lds_cap);
  const hipKernelNodeParams cp = count_node(cargs, count);
  const size_t bytes = count * sizeof(FlowBatchItem);
  if (!g->exec) {
    if ((e = hipGraphCreate(&g->graph, 0)) != hipSuccess) return e;
    if ((e = hipGraphAddMemcpyNode1D(&g->copy, g->graph, nullptr, 0, g->dev, g->host, bytes,
                                     hipMemcpyHostToDevice)) != hipSuccess)
      return e;
    if ((e = hipGraphAddKernelNode(&g->flow, g->graph, &g->copy, 1, &fp)) != hipSuccess)
      return e;
    if ((e = hipGraphAddKernelNode(&g->count, g->graph, &g->flow, 1, &cp)) != hipSuccess)
      return e;
    if ((e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess)
      return e;
  } else {
    if ((e = hipGraphExecMemcpyNodeSetParams1D(g->exec, g->copy, g->dev, g->host, bytes,
                                               hipMemcpyHostToDevice)) != hipSuccess)
      return e;
    if ((e = hipGraphExecKernelNodeSetParams(g->exec, g->flow, &fp)) != hipSuccess) return e;
    if ((e = hipGraphExecKernelNodeSetParams(g->exec, g->count, &cp)) != hipSuccess) return e;
  }
  if ((e = hipGraphLaunch(g->exec, s)) != hipSuccess) return e;
  ++g->launches;
  return hipSuccess;
}

unsigned flow_batch_capacity(uint32_t max_v, int device) {
  const uint32_t lds_cap = max_v < kGridLdsVertices ? max_v : kGridLdsVertices;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void *>(k_flow_batch), kFlowBatchThreads,
          (size_t)lds_cap * sizeof(long long)) != hipSuccess ||
      per_cu < 1)
    return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) !=
      hipSuccess)
    return 0;
  return (unsigned)cus;
}

unsigned flow_grid_blocks(const FlowProblem &P, int device, unsigned blocks, bool wide,
                          bool imp) {
  int per_cu = 0, cus = 0;
  const size_t lds = P.V <= kGridLdsVertices
                         ? (size_t)P.V * (imp ? 12 : sizeof(long long))
                         : 0;
  const void *fn = imp   ? (P.by_clique ? reinterpret_cast<const void *>(k_grid_flow_imp_wide)
                                        : reinterpret_cast<const void *>(k_grid_flow_imp))
                   : wide ? reinterpret_cast<const void *>(k_grid_flow_wide)
                          : reinterpret_cast<const void *>(k_grid_flow);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, fn, !imp && wide ? kGridFlowWideThreads : kGridFlowThreads, lds) !=
              hipSuccess ||
      per_cu < 1)
    return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) !=
      hipSuccess)
    return 0;
  // One 1024-thread workgroup per 64k (clique, member) arcs, up to one per
  // CU: each of the four grid barriers of a round costs a fence pair per
  // workgroup, so small networks run on one workgroup, while on large ones
  // the vertex discharges dominate and scale with the CUs.
  const uint64_t cap = cus > 1 ? (uint64_t)cus : 1;
  const uint64_t arcs = (uint64_t)P.T * P.k;
  uint64_t want = (arcs + 65535) / 65536;
  if (want > cap) want = cap;
  if (want < 1) want = 1;
  if (blocks >= 1 && blocks <= cap) want = blocks;
  return (unsigned)want;
}

hipError_t grid_max_flow(const FlowProblem &P, const GridFlowState &S,
                         long long cap_t, long long unit, uint32_t max_rounds,
                         unsigned *grid, hipStream_t s) {
  if (*grid < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_grid_ctl_init, dim3(1), dim3(1), 0, s, S.ctl, barrier_limit_ticks());
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t lds = P.V <= kGridLdsVertices ? (size_t)P.V * sizeof(long long) : 0;
  FlowProblem p = P;
  GridFlowState st = S;
  void *args[] = {&p, &st, &cap_t, &unit, &max_rounds};
  if (S.rec32) {
    if (P.k < 2 || P.k > kImpMaxK) return hipErrorInvalidValue;
    const size_t ilds = P.V <= kGridLdsVertices ? (size_t)P.V * 12 : 0;
    const void *ifn = P.by_clique ? reinterpret_cast<const void *>(k_grid_flow_imp_wide)
                                  : reinterpret_cast<const void *>(k_grid_flow_imp);
    return launch_grid_shrinking(ifn, grid, kGridFlowThreads, args, ilds, s);
  }
  const void *fn = S.x32   ? reinterpret_cast<const void *>(k_grid_flow_wide)
                           : reinterpret_cast<const void *>(k_grid_flow);
  return launch_grid_shrinking(fn, grid, S.x32 ? kGridFlowWideThreads : kGridFlowThreads,
                               args, lds, s);
}

}  // namespace cds
