// Host-callable launchers of the HIP kernels (implemented in *.hip).
// Every launcher enqueues work on `s` and returns the launch status; none
// synchronises unless its comment says so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/cds.h"

namespace cds {

constexpr uint32_t kAliveHost = 0xFFFFFFFFu;  // same value as kAlive (dev.h)

// Stream-ordered device allocations (hipMallocAsync on the device pool).
// Blocks of kArenaCacheMin bytes or more are kept on release, up to
// cache_cap bytes, and handed out again to a request of at least half
// their size: freeing and re-mapping tens of GB per solve cost seconds
// (hipFreeAsync of a 13 GB block blocked the host for 1.2-1.8 s).  All
// use is on the arena's one stream, so a reused block is stream-ordered.
// A failed allocation empties the cache and retries once.
constexpr size_t kArenaCacheMin = 64ull << 20;
class DeviceArena {
 public:
  explicit DeviceArena(hipStream_t s, size_t cache_cap = 0)
      : stream_(s), cache_cap_(cache_cap) {}
  ~DeviceArena();
  hipError_t alloc(void **p, size_t bytes);
  void release(void *p);
  void trim();  // frees every cached block
  size_t live_bytes() const { return live_; }
  size_t peak_bytes() const { return peak_; }
  size_t cached_bytes() const { return cached_; }
  hipStream_t stream() const { return stream_; }

 private:
  struct Block {
    void *p;
    size_t size;  // bytes requested
    size_t real;  // bytes of the underlying allocation
  };
  hipStream_t stream_;
  size_t cache_cap_;
  size_t live_ = 0, peak_ = 0, cached_ = 0;
  std::vector<Block> live_blocks_, cache_;
};

// ---- scan / reduce (scan.hip) ----
// out[0..n]: exclusive prefix sums of in[0..n), out[n] = total.
hipError_t scan_u32(const uint32_t *in, uint64_t *out, size_t n,
                    DeviceArena &a);
hipError_t scan_u64(const uint64_t *in, uint64_t *out, size_t n,
                    DeviceArena &a);
// Same with 32-bit output (caller guarantees total < 2^32).
hipError_t scan_u32_32(const uint32_t *in, uint32_t *out, size_t n,
                       DeviceArena &a);

// The same scans with their block sums in a caller-given workspace of
// scan_ws_words(n) elements of the output type (no arena, no allocation).
size_t scan_ws_words(size_t n);
hipError_t scan_u32_ws(const uint32_t *in, uint64_t *out, size_t n, uint64_t *ws,
                       size_t ws_words, hipStream_t s);
hipError_t scan_u32_32_ws(const uint32_t *in, uint32_t *out, size_t n, uint32_t *ws,
                          size_t ws_words, hipStream_t s);

// ---- zero-allocation scratch (noalloc.hip) ----
// Bytes of the static device array behind cds_count_cliques_noalloc.
constexpr size_t kNoallocScratchBytes = CDS_NOALLOC_SCRATCH_BYTES;
// Records this device's address of the array (hipGetSymbolAddress); call
// once per device with that device current.
hipError_t noalloc_attach(int device);
// The address noalloc_attach recorded for `device` (null before it ran).
void *noalloc_scratch(int device);
// a[0..n) = v, by a kernel.
hipError_t fill_u64(uint64_t *a, size_t n, uint64_t v, hipStream_t s);
// *out = sum of in[0..n) (out is cleared by a kernel first).
hipError_t sum_u32(const uint32_t *in, size_t n, unsigned long long *out, hipStream_t s);

// ---- core pruning and the dense-core peel (prune.hip) ----
// Vertices of the largest core peeled on adjacency bitsets.
constexpr uint32_t kDenseMaxVertices = 16384;
// Induced subgraph on the vertices with keep[v] != 0, in two calls:
// fill = false computes newid (exclusive scan of keep) and cnt[newid] (the
// induced degrees); the caller scans cnt into off2; fill = true writes adj2
// (new ids, ascending) and sel[newid] = old id.
hipError_t induced_subgraph(uint32_t n, const uint32_t *off, const uint32_t *adj,
                            const uint8_t *keep, uint32_t *newid, uint32_t *cnt,
                            uint32_t *off2, uint32_t *adj2, uint32_t *sel, bool fill,
                            DeviceArena &a);
// Clique count (*total) and clique degrees (cd, ns words) of a graph of
// ns <= kBitsCountMax vertices, k = 3 or 4, from its adjacency bitsets
// (bits: ns * ceil(ns / 64) words; src: arcs words of workspace).
constexpr uint32_t kBitsCountMax = 4096;
hipError_t bits_clique_count(uint32_t k, uint32_t ns, const uint32_t *off,
                             const uint32_t *adj, uint32_t arcs, unsigned long long *bits,
                             uint32_t *src, unsigned long long *cd,
                             unsigned long long *total, hipStream_t s);
// In-place drops of the pruned path, on the current induced subgraph with
// a state byte per vertex (1 in S, 2 dropped this round, 0 dropped
// before).  prune_select moves the vertices in S with cd < tau to state 2
// and lists them in D (*nd, cleared by the caller), stats[0] += their
// clique degrees, stats[1] = max(stats[1], their degree).
hipError_t prune_select(uint32_t ns, const uint64_t *cd, uint64_t tau, const uint32_t *off,
                        uint8_t *state, uint32_t *D, uint32_t *nd,
                        unsigned long long *stats, hipStream_t s);
// Largest degree of a dropped vertex prune_unlink takes (its neighbour
// lists are staged in LDS, 4 per workgroup).
constexpr uint32_t kUnlinkMaxDegree = 2048;
// Removes the k-cliques (k = 3, 4) that contain a vertex of D from the
// clique degrees of their members in S (each clique once, from its
// smallest dropped member); *removed += how many.  deg (nd words) and
// poff (nd + 1) are workspace.
hipError_t prune_unlink(uint32_t k, uint32_t nd, const uint32_t *D, uint32_t max_deg,
                        const uint32_t *off, const uint32_t *adj, const uint8_t *state,
                        unsigned long long *cd, unsigned long long *removed, uint32_t *deg,
                        uint32_t *poff, DeviceArena &a);
// state[D[j]] = 0 and keep[sel[D[j]]] = 0.
hipError_t prune_finish(uint32_t nd, const uint32_t *D, const uint32_t *sel, uint8_t *state,
                        uint8_t *keep, hipStream_t s);
// The sequential peel of a graph of ns <= kDenseMaxVertices vertices with
// clique degrees cd: rv[i], rval[i] = the vertex removed i-th and its clique
// count then.  bits: ns * ceil(ns / 64) words; alive: ceil(ns / 64); cur: ns.
hipError_t dense_peel(uint32_t ns, uint32_t k, const uint32_t *off, const uint32_t *adj,
                      const uint64_t *cd, unsigned long long *bits, unsigned long long *alive,
                      unsigned long long *cur, uint32_t *step_v, uint32_t *rv,
                      unsigned long long *rval, hipStream_t s);

// ---- DAG (dag.hip) ----
// Degree-ordered orientation: u -> v iff (deg u, u) < (deg v, v).
hipError_t dag_count(uint32_t n, const uint32_t *off, const uint32_t *adj,
                     uint32_t *dcnt, hipStream_t s);
hipError_t dag_fill(uint32_t n, const uint32_t *off, const uint32_t *adj,
                    const uint32_t *doff, uint32_t *dadj, hipStream_t s);
// out[0] = max out-degree, out[1] = the listing work sum_v indeg(v) d+(v),
// out[2] = sum over DAG arcs (u, v) of min(d+(u), d+(v)), a bound on the
// triangle count (three device words, cleared here).
hipError_t dag_stats(uint32_t n, const uint32_t *off, const uint32_t *dcnt,
                     const uint32_t *doff, const uint32_t *dadj, unsigned long long *out,
                     hipStream_t s);
// max over dcnt (result written to *out, device)
hipError_t reduce_max_u32(const uint32_t *in, size_t n, uint32_t *out,
                          hipStream_t s);

// ---- listing (list.hip) ----
// Work items: (u, first position in N+(u)); each covers up to `chunk`
// out-neighbours of u.  The tile shape (workgroup lanes, chunk) is one of
// (256, 64) [default], (256, 32), (512, 64), (512, 32), (128, 32).
struct ListShape {
  uint32_t threads = 256;
  uint32_t chunk = 64;
};
// N+(u) is staged in LDS up to this many entries (k = 3: list + counters,
// 64000 B; larger lists are searched in global memory).  k = 4 needs three
// LDS arrays and requires every out-degree to fit.
constexpr uint32_t kListLdsCap = 8000;
constexpr uint32_t kList4LdsCap = 4000;
hipError_t list_item_counts(uint32_t n, uint32_t chunk, const uint32_t *doff,
                            uint32_t *items_per_vertex, hipStream_t s);
hipError_t list_item_fill(uint32_t n, uint32_t chunk, const uint32_t *doff,
                          const uint64_t *item_off, uint2 *items,
                          hipStream_t s);
// k = 3 or 4.  Count mode (out == nullptr): item_count[i] = cliques found by
// item i, clique_degree += per-vertex contributions.  Write mode: cliques
// written AoS (k ids each) at out + k * item_base[i].  hipErrorInvalidValue
// for a shape not listed above.
// Staging (k = 3, count mode; st.buf null: none): each item reserves its
// worst case, sum over its arcs (u, v) of min(d+(u), d+(v)) triangles, at
// *st.cursor in st.buf (st.cap triangles) and writes the triangles it finds
// there (3 ids each, item_tmp[i] = the reservation, ~0 and *st.over = 1 if
// it did not fit), so the cliques can be moved into place by
// list_unstage instead of being listed again.
struct ListStage {
  uint32_t *buf = nullptr;
  unsigned long long *cursor = nullptr;
  uint64_t cap = 0;
  uint64_t *item_tmp = nullptr;
  uint32_t *over = nullptr;
};
hipError_t list_cliques(uint32_t k, ListShape shape, uint64_t num_items,
                        const uint2 *items, const uint32_t *doff,
                        const uint32_t *dadj, uint32_t max_out_degree,
                        uint32_t *item_count, uint64_t *clique_degree,
                        const uint64_t *item_base, uint32_t *out, hipStream_t s,
                        ListStage st = ListStage());
// The staged triangles of every item, item_count[i] of them at
// item_tmp[i], to their place item_base[i] in out (3 ids each).
hipError_t list_unstage(uint64_t num_items, const uint64_t *item_tmp,
                        const uint32_t *item_count, const uint64_t *item_base,
                        const uint32_t *stage, uint32_t *out, hipStream_t s);
// k = 2: every DAG arc is a clique.
hipError_t list_edges(uint32_t n, const uint32_t *doff, const uint32_t *dadj,
                      uint64_t *clique_degree, uint32_t *out, hipStream_t s);

// ---- peel (peel.hip) ----
hipError_t build_incidence(uint64_t T, uint32_t k, const uint32_t *cliques,
                           const uint64_t *ioff, uint32_t *icur,
                           uint32_t *inc, hipStream_t s);
hipError_t peel_init(uint32_t n, uint64_t T, const uint64_t *cd,
                     uint32_t *deg, uint32_t *core, uint32_t *kill,
                     hipStream_t s);
// *out = min deg over vertices with core == kAlive (kAlive if none).
hipError_t peel_min_alive(uint32_t n, const uint32_t *deg,
                          const uint32_t *core, uint32_t *out, hipStream_t s);
// Appends alive vertices with deg <= level to frontier, sets core = level.
hipError_t peel_frontier(uint32_t n, uint32_t level, const uint32_t *deg,
                         uint32_t *core, uint32_t *frontier,
                         uint32_t *count, hipStream_t s);
// Processes a frontier: kills alive cliques, decrements members, appends
// vertices that drop to `level` to next; counts killed cliques.
hipError_t peel_process(uint32_t k, uint32_t level, const uint32_t *frontier,
                        const uint32_t *fcount, uint32_t fmax,
                        const uint64_t *ioff, const uint32_t *inc,
                        const uint32_t *cliques, uint32_t *deg,
                        uint32_t *core, uint32_t *kill, uint32_t *next,
                        uint32_t *ncount, unsigned long long *killed,
                        hipStream_t s);
// Grid barrier state of a cooperative launch: a generation word and
// arrival counters (one per group of workgroups blockIdx % 8, plus one for
// the groups), each in its own 128-byte line.  Counters only grow; zero
// them before the launch.  A wait lasts at most `limit` wall-clock ticks
// (0: kBarrierDefaultTicks); a workgroup whose wait runs out sets `abort`
// and ends, and so does every workgroup that reaches a barrier after that.
// The host reads `abort` after the launch and fails the call instead of
// hanging when workgroups are not co-resident.
struct GridBarrier {
  unsigned int gen, pad0[31];
  unsigned int top, pad1[31];
  unsigned int sub[8 * 32];
  unsigned int abort, pad2[31];
  unsigned long long limit, pad3[15];
};

// Whole peel in one cooperative launch (see peel.hip).  Control block:
struct PeelCtl {
  GridBarrier bar;
  unsigned int gmin[2], pad0[30];
  unsigned int alive_n[2], pad1[30];
  unsigned long long fpack[3], pad2[13];  // per frontier: count<<40 | chunks
  unsigned int nlevels, status;
  unsigned long long removed;
  unsigned long long rounds;
  // leader-thread time split (wall_clock64 ticks): barriers, alive scans,
  // frontier chunk processing, LDS flushes
  unsigned long long t_bar, t_scan, t_proc, t_flush;
};
// Scratch: u32 6n (two alive lists, three frontiers, alive positions),
// u64 3n (chunk starts).
// Level tables (capacity n, zeroed here), one entry per step: level (the
// core value given to the step's vertices and cliques), lv_min (the
// minimum degree that opened it; == level for an exact level), removed,
// killed.  cshift = 0 peels exact levels; cshift > 0 coarse steps covering
// [m, min(m + (m >> cshift), max_level)] (max_level bounds the values given
// out when the vertices' core numbers are known not to exceed it).
// sdeg[v] (n entries): v's degree at the start of the step removing it.
// alive0 (na0 vertices, or null for all n): the vertices alive at the
// start; every other vertex must have core != kAlive and every clique
// containing one must be dead or never reached.
// kill[] (T words) is not touched: once no launch is left (and no vertex
// parked), peel_derive_kill sets every clique's word to the level that
// killed it, the smallest core value among its members.  stamp[] (n words) holds each
// vertex's sub-round stamp; sbase is the launch's first sub-round: 1 for the
// first launch on a graph, above every stamp of earlier launches (their
// sub-round counts + 1) for a re-peel.
// Requires n < 2^24.  Synchronises nothing; read ctl after the stream.
hipError_t peel_derive_kill(uint64_t T, uint32_t k, const uint32_t *cliques,
                            const uint32_t *core, uint32_t *kill, hipStream_t s);
hipError_t peel_persistent(uint32_t n, uint32_t k, uint64_t T, const uint64_t *ioff,
                           const uint32_t *inc, const uint32_t *cliques,
                           uint32_t *deg, uint32_t *core, uint32_t *kill,
                           uint32_t *stamp, uint32_t sbase, uint32_t *scratch32,
                           uint64_t *scratch64,
                           PeelCtl *ctl, uint32_t *lv_level, uint32_t *lv_min,
                           unsigned long long *lv_removed,
                           unsigned long long *lv_killed, uint32_t *sdeg,
                           const uint32_t *alive0, uint32_t na0, uint32_t cshift,
                           uint32_t max_level, unsigned *grid, hipStream_t s);
// Prepares the exact re-peel of one coarse step whose vertices and cliques
// carry the value `level`: those vertices become alive (core = kAlive,
// degree = sdeg, listed in alive[], count in *count), vertices of later
// steps (core > level) are parked (core = kParked, their core value kept
// in deg) and, when kill is given, the step's cliques become alive again
// (kill null: the kill words are derived afresh after the last launch).
constexpr uint32_t kParked = 0xFFFFFFFEu;
hipError_t refine_prepare(uint32_t n, uint64_t T, uint32_t level,
                          uint32_t *core, uint32_t *deg, const uint32_t *sdeg,
                          uint32_t *kill, uint32_t *alive, uint32_t *count,
                          hipStream_t s);
// Undoes the parking after the re-peel.
hipError_t refine_restore(uint32_t n, uint32_t *core, const uint32_t *deg,
                          hipStream_t s);
// Workgroups for peel_persistent on this graph (0: cannot launch).
// force >= 1: that many (at most the cap) instead of the sizing rule.
unsigned peel_grid_blocks(uint32_t n, uint64_t T, uint32_t k, int device,
                          unsigned force = 0);

// Exact sequential replay of one peel phase (level `level`); see peel.hip.
struct ReplayOut {
  unsigned long long best_num, best_den;  // max intermediate density
  uint32_t steps, found;
  uint32_t stopped, pad;  // 1: ended early, the rows left cannot exceed kp
};
// One phase of replay_phases_multi: level, its vertices at positions
// [base, base + M) of the shared position arrays, and the state C_level.
struct ReplayPhase {
  uint32_t level, base, M, pad;
  unsigned long long T_start, N_start;
};
// Replays many phases at once, one workgroup each (phases of different
// graphs, or of different levels: disjoint vertices and cliques).  pv holds
// npos positions (each phase's vertices ascending), posmap maps a vertex to
// its position, sdeg gives the level-start degrees; pdeg/palive/pioff are
// per position (pioff npos+1); pinc >= k * sum(pdeg) words.  out[p] per phase.
hipError_t replay_phases_multi(uint32_t k, uint32_t nphases,
                               const ReplayPhase *phases, uint32_t npos,
                               const uint32_t *pv, const uint32_t *sdeg,
                               uint32_t *pdeg, const uint32_t *posmap,
                               uint8_t *palive, const uint32_t *core,
                               const uint32_t *kill, const uint64_t *ioff,
                               const uint32_t *inc, uint64_t *pioff,
                               uint32_t *pinc, uint32_t *rdead, ReplayOut *out,
                               DeviceArena &arena, hipStream_t s);
// Replays level `level` (M vertices) in the oracle's removal order.
// One workgroup (grid = 0): the phase-clique incidence is first compacted
// into pinc/pioff (pinc >= k * k * cliques killed at the level, pioff M+1).
// Whole GPU (grid workgroups, M <= kReplayGridMaxVertices): works on the
// full incidence lists; partial holds grid * M words, bar one GridBarrier,
// and sdeg (if not null) the degrees recorded by peel_persistent.
// `prepared` (if not null) is recorded before the replay kernel.
// Whole-GPU replay of one phase given its vertex list pv[0, M) (ascending),
// for a phase of one graph inside a disjoint union, where the level value
// does not identify the phase: posmap entries of pv's vertices are set to
// their positions; sdeg holds the degrees at the level's start.
hipError_t replay_phase_list(uint32_t k, uint32_t level, const uint32_t *pv, uint32_t M,
                             const uint32_t *core, const uint32_t *kill,
                             const uint64_t *ioff, const uint32_t *inc,
                             uint64_t T_start, uint64_t N_start, uint32_t *pdeg,
                             uint32_t *posmap, uint8_t *palive, uint32_t *rdead,
                             uint32_t *pcount, ReplayOut *out, uint32_t *partial,
                             GridBarrier *bar, unsigned *grid, const uint32_t *sdeg,
                             uint64_t kp, uint64_t nstar, hipStream_t s);
// Replays level `level` (M phase vertices) on adjacency bitsets, for k = 2
// or 3 when the N vertices alive at the level's start (core >= level) are
// at most kReplayBitsMaxVertices: one workgroup, no incidence lists, no grid
// barriers.  doff/dadj: the degree-ordered DAG; sdeg: the degrees at the
// level's start; av and posmap hold n entries, bits N * ceil(N/64) words.  The
// kernel checks that N vertices are alive and otherwise reports 0 steps.
// seq (2 M words, or null): the j-th removal's vertex and its clique count
// then, at seq[2j] and seq[2j+1].
constexpr uint32_t kReplayBitsMaxVertices = 4096;
hipError_t replay_phase_bits(uint32_t k, uint32_t level, uint32_t n, const uint32_t *doff,
                             const uint32_t *dadj, const uint32_t *core, const uint32_t *sdeg,
                             uint64_t T_start, uint32_t N, uint32_t M, uint32_t *av,
                             uint32_t *posmap, uint32_t *pcount, unsigned long long *bits,
                             ReplayOut *out, hipEvent_t prepared, uint64_t kp, uint64_t nstar,
                             hipStream_t s, uint32_t *seq = nullptr);
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
                        hipStream_t s);
// Early end of a whole-GPU replay (nstar = 0: none).  Once T cliques
// remain on N vertices, every later row holds at most T cliques and at most
// C(N'-1, k-1)/k density on its N' < N vertices; nstar is the least N' with
// C(N'-1, k-1) > k * kp, so the replay may stop when N <= nstar or
// T <= kp * nstar: no later row can exceed kp.
uint64_t replay_nstar(uint32_t k, uint64_t kp);
constexpr uint32_t kReplayGridMaxVertices = 16384;
// Workgroups of the whole-GPU replay (half the CUs), 0 if unavailable.
// Workgroups for the whole-GPU replay of a phase whose removals touch about
// entries_per_step incidence entries each (CDS_REPLAY_BLOCKS overrides).
unsigned replay_grid_blocks(int device, uint64_t entries_per_step, unsigned force = 0);

// The densest core's vertices (core >= kp): cidx[v] = their numbering
// (0xFFFFFFFF off the core), cv[slot] = v for the first cap slots, *count =
// how many (the order of the slots is not fixed).
hipError_t core_list(uint32_t n, const uint32_t *core, uint32_t kp, uint32_t *cidx,
                     uint32_t *cv, uint32_t cap, uint32_t *count, hipStream_t s);
// Directed edges among core vertices: over every vertex (core >= kp), or
// over the n vertices listed in cv when it is given.
hipError_t core_edge_count(uint32_t n, const uint32_t *off,
                           const uint32_t *adj, const uint32_t *core,
                           uint32_t kp, unsigned long long *out,
                           hipStream_t s, const uint32_t *cv = nullptr);
// *out = cliques with kill level >= kp (the core's cliques; with kill
// null, those whose k members all have core >= kp); with cc their ids are
// also listed there (order not fixed).
hipError_t core_cliques(uint64_t T, uint32_t k, const uint32_t *cliques, const uint32_t *core,
                        const uint32_t *kill, uint32_t kp, uint32_t *cc,
                        unsigned long long *out, hipStream_t s);
// The same list (cc, *out) from the cn core vertices cv only: a core
// clique's first listed member is a core vertex, and each vertex's cliques
// are one range of the clique array (k = 2: DAG arc order, doff; k >= 3:
// listing item order, item_off / item_base).
hipError_t core_cliques_ranged(uint32_t cn, const uint32_t *cv, uint32_t k,
                               const uint32_t *doff, const uint64_t *item_off,
                               const uint64_t *item_base, const uint32_t *cliques,
                               const uint32_t *core, uint32_t kp, uint32_t *cc,
                               unsigned long long *out, DeviceArena &a);
// dst[idx[i]] = a[i], i < m.
hipError_t scatter_u32(uint32_t m, const uint32_t *idx, const uint32_t *a, uint32_t *dst,
                       hipStream_t s);
// dst[i] = src[idx[i]] and da[idx[i]] = a[i], db[idx[i]] = b[i], i < m.
hipError_t gather_u32(uint32_t m, const uint32_t *idx, const uint32_t *src, uint32_t *dst,
                      hipStream_t s);
hipError_t scatter_u32x2(uint32_t m, const uint32_t *idx, const uint32_t *a, const uint32_t *b,
                         uint32_t *da, uint32_t *db, hipStream_t s);
hipError_t uf_init(uint32_t n, uint32_t *parent, hipStream_t s);
// Unions the endpoints of the marked (clique-covered) edges of core vertices
// (every vertex with core >= kp, or the n listed in cv).
hipError_t uf_union_edges(uint32_t n, const uint32_t *off, const uint32_t *adj,
                          const uint8_t *mark, const uint32_t *core,
                          uint32_t kp, uint32_t *parent, hipStream_t s,
                          const uint32_t *cv = nullptr);
hipError_t uf_flatten(uint32_t n, uint32_t *parent, hipStream_t s);
// Same marks for a densest core of cn <= kDenseCoreMax vertices (cv, cidx
// from core_list) and its Tc cliques listed in cc, through a cn x cn matrix
// of edge positions (cmat, cn * cn words of scratch); cc null: all T
// cliques are core cliques.
constexpr uint32_t kDenseCoreMax = 4096;
hipError_t prune_mark_dense(uint64_t Tc, uint32_t k, const uint32_t *cliques,
                            const uint32_t *cc, const uint32_t *cv, uint32_t cn,
                            const uint32_t *cidx, const uint32_t *off, const uint32_t *adj,
                            uint32_t *cmat, uint8_t *mark, hipStream_t s);
// The same marks for k = 2, 3 or 4 from the core's adjacency alone: bits is
// cn * ceil(cn / 64) words of scratch for the core's bitset rows.
hipError_t prune_mark_bits(uint32_t k, const uint32_t *cv, uint32_t cn, const uint32_t *cidx,
                           const uint32_t *off, const uint32_t *adj,
                           unsigned long long *bits, uint8_t *mark, hipStream_t s);
// Marks every edge of a core clique: the cliques with kill >= kp among T,
// or the T listed in cc when it is given, or all T when both are null.
hipError_t prune_mark(uint64_t T, uint32_t k, const uint32_t *cliques,
                      const uint32_t *kill, uint32_t kp, const uint32_t *off,
                      const uint32_t *adj, uint8_t *mark, hipStream_t s,
                      const uint32_t *cc = nullptr);
hipError_t count_nonzero_u8(size_t n, const uint8_t *a,
                            unsigned long long *out, hipStream_t s);
// Per-component clique buckets of ncomp components (reservations summed
// per workgroup in LDS when ncomp is small, else wave-aggregated):
//  comp_clique_totals  ccount[comp] += 1 per core clique
//  comp_cliques_build  fmem, and cslot[c] = the flow index of core clique c
//                      when cslot is given
// The core cliques are those with kill >= kp among T, or the T listed in cc
// when it is given, or all T when both are null.
hipError_t comp_clique_totals(uint64_t T, uint32_t k, const uint32_t *cliques,
                              const uint32_t *kill, uint32_t kp,
                              const uint32_t *comp_of, uint32_t ncomp,
                              uint32_t *ccount, hipStream_t s,
                              const uint32_t *cc = nullptr);
hipError_t comp_cliques_build(uint64_t T, uint32_t k, const uint32_t *cliques,
                              const uint32_t *kill, uint32_t kp,
                              const uint32_t *comp_of, const uint32_t *fv_of,
                              uint32_t ncomp, const uint64_t *coff,
                              uint32_t *ccur, uint32_t *fmem, uint32_t *cslot,
                              hipStream_t s, uint32_t ms = 0,
                              const uint32_t *cc = nullptr);
// Flow-network incidence from the core's clique bucket fmem (Tc cliques,
// k flow-vertex ids each, nv flow vertices): fcd[v] = number of pairs of
// vertex v (zeroed here); then, with fioff the exclusive prefix sum of fcd,
// finc[fioff[v] ..) = the pair indices s*k+j whose vertex is v, or with
// by_clique their cliques s (for networks of 2^32 pairs or more; fcur is
// nv words of scratch, zeroed here).
// ms: words per clique in fmem (0: k; clique records hold more).
// fcd[fv_of[cv[i]]] = cd[cv[i]] for the cn core vertices: the flow
// incidence degrees when every clique of the graph is a core clique.
hipError_t core_flow_degrees(uint32_t cn, const uint32_t *cv, const uint64_t *cd,
                             const uint32_t *fv_of, uint32_t *fcd, hipStream_t s);
hipError_t flow_incidence_degrees(uint64_t Tc, uint32_t k, const uint32_t *fmem,
                                  uint32_t nv, uint32_t *fcd, hipStream_t s,
                                  uint32_t ms = 0);
hipError_t flow_incidence_fill(uint64_t Tc, uint32_t k, const uint32_t *fmem,
                               const uint64_t *fioff, uint32_t nv, uint32_t *fcur,
                               uint32_t *finc, bool by_clique, hipStream_t s,
                               uint32_t ms = 0);

// ---- batched solves: the disjoint union of many graphs (batch.hip) ----
struct UnionPart {
  const uint32_t *off, *adj;  // the graph's own CSR (device)
  uint32_t n, vbase, abase, pad;  // its vertices / first vertex / first slot
};
// Union CSR (off n_total+1, adj a_total) and gid[v] = graph of v.
hipError_t union_build(uint32_t B, const UnionPart *parts, uint32_t n_total,
                       uint32_t a_total, uint32_t *off, uint32_t *adj,
                       uint32_t *gid, hipStream_t s);
// Per graph: sum and max of clique degrees (outputs zeroed by the caller).
hipError_t graph_clique_degrees(uint32_t n, const uint32_t *gid, const uint64_t *cd,
                                unsigned long long *sum, unsigned long long *mx,
                                hipStream_t s);
// Level runs: vertices sorted by (graph, core) (skey/sval, n each; key/val
// scratch n each); ukey[r], counts[r] for each distinct (graph, core) pair
// in ascending order, *nruns (device) their number.
hipError_t level_runs(uint32_t n, const uint32_t *gid, const uint32_t *core,
                      uint64_t *key, uint32_t *val, uint64_t *skey, uint32_t *sval,
                      uint64_t *ukey, uint32_t *counts, uint32_t *nruns,
                      DeviceArena &arena, hipStream_t s);
// killed[r] = cliques whose (graph, kill level) is run r.
hipError_t run_killed(uint64_t T, uint32_t k, const uint32_t *cl,
                      const uint32_t *kill, const uint32_t *gid,
                      const uint64_t *ukeys, uint32_t nruns, uint32_t *killed,
                      hipStream_t s);
// posmap[sval[i]] = i.
hipError_t positions_of(uint32_t n, const uint32_t *sval, uint32_t *posmap,
                        hipStream_t s);
// core01[v] = core[v] >= kp[graph of v]; kill01[c] likewise with the graph
// of the clique; both are then compared against a threshold of 1.
hipError_t densest_flags(uint32_t n, uint64_t T, uint32_t k, const uint32_t *cl,
                         const uint32_t *gid, const uint32_t *core,
                         const uint32_t *kill, const uint32_t *kp,
                         uint32_t *core01, uint32_t *kill01, hipStream_t s);
// Per graph: adjacency entries between flagged vertices, and marked entries
// (mark may be null).  Outputs zeroed by the caller.
hipError_t graph_edge_counts(uint32_t n, const uint32_t *gid, const uint32_t *off,
                             const uint32_t *adj, const uint32_t *flag,
                             const uint8_t *mark, unsigned long long *core_e,
                             unsigned long long *kept_e, hipStream_t s);
// Per graph: cliques whose members all have mask != 0 (out zeroed by caller).
hipError_t graph_cliques_in_mask(uint64_t T, uint32_t k, const uint32_t *cl,
                                 const uint32_t *gid, const uint8_t *mask,
                                 uint32_t *out, hipStream_t s);

// ---- max flow (flow_grid.hip) and set accounting (flow.hip) ----
struct FlowProblem {
  uint32_t k, V, T;       // vertices, cliques of this component
  uint32_t vbase;         // flow-vertex id of local vertex 0
  uint64_t sbase;         // slot id of local slot 0 (first clique * k)
  const uint32_t *mem;    // member flow-vertex ids (subtract vbase), mstride
                          // words per clique, the first k of them
  const uint64_t *ioff;   // V+1 offsets into inc, for local vertices
  const uint32_t *inc;    // slot ids (subtract sbase) of each vertex
  const uint32_t *cd;     // V clique degrees
  // inc holds clique ids (from sbase / k) instead of slot ids
  uint32_t by_clique = 0;
  uint32_t mstride = 0;   // words per clique in mem (0: k)
};

// Exact max flow of one component in one cooperative launch on as many
// workgroups as its size needs.
struct GridFlowCtl {
  GridBarrier bar;
  unsigned long long lpack[2];  // vertex work lists: count<<40 | chunks
  unsigned int cqn[2];          // clique BFS frontier sizes
  unsigned long long active;
  unsigned long long sink;      // flow into the sink (scaled)
  unsigned int rounds, limit, relabels, pad;
  // work done, summed over rounds: clique work-list entries visited in
  // phase 1, incidence chunks posted by active vertices in phase 2
  unsigned long long clique_visits, chunk_visits;
};
// Arcs per vertex incidence chunk of the max-flow kernels.
constexpr uint32_t kFlowChunkArcs = 512;
struct GridFlowState {
  long long *x;            // T*k
  long long *rvs, *exv, *rem;  // V each
  long long *exc;          // T
  // networks of 2^32 clique-member slots or more: 32-bit residuals and
  // clique excess here (x, exc unused), and P.inc holds clique ids from
  // P.sbase / k (the wide kernel); null otherwise
  int *x32 = nullptr, *exc32 = nullptr;
  // implicit clique nodes (k_grid_flow_imp): records of imp_record_words(k)
  // words per clique, the k member ids then the k residual words, from this
  // problem's first clique (P.mem may point at the same records, with
  // P.mstride = imp_record_words(k)); inflow: V words; x, exc, hc, cq0, cq1
  // unused
  uint32_t *rec32 = nullptr;
  long long *inflow = nullptr;
  int *pend = nullptr;  // V words: pending relabels of the implicit flow
  int *hv, *hmin;          // V
  int *hc;                 // T
  uint32_t *list0, *list1; // V each
  uint64_t *lchunk0, *lchunk1;  // V each
  uint32_t *cq0, *cq1;     // T each
  GridFlowCtl *ctl;
  int *reach;              // V
  uint64_t *arc;           // V: current-arc window (first chunk | chunks << 32)
  uint32_t *astop;         // V: first chunk of the window left unsaturated
  // Optional per-round record (nullptr: none): 8 words per round, the
  // clique work-list size, active vertices, their chunks, the round's
  // active-node count, and the wall_clock64 ticks of phase 1, phases 2-3,
  // phase 4 and the global relabel after it; trace_cap rounds at most.
  unsigned long long *trace = nullptr;
  uint32_t trace_cap = 0;
};
// Many independent max flows in one launch (same answers as
// grid_max_flow): item i runs on workgroups [first_block, first_block +
// nblocks), items ascending by first_block, total_blocks in all.  Items on
// several workgroups need `cooperative` (all co-resident: total_blocks <=
// flow_batch_capacity).  Every item's S.ctl must be zeroed before the
// launch.  max_v = largest P.V (sizes the LDS excess accumulator);
// reach_cliques[i] = cliques of item i inside its min-cut side.
struct FlowBatchItem {
  FlowProblem P;
  GridFlowState S;
  long long capT, U;
  uint32_t first_block, nblocks;
};
hipError_t flow_batch(const FlowBatchItem *items, uint32_t count,
                      uint32_t total_blocks, bool cooperative, uint32_t max_rounds,
                      uint32_t max_v, unsigned long long *reach_cliques,
                      hipStream_t s);
// Workgroups of one cooperative flow_batch launch that are surely
// co-resident (one per CU), 0 if the kernel cannot launch.
unsigned flow_batch_capacity(uint32_t max_v, int device);
// flow_batch of single-workgroup problems only (a plain launch) as one HIP
// graph launch: *g (created on the first call, freed by
// flow_batch_graph_free) keeps the graph, its pinned staging and device
// copy of the items; `items` are host items (copied before return).
struct FlowBatchGraph;
hipError_t flow_batch_graph(FlowBatchGraph **g, const FlowBatchItem *items, uint32_t count,
                            uint32_t total_blocks, uint32_t max_rounds, uint32_t max_v,
                            unsigned long long *reach_cliques, hipStream_t s);
void flow_batch_graph_free(FlowBatchGraph *g);
unsigned long long flow_batch_graph_launches(const FlowBatchGraph *g);
hipError_t grid_max_flow(const FlowProblem &P, const GridFlowState &S,
                         long long cap_t, long long unit, uint32_t max_rounds,
                         unsigned *grid, hipStream_t s);
// Workgroups for grid_max_flow on P (0: cannot launch); `blocks` != 0
// overrides the size-based choice; `wide`: the kernel for 2^32 slots or more.
unsigned flow_grid_blocks(const FlowProblem &P, int device, unsigned blocks,
                          bool wide = false, bool imp = false);
// Words per clique record of the implicit-clique max flow: the k member
// ids, then the k residual words, padded so no record crosses a 64-byte line.
uint32_t imp_record_words(uint32_t k);

// Cliques of a component whose members all have reach != 0.
// ms: words per clique in mem / cl (0: k, plain member lists).
hipError_t count_cliques_in_reach(uint32_t T, uint32_t k, const uint32_t *mem,
                                  uint32_t vbase, const int *reach,
                                  unsigned long long *out, hipStream_t s, uint32_t ms = 0);
// Cliques (AoS, k ids) whose members all have mask[id] != 0.
hipError_t count_cliques_in_mask(uint64_t T, uint32_t k, const uint32_t *cl,
                                 const uint8_t *mask, unsigned long long *out,
                                 hipStream_t s, uint32_t ms = 0);

}  // namespace cds
