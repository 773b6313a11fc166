/*
 * cds.h — C ABI of the MI355X k-clique densest-subgraph (DSD) engine.
 *
 * The reference (wajidmanzoor/CDS) has no library boundary: its GPU path is a
 * process, `main <graph> <k> <pSize> <cpSize> <glBufferSize> <partitionSize>
 * <earlyStop>` (cudaCode/main.cu:865-887), reading the adjacency-list format
 * of cudaCode/src/graph.cpp:7-40 and printing the report of
 * cudaCode/main.cu:1069-1104.  Its correctness oracle is the serial CPU path
 * `CDS::DSD` (cpuCode/c++/src/helpers.cpp:1422-1589).  The entry points below
 * are what a foreign-function binding of that path needs: load a graph, solve
 * it, read back the oracle's result block.  The repo's own `cds_main` binary
 * is the drop-in for the reference's CLI and is built on these calls.
 *
 * Conventions
 *   - Every call returns a cds_status; nothing calls exit().  The message of
 *     the last failure on the calling thread is cds_last_error().
 *   - Host arrays passed in stay owned by the caller; arrays returned inside
 *     cds_graph / cds_result are owned by the library and released by
 *     cds_graph_free / cds_result_free.
 *   - Device memory is owned by the engine.  One engine drives one GPU on one
 *     HIP stream; calls on one engine are serialised internally.
 *   - Vertex ids are uint32; clique counts are uint64.
 */
#ifndef CDS_H
#define CDS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  CDS_OK = 0,
  CDS_ERR_INVALID = 1,   /* bad argument / malformed input */
  CDS_ERR_IO = 2,        /* file could not be opened or parsed */
  CDS_ERR_HIP = 3,       /* HIP runtime error (message has the HIP string) */
  CDS_ERR_OOM = 4,       /* device allocation or HBM budget exceeded */
  CDS_ERR_RANGE = 5,     /* value outside the exact range of the engine */
  CDS_ERR_NOT_CONVERGED = 6, /* an iteration bound was hit (never a hang) */
  CDS_ERR_UNSUPPORTED = 7,
  CDS_ERR_BUSY = 8,      /* serving: no slot / budget right now */
  CDS_ERR_CANCELLED = 9,
  CDS_ERR_INTERNAL = 10  /* an internal invariant failed (a bug) */
} cds_status;

/* Host CSR graph: undirected, symmetric, no self loops.
 * Replaces the reference's `Graph` (cudaCode/inc/graph.h:11-29), whose CSR is
 * `offset[n+1]`, `neighbors[2m]` of `ui` (cudaCode/src/graph.cpp:20-33). */
typedef struct {
  uint32_t n;          /* vertices */
  uint64_t m;          /* undirected edges = offsets[n] / 2 */
  uint64_t m_header;   /* m as written in the file header (0 if from CSR) */
  uint32_t *offsets;   /* n+1 entries */
  uint32_t *neighbors; /* offsets[n] entries, each list sorted ascending */
  uint32_t max_degree;
} cds_graph;

/* Solve parameters: the reference CLI's positional arguments
 * (cudaCode/main.cu:875-885) plus the oracle's bound switches
 * (cpuCode/c++/main.cpp:11-12, both on in the reference's canonical run). */
typedef struct {
  uint32_t k;              /* clique size, 2..CDS_MAX_K */
  uint32_t p_size;         /* pSize: partial-clique partition (capacity hint) */
  uint32_t cp_size;        /* cpSize: candidate partition (capacity hint) */
  uint32_t gl_buffer_size; /* glBufferSize: peel frontier buffer per block */
  uint32_t partition_size; /* partitionSize: active-node partition (flow) */
  int32_t early_stop;      /* earlyStop; -1 = off (exact) */
  uint32_t ub1;            /* apply the factorial upper bound (UB1) */
  uint32_t ub2;            /* apply the clique-degree upper bound (UB2) */
  uint32_t want_vertex_arrays; /* fill clique_degree[] and core[] */
} cds_params;

#define CDS_MAX_K 6

/* Per-stage wall times (ms, measured with HIP events on the engine stream). */
typedef struct {
  double upload, dag, listing, core_decompose, densest_core, prune,
      components, exact, total;
} cds_times;

/* Result of one solve.  Field names follow the reference's report keys
 * (cudaCode/main.cu:1069-1104 and cpuCode/c++/src/helpers.cpp:1553-1588). */
typedef struct {
  uint32_t k;
  uint32_t n;
  uint64_t m;
  uint32_t max_degree;
  uint64_t total_cliques;           /* Total Cliques / Total_cliques */
  uint64_t *clique_degree;          /* n entries (optional) */
  uint32_t *core;                   /* n entries: clique-core numbers (opt.) */
  uint64_t kmax;                    /* Denseset_core_kmax (max clique core) */
  uint64_t k_prime;                 /* ceil(max peel-prefix density) */
  double peel_max_density;          /* max peel-prefix density over the level
                                       boundaries and the replayed levels
                                       (levels that cannot raise k_prime are
                                       not replayed); ceil = k_prime */
  double core_density;              /* Denseset_core_density */
  uint32_t core_size;               /* Densest_core_size */
  uint64_t core_edges;              /* Densest_core_edges_count */
  uint64_t core_cliques;            /* cliques inside the densest core */
  uint64_t deleted_cliques;         /* Densest_core_deleted_cliques */
  uint64_t pruned_edges;            /* Remaining_edges_after_prune */
  uint32_t num_components;          /* Component_number / Total Components */
  double paper_lower_bound;         /* Paper_lower_bound */
  float ub1_val, ub2_val;           /* last UB1 / UB2 evaluated */
  /* final answer, as the oracle reports it */
  float density;                    /* Density: (float, oracle semantics) */
  uint32_t size;                    /* Size: */
  uint32_t *vertices;               /* densest subgraph, original ids */
  uint64_t set_cliques;             /* k-cliques inside `vertices` */
  uint32_t set_size;                /* |vertices| (exact fraction denom.) */
  /* density binary-search trace, one entry per (component, alpha step) */
  uint32_t trace_len;
  float *trace_lb, *trace_ub, *trace_alpha;
  uint32_t *trace_component;
  double *trace_margin;             /* kT_c - F(alpha), exact */
  uint32_t flow_networks_solved;    /* components that ran a search */
  uint64_t flow_sweeps;             /* push-relabel sweeps, all alphas */
  uint32_t peel_levels;             /* level-synchronous peel rounds */
  uint32_t replay_phases;           /* phases replayed in exact order */
  uint32_t replay_grid_phases;      /* ... of which on the whole GPU */
  uint32_t oracle_quirk_flags;      /* bit2: oracle's int clique total wraps */
  /* listing statistics (algorithmic-bytes inputs) */
  uint64_t dag_arcs;                /* m */
  uint64_t list_work;               /* W = sum over DAG arcs (u,v) of d+(v) */
  uint32_t max_out_degree;          /* max d+ */
  uint64_t list_items;              /* listing work items */
  double t_list_count_ms;           /* listing count kernel (HIP events) */
  double t_list_write_ms;           /* listing write kernel (HIP events) */
  uint64_t peak_device_bytes;       /* engine arena peak during the solve */
  uint32_t peel_grid;               /* workgroups of the persistent peel */
  uint32_t flow_launches;           /* exact max flows computed */
  uint32_t density_certificates;    /* components proven their own densest */
  uint32_t peel_distinct_levels;    /* level-table entries: exact levels,
                                       plus coarse steps left unrefined */
  uint32_t peel_coarse_shift;       /* 0: exact levels; s: first pass in
                                       steps [m, m + m>>s] */
  uint32_t peel_refined_steps;      /* coarse steps re-peeled exactly */
  uint32_t peel_launches;           /* persistent peel launches (first pass
                                       + re-peels) */
  uint64_t peel_alive_scanned;      /* sum over levels of alive vertices */
  double peel_split_ms[4];          /* one peel thread: barrier, scans,
                                       chunk processing, LDS flushes */
  double t_incidence_ms;            /* clique-incidence build (HIP events) */
  double t_peel_kernel_ms;          /* persistent peel launches, summed
                                       (HIP events) */
  double t_replay_ms;               /* exact phase replays (HIP events) */
  double t_replay_prep_ms;          /* ... of which phase incidence compaction */
  double t_flow_kernel_ms;          /* whole-network max-flow launches, summed
                                       (HIP events) */
  uint64_t flow_clique_visits;      /* clique work-list entries visited by
                                       those flows (phase 1, all rounds) */
  uint64_t flow_arc_visits;         /* vertex incidence arcs posted for
                                       discharge (phase 2, all rounds) */
  uint64_t flow_relabels;           /* global relabel passes of those flows */
  cds_times times;
  /* Cooperative launches (persistent peel, whole-network max flow,
   * whole-GPU replay) keep all their workgroups resident.  peel_grid above,
   * flow_grid and replay_grid are the widths they actually ran at (the
   * smallest over the solve's launches of that kernel; 0: none ran);
   * coop_shrinks counts launches that ran narrower than asked because the
   * runtime refused the cooperative launch (another process on the GPU). */
  uint32_t flow_grid;
  uint32_t replay_grid;
  uint32_t coop_shrinks;
  /* Float fragility of the density search.  The reference decides each
   * alpha step with a float Edmonds-Karp (cpuCode/c++/src/helpers.cpp:
   * 1279-1290, |F - kT| < 1e-2f); this engine decides exactly.
   * trace_drift[i] = 2^-23 (kT_c)^1.5 bounds how far a float flow sum
   * strays from the exact F at step i (calibrated on the reference's float
   * arithmetic over 1110 steps of 147 seeded graphs, where the largest
   * drift was 0.46 of it); trace_fragile[i] = 1 when the exact margin
   * trace_margin[i] lies within that bound of the 1e-2 threshold, i.e. the
   * reference may take the other branch there.  fragile_decisions counts
   * them; when it is 0 the reference's float search is expected to take
   * the same decisions as this result. */
  uint32_t fragile_decisions;
  uint8_t *trace_fragile;           /* trace_len entries */
  double *trace_drift;              /* trace_len entries */
  /* Core pruning, for graphs whose cliques are too many to materialise
   * (k*T slots beyond 32 bits or beyond the device): the peel and the
   * search ran on the prune_tau-clique-core (prune_vertices vertices, found
   * by prune_passes clique-count passes), which holds the densest core
   * whenever prune_tau <= k_prime; 0 when the whole graph was materialised.
   * core[] then holds 0xFFFFFFFF for the vertices outside that core (core
   * number below prune_tau).  dense_peel = 1: the core was peeled on
   * adjacency bitsets, one removal at a time. */
  uint64_t prune_tau;
  uint32_t prune_vertices;
  uint32_t prune_passes;
  uint32_t dense_peel;
  /* Max flows that ran with implicit clique nodes (clique records of member
   * ids and residuals, no clique heights, excess or work lists; the default
   * whenever k*k*U fits 31 bits); the others ran the explicit network. */
  uint32_t flow_implicit;
  /* Batched solves (cds_solve_batch): flow steps of the call whose
   * single-workgroup max flows ran as one HIP graph launch (the same for
   * every graph of the call); 0 for cds_solve. */
  uint32_t flow_graph_launches;
  /* Replayed levels (of replay_phases) that ran on adjacency bitsets in one
   * workgroup (k <= 3, at most 4096 vertices alive at the level's start). */
  uint32_t replay_bits_phases;
  /* Coarse peel steps (of peel_refined_steps) re-peeled exactly on
   * adjacency bitsets in one workgroup, their levels' replays included. */
  uint32_t refine_bits_steps;
  /* 1: the triangle count pass also staged the triangles (t_list_write_ms
   * then times moving them into place); 0: they were listed twice. */
  uint32_t list_staged;
  /* Rounds of the pruned path that dropped vertices in place, removing the
   * cliques through them from the counts of the rest, instead of counting
   * a rebuilt subgraph again (k = 3, 4; few cliques through the drops). */
  uint32_t prune_unlinks;
} cds_result;

typedef struct {
  int device;               /* HIP device ordinal */
  uint64_t hbm_budget_bytes;/* 0 = no budget */
  uint32_t profile_kernels; /* record per-kernel HIP-event times */
  const char *tune_cache;   /* autotune cache file loaded at creation (may be
                               NULL; the CDS_TUNE_CACHE environment variable
                               names one otherwise); a missing file is not an
                               error */
} cds_opts;

/* Launch shapes of the stages that have a choice (no counterpart in the
 * reference, whose shapes are compile-time: BLK_NUMS = 216 blocks of
 * BLK_DIM = 1024 threads, cudaCode/inc/common.h:30-36).  Zero in a field
 * means "the engine's rule". */
typedef struct {
  uint32_t list_threads;      /* listing workgroup lanes: 128, 256 or 512 */
  uint32_t list_chunk;        /* out-neighbours per listing work item: 32 or
                                 64 ((256, 64) by default; 128 lanes need 32) */
  uint32_t peel_blocks;       /* persistent-peel workgroups (<= half the CUs) */
  int32_t peel_coarse_shift;  /* -1 rule (coarse steps from 2^28 incidence
                                 entries), 0 exact levels, s in 1..8: coarse
                                 steps [m, m + m >> s] */
  uint32_t replay_blocks;     /* whole-GPU replay workgroups */
  uint32_t flow_blocks;       /* max-flow workgroups per network */
} cds_tuning;

typedef struct cds_engine cds_engine;
typedef struct cds_dgraph cds_dgraph; /* device-resident graph */

const char *cds_last_error(void);
const char *cds_version(void);
/* sizeof of the ABI structs (0 cds_graph, 1 cds_params, 2 cds_result,
 * 3 cds_opts, 4 cds_tuning) so foreign bindings can verify their layouts. */
size_t cds_struct_size(int which);

/* ---- graph I/O (replaces Graph::Graph, cudaCode/src/graph.cpp:7-40) ---- */
cds_status cds_graph_load_adjlist(const char *path, cds_graph *out);
cds_status cds_graph_from_csr(uint32_t n, const uint32_t *offsets,
                              const uint32_t *neighbors, cds_graph *out);
cds_status cds_graph_write_adjlist(const cds_graph *g, const char *path);
void cds_graph_free(cds_graph *g);
/* Binary CSR file (no counterpart in the reference, whose loader re-parses
 * the text on every run): the cds_graph as stored, for reloading a large
 * graph without parsing.  Layout, little-endian: magic "CDSCSR\0\1",
 * u32 n, u32 max_degree, u64 m, u64 m_header, u32 offsets[n+1],
 * u32 neighbors[2m].  Loading checks every invariant of cds_graph (monotone
 * offsets, ids < n, strictly ascending lists, no self loops, symmetry) and
 * fails with CDS_ERR_IO otherwise. */
cds_status cds_graph_save_csr(const cds_graph *g, const char *path);
cds_status cds_graph_load_csr(const char *path, cds_graph *out);

/* ---- engine (replaces the global device structs of
 *      cudaCode/src/common.cpp:4-12 and gpuMemoryAllocation.cu) ---- */
cds_status cds_engine_create(const cds_opts *opts, cds_engine **out);
void cds_engine_destroy(cds_engine *e);
int cds_engine_device(const cds_engine *e);

/* ---- autotuner (no counterpart in the reference) ----
 * Tunings are kept per (device architecture, CU count, k, edge class), the
 * edge class being floor(log2(m)); a solve uses the entry of its k whose
 * class is nearest to its graph's (within 2), else the engine's default.
 * cds_engine_set_tuning sets the default (NULL: the built-in rules).
 * cds_engine_get_tuning reports what a solve of (k, m) would use. */
cds_status cds_engine_set_tuning(cds_engine *e, const cds_tuning *t);
cds_status cds_engine_get_tuning(cds_engine *e, uint32_t k, uint64_t m,
                                 cds_tuning *out);
/* Loads every entry of the cache file made for this engine's device
 * (returns the number loaded in *loaded, may be NULL). */
cds_status cds_engine_load_tuning(cds_engine *e, const char *path,
                                  uint32_t *loaded);
/* Sweeps the shapes one field at a time on graph dg (each candidate solved
 * and timed; every candidate's answer must equal the first solve's, else
 * CDS_ERR_INTERNAL), keeps the fastest for (k, class of dg) in the engine
 * and, when path is not NULL, merges it into that cache file (written to a
 * temporary name, then renamed).  best / best_ms / base_ms may be NULL. */
cds_status cds_autotune(cds_engine *e, const cds_dgraph *dg, const cds_params *p,
                        const char *path, cds_tuning *best, double *best_ms,
                        double *base_ms);
/* Host-only lookup in a cache file, for tools and tests: the entry for
 * (arch, cus, k) nearest to class floor(log2(m)); CDS_ERR_INVALID if none
 * is within 2 classes, CDS_ERR_IO if the file is unreadable or malformed. */
cds_status cds_tuning_lookup(const char *path, const char *arch, uint32_t cus,
                             uint32_t k, uint64_t m, cds_tuning *out);

/* Upload a host graph to HBM once; solves then start from device memory
 * (replaces memoryAllocationGraph, cudaCode/src/gpuMemoryAllocation.cu:5-27). */
cds_status cds_graph_upload(cds_engine *e, const cds_graph *g,
                            cds_dgraph **out);
void cds_dgraph_free(cds_engine *e, cds_dgraph *dg);

/* Plain path, synchronous: the whole pipeline of cudaCode/main.cu:888-1060
 * with the oracle's semantics (cpuCode/c++/src/helpers.cpp:1422-1589). */
cds_status cds_solve(cds_engine *e, const cds_dgraph *dg,
                     const cds_params *p, cds_result *out);
void cds_result_free(cds_result *r);
/* The result as one JSON object (answer with the exact fraction, core and
 * search statistics, the density-search trace, stage and kernel times).
 * The reference prints only stdout lines (cudaCode/main.cu:1069-1104,
 * cpuCode/c++/src/helpers.cpp:1580-1588); this is their machine-readable
 * form.  snprintf convention: writes at most cap-1 bytes and a NUL, returns
 * the full length (call with cap 0 to size the buffer). */
size_t cds_result_json(const cds_result *r, char *buf, size_t cap);

/* Stage entry points (diagnostics and per-stage parity).
 * cds_count_cliques: T and per-vertex clique degrees
 *   (listAllCliques + countCliques, cudaCode/main.cu:59-250;
 *    oracle listCliques, cpuCode/c++/src/helpers.cpp:373-511).
 *   clique_degree may be NULL.  kernel_ms (may be NULL) receives the time of
 *   the listing kernel(s) measured with HIP events. */
cds_status cds_count_cliques(cds_engine *e, const cds_dgraph *dg, uint32_t k,
                             uint64_t *total, uint64_t *clique_degree,
                             double *kernel_ms);

/* cds_count_cliques without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as cds_count_cliques; every scratch buffer is
 * a slice of one static device array of CDS_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends), and results come back through the
 * engine's pinned block.  A graph whose scratch (about 20n + 4m + 12
 * (n + m/64) bytes) exceeds the array returns CDS_ERR_RANGE.  Calls on one
 * GPU share the array: calls that overlap on one GPU (several engines on
 * one device, driven from several threads) race.  kernel_ms is host wall
 * time around the listing (no events are created). */
#define CDS_NOALLOC_SCRATCH_BYTES (64ull << 20)
cds_status cds_count_cliques_noalloc(cds_engine *e, const cds_dgraph *dg, uint32_t k,
                                     uint64_t *total, uint64_t *clique_degree,
                                     double *kernel_ms);

/* Clique-core stage only: listing, peel and densest-core extraction
 * (cliqueCoreDecompose + generateDensestCore, cudaCode/main.cu:252-399;
 * oracle cliqueCoreDecompose + locateDensestCore,
 * cpuCode/c++/src/helpers.cpp:16-150, :710-788).  Fills the result up to
 * core_size / core_edges / core_cliques / deleted_cliques (and the vertex
 * arrays when asked); the flow fields stay zero. */
cds_status cds_core_decompose(cds_engine *e, const cds_dgraph *dg,
                              const cds_params *p, cds_result *out);

/* Batched plain path: many independent graphs, solved with shared launches.
 * results[i] is identical to cds_solve on graphs[i] with params[i]. */
cds_status cds_solve_batch(cds_engine *e, const cds_dgraph *const *graphs,
                           const cds_params *params, size_t count,
                           cds_result *results);

/* ---- serving (no counterpart in the reference, which solves one graph
 *      per process): a request loop over one engine with HBM-budget
 *      admission and continuous batching.  Requests submitted while a batch
 *      is on the GPU join the next batch; a request's answer equals
 *      cds_solve on its graph whatever it is batched with. ---- */
typedef struct cds_server cds_server;
typedef uint64_t cds_ticket;

typedef struct {
  uint64_t hbm_budget_bytes; /* estimated device bytes of all admitted,
                                unfinished requests; 0 = unlimited */
  uint32_t max_batch;        /* requests per batch (0 = 1024) */
  uint32_t max_wait_us;      /* after the first arrival, wait this long for
                                more before launching (0 = launch at once) */
  uint32_t large_vertices;   /* graphs with at least this many vertices go to
                                the large lane, one at a time (0 = 65536) */
} cds_server_opts;

typedef struct {
  uint64_t submitted, admitted, completed, failed, batches;
  uint64_t max_batch_seen, peak_admitted_bytes, queued;
} cds_server_stats;

typedef struct {
  double queue_ms;   /* submit -> admitted into a batch */
  double service_ms; /* admitted -> completed */
} cds_request_times;

/* Two lanes, each a loop thread with its own engine: the batch lane (on
 * `e`) admits and batches graphs below large_vertices; the large lane (an
 * engine the server creates on e's device) solves bigger graphs one at a
 * time.  `e` must outlive the server and should not be driven by other
 * threads meanwhile (calls would queue on its lock). */
cds_status cds_server_create(cds_engine *e, const cds_server_opts *opts,
                             cds_server **out);
/* Stops the loop (queued requests end CDS_ERR_CANCELLED) and frees results
 * nobody waited for. */
void cds_server_destroy(cds_server *s);
/* Thread-safe.  The graph is copied; the caller keeps its arrays.
 * CDS_ERR_OOM if the request alone exceeds the budget. */
cds_status cds_submit(cds_server *s, const cds_graph *g, const cds_params *p,
                      cds_ticket *out);
cds_status cds_poll(cds_server *s, cds_ticket t, int *done);
/* Waits (timeout_us < 0: forever) and hands the result over (free it with
 * cds_result_free); CDS_ERR_BUSY on timeout, else the request's status.
 * times (may be NULL) receives its queueing and service times.  A ticket
 * can be waited for once. */
cds_status cds_wait(cds_server *s, cds_ticket t, int64_t timeout_us,
                    cds_result *out, cds_request_times *times);
/* Removes a queued request (CDS_ERR_BUSY once it is on the GPU). */
cds_status cds_cancel(cds_server *s, cds_ticket t);
cds_status cds_server_get_stats(cds_server *s, cds_server_stats *out);

/* ---- generators (seeded synthetic inputs; not part of the reference) ---- */
cds_status cds_gen_chung_lu(uint32_t n, double exponent, double mean_degree,
                            uint32_t planted_size, double planted_p,
                            uint64_t seed, cds_graph *out);
cds_status cds_gen_rmat(uint32_t scale, uint32_t edge_factor, double a,
                        double b, double c, uint32_t planted_count,
                        uint32_t planted_min, uint32_t planted_max,
                        double planted_p, uint64_t seed, cds_graph *out);

#ifdef __cplusplus
}
#endif
#endif /* CDS_H */
