// Engine: host orchestration of the HIP pipeline behind the C ABI (cds.h).
//
// Stage order follows main() of the reference GPU path
// (cudaCode/main.cu:888-1060): DAG, listing, clique-core peel, densest
// core, prune, components, per-component exact search.  The result is the
// one the reference's CPU oracle reports (CDS::DSD,
// cpuCode/c++/src/helpers.cpp:1422-1589): its peel-prefix rule for k', its
// densest core, its component order, its float bound / alpha / density
// arithmetic, with every max-flow evaluated exactly on the device.
//
// This file is compiled with -ffp-contract=off and without fast-math: the
// float expressions below must round exactly as written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <optional>
#include <unordered_map>
#include <string>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "tune.h"

using namespace cds;

struct cds_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  DeviceArena *arena = nullptr;
  uint64_t budget = 0;
  bool profile = false;
  std::mutex mu;
  void *pinned = nullptr;  // small readbacks
  // phases whose compacted incidence holds at least this many entries per
  // vertex are replayed on the whole GPU (CDS_REPLAY_GRID_WORK overrides)
  uint64_t replay_grid_work = 2048;
  // workgroups of every max-flow launch; 0 = sized from the network
  // (CDS_FLOW_BLOCKS overrides, e.g. to exercise several on small graphs)
  unsigned flow_blocks = 0;
  // Peel in coarse steps of [m, m + m >> peel_shift] and re-peel exactly
  // only the steps the answer depends on: -1 = when k*T >= kCoarseMinSlots,
  // 0 = never, 1 = always (CDS_PEEL_COARSE, CDS_PEEL_COARSE_SHIFT).  Shift 1
  // (steps [m, 1.5 m]) measured fastest of 1..5 on R-MAT scale 20 k=3,
  // scale 18 k=3 and scale 17 k=4.
  int peel_coarse = -1;
  uint32_t peel_shift = 1;
  bool peel_shift_env = false;
  // launch shapes: device identity for the tune cache, the default tuning,
  // the loaded cache entries, and the tuning of the solve in progress
  std::string arch;
  uint32_t cus = 0;
  cds_tuning deflt{0, 0, 0, -1, 0, 0};
  std::vector<TuneEntry> tuned;
  cds_tuning cur{0, 0, 0, -1, 0, 0};
  bool forced = false;  // autotune sweep: `cur` is set by the sweep
  // Engines that share a device with another engine (the server's lanes)
  // keep every cooperative kernel within half the CUs, so neither lane's
  // kernel needs a CU the other lane's kernels may hold.
  bool coop_half = false;
  // the batched search's single-workgroup flow steps as one HIP graph
  // (built on first use; CDS_BATCH_GRAPH=0 launches them eagerly)
  FlowBatchGraph *flow_graph = nullptr;
};

namespace {
// The tuning a solve of (k, m) uses: the nearest cache entry of k, else the
// engine default.  Called under the engine lock.
cds_tuning tuning_for(const cds_engine *e, uint32_t k, uint64_t m) {
  if (const TuneEntry *t = nearest_entry(e->tuned, e->arch, e->cus, k, m)) return t->t;
  return e->deflt;
}
void begin_solve(cds_engine *e, uint32_t k, uint64_t m) {
  if (!e->forced) e->cur = tuning_for(e, k, m);
}
// Cooperative width of this engine: `grid` capped at half the CUs when the
// engine shares its device.
unsigned coop_width(const cds_engine *e, unsigned grid) {
  if (!e->coop_half || e->cus < 2) return grid;
  return std::min(grid, std::max(1u, e->cus / 2));
}
ListShape list_shape(const cds_engine *e) {
  ListShape s;
  if (e->cur.list_threads) {
    s.threads = e->cur.list_threads;
    s.chunk = e->cur.list_chunk;
  }
  return s;
}
unsigned flow_blocks_of(const cds_engine *e) {
  return e->flow_blocks ? e->flow_blocks : e->cur.flow_blocks;  // CDS_FLOW_BLOCKS first
}
}  // namespace

struct cds_dgraph {
  uint32_t n = 0;
  uint64_t m = 0;
  uint32_t max_degree = 0;
  uint32_t *off = nullptr;
  uint32_t *adj = nullptr;
  std::vector<uint32_t> host_off;  // for host-side statistics
};

namespace {

#define HIPCHK(expr)                                                       \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess)                                                  \
      return fail(e_ == hipErrorOutOfMemory ? CDS_ERR_OOM : CDS_ERR_HIP,   \
                  "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                  __FILE__, __LINE__);                                     \
  } while (0)

#define CHK(expr)                       \
  do {                                  \
    cds_status s_ = (expr);             \
    if (s_ != CDS_OK) return s_;        \
  } while (0)

// Device buffer owned by the arena for the duration of a solve.
template <class T> struct DBuf {
  DeviceArena *a = nullptr;
  T *p = nullptr;
  size_t count = 0;
  explicit DBuf(DeviceArena *arena) : a(arena) {}
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { reset(); }
  void reset() {
    if (p) a->release(p);
    p = nullptr;
    count = 0;
  }
  hipError_t alloc(size_t n) {
    reset();
    count = n;
    return a->alloc((void **)&p, (n ? n : 1) * sizeof(T));
  }
  hipError_t zero() {
    return hipMemsetAsync(p, 0, (count ? count : 1) * sizeof(T), a->stream());
  }
};

struct Timer {
  hipEvent_t ev[24];  // 22, 23: spare pair for sub-stage timing
  int used = 0;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st) {
    for (auto &e : ev) (void)hipEventCreate(&e);
  }
  ~Timer() {
    for (auto &e : ev) (void)hipEventDestroy(e);
  }
  int mark() {
    if (used >= 22) return 21;  // out of events: later marks alias the last
    (void)hipEventRecord(ev[used], s);
    return used++;
  }
  double ms(int a, int b) {
    float t = 0;
    (void)hipEventSynchronize(ev[b]);  // a mark may still be in flight
    (void)hipEventElapsedTime(&t, ev[a], ev[b]);
    return t;
  }
};

// Cooperative launches need every workgroup co-resident.  Several engines
// may share a device (one per host thread, or the server's two lanes), so
// the cooperative workgroups in flight on a device are budgeted per XCD:
// the dispatcher deals a launch's workgroups round-robin over the XCDs
// (eight on MI355X), so a width-G launch places at most ceil(G / 8) on any
// one XCD whatever XCD it starts on.  Every cooperative kernel here runs
// at most one workgroup per CU, so launches whose ceil(G / 8) sum to at
// most CUs / 8 fit on every XCD at once (other kernels only delay them).
// A gate holds its share from the launch until the engine's stream has
// drained.  Widths that are multiples of 8 lose nothing to the rounding.
constexpr int kXcds = 8;

struct CoopBudget {
  std::mutex mu;
  std::condition_variable cv;
  int cap = -1, avail = 0;  // per-XCD workgroup slots; -1: not yet read
};

CoopBudget &coop_budget(int device) {
  static CoopBudget b[64];
  return b[device & 63];
}

struct CoopGate {
  cds_engine *e;
  int slots;  // per-XCD workgroups reserved
  CoopGate(cds_engine *eng, unsigned want) : e(eng), slots(0) {
    if (want <= 1) return;  // one workgroup: a plain launch, nothing to reserve
    CoopBudget &b = coop_budget(e->device);
    std::unique_lock<std::mutex> lk(b.mu);
    if (b.cap < 0) {
      int cus = 0;
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device);
      b.cap = b.avail = std::max(1, cus / kXcds);
    }
    slots = std::min((int)((want + kXcds - 1) / kXcds), b.cap);
    b.cv.wait(lk, [&] { return b.avail >= slots; });
    b.avail -= slots;
  }
  ~CoopGate() {
    if (slots == 0) return;
    (void)hipStreamSynchronize(e->stream);
    CoopBudget &b = coop_budget(e->device);
    {
      std::lock_guard<std::mutex> lk(b.mu);
      b.avail += slots;
    }
    b.cv.notify_all();
  }
  CoopGate(const CoopGate &) = delete;
  CoopGate &operator=(const CoopGate &) = delete;
};

// The engine's pinned host block: the first kPinnedScratch bytes are free
// for callers, the rest stages small readbacks (a copy into pageable
// memory goes through the runtime's shared staging buffer instead).
constexpr size_t kPinnedScratch = 4096;
constexpr size_t kPinnedBytes = 64 * 1024;

template <class T>
cds_status readback(cds_engine *e, T *host, const T *dev, size_t count) {
  const size_t bytes = count * sizeof(T);
  char *stage = static_cast<char *>(e->pinned) + kPinnedScratch;
  const bool in_pinned = (char *)host >= (char *)e->pinned &&
                         (char *)host < (char *)e->pinned + kPinnedBytes;
  if (bytes <= kPinnedBytes - kPinnedScratch && !in_pinned) {
    HIPCHK(hipMemcpyAsync(stage, dev, bytes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    memcpy(host, stage, bytes);
    return CDS_OK;
  }
  HIPCHK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return CDS_OK;
}

template <class T>
cds_status upload(cds_engine *e, T *dev, const T *host, size_t count) {
  HIPCHK(hipMemcpyAsync(dev, host, count * sizeof(T), hipMemcpyHostToDevice,
                        e->stream));
  return CDS_OK;
}

// ---- float arithmetic of the oracle's bounds and search --------------------

// UB1 (cpuCode/c++/src/helpers.cpp:1168-1181) as the reference build
// evaluates it: powf(T, (k-1)/k) * expf(-(sum_i logf(i)) / k).
float ub1_value(uint32_t k, uint64_t total) {
  const float kf = (float)k;
  float log_fact = 0.0f;
  for (uint32_t i = 1; i <= k; ++i) log_fact = log_fact + logf((float)i);
  const float num = powf((float)total, (kf - 1.0f) / kf);
  return num * expf(-(log_fact / kf));
}

// UB2 (helpers.cpp:1186-1198) as the reference build evaluates it:
// T * sum_v fma((cd_v * (1/(k T)))^2).
float ub2_value(uint32_t k, uint64_t total, const uint32_t *cd, uint32_t V) {
  const float dem = (float)((uint64_t)k * total);
  const float inv = 1.0f / dem;
  float acc = 0.0f;
  for (uint32_t v = 0; v < V; ++v) {
    const float pv = (float)cd[v] * inv;
    acc = fmaf(pv, pv, acc);
  }
  return acc * (float)total;
}

// |F - target| < 1e-2f evaluated exactly; F_scaled = F * 2^sc, target is
// an integral float, 1e-2f = 5368709 * 2^-29.
bool flow_equal(__int128 f_scaled, float target, int sc) {
  const __int128 t = (__int128)(long long)target << sc;
  __int128 d = f_scaled - t;
  if (d < 0) d = -d;
  return d * ((__int128)1 << 29) < (__int128)5368709 << sc;
}

// Records one cooperative launch: *field keeps the smallest width launched,
// coop_shrinks counts launches the runtime made run narrower than asked.
void note_width(cds_result *R, uint32_t *field, unsigned asked, unsigned launched) {
  if (launched < asked) R->coop_shrinks++;
  if (*field == 0 || launched < *field) *field = launched;
}

// A cooperative kernel ended because a grid barrier wait ran out.
cds_status barrier_failed(const char *kernel, unsigned grid) {
  return fail(CDS_ERR_HIP,
              "%s on %u workgroups: a grid barrier wait ran out (the workgroups were "
              "not all resident at once, or one stalled); CDS_BARRIER_TIMEOUT_US sets "
              "the limit",
              kernel, grid);
}

// Scale of the float Edmonds-Karp drift bound (float_ek_drift): over 1110
// alpha steps of 147 seeded graphs (R-MAT scale 8 and Chung-Lu, k = 2..4),
// |F_float - F_exact| stayed below 0.92 u (kT)^1.5, u = 2^-24, the size of
// a random walk of about kT augmentations each rounding by about u kT; the
// bound doubles that.  The worst case of float Edmonds-Karp is far larger,
// so the flag is calibrated, not proven.
constexpr double kFloatDriftScale = 2.0;

// Round bound of the whole-GPU flow kernel (NOT_CONVERGED beyond it).
constexpr uint32_t kGridFlowRounds = 20000;
// Round limit of the implicit-clique max flow before the explicit network
// takes the flow over.
constexpr uint32_t kImpFlowRounds = 4096;
// Incidence entries (k*T) from which the plain path peels in coarse steps.
constexpr uint64_t kCoarseMinSlots = 1ull << 24;
constexpr uint32_t kFlowTraceRounds = 4096;

// Whole-GPU replay of heavy phases inside batched unions: off unless
// CDS_UNION_GRID_REPLAY=1.  A served stream (batch lane and large lane
// running side by side) stalled with it on; the cause is not known.
static bool union_grid_replay() {
  static const bool on = [] {
    const char *v = getenv("CDS_UNION_GRID_REPLAY");
    return v && *v && *v != '0';
  }();
  return on;
}

// Bitset replay of levels with few vertices alive (k <= 3): -1 where the
// whole-GPU replay would run (the default), 0 never (CDS_REPLAY_BITS=0),
// 1 for every level it can take (CDS_REPLAY_BITS=1).
static int replay_bits_mode() {
  const char *v = getenv("CDS_REPLAY_BITS");
  if (!v || !*v) return -1;
  return atoi(v) != 0 ? 1 : 0;
}

// CDS_PEEL_TRACE=1: one stderr line per replayed level (diagnostics).
static bool peel_trace_on() {
  const char *v = getenv("CDS_PEEL_TRACE");
  return v && *v && *v != '0';
}

static bool flow_trace_on() {
  const char *v = getenv("CDS_FLOW_TRACE");
  return v && *v && *v != '0';
}

struct Component {
  uint32_t V = 0;        // vertices
  uint64_t T = 0;        // cliques
  uint32_t vbase = 0;    // first flow-vertex id
  uint64_t cbase = 0;    // first clique index
  uint32_t first = 0;    // smallest original vertex id
};

// One max flow the search needs: component `comp`, sink capacity capT and
// unit U (every capacity scaled by U).
struct FlowRequest {
  uint32_t comp = 0;
  long long capT = 0, U = 0;
};

// The oracle's dynamicExact (cpuCode/c++/src/helpers.cpp:1120-1254) for one
// graph, as a state machine so one driver can run it alone and another can
// advance many graphs in lockstep with batched flows.  The arithmetic is the
// oracle's: float lb/ub/alpha/bias, cumulative ub over components (UB1/UB2),
// decision |F - kT| < 1e-2f evaluated exactly on the scaled integer flow,
// strict improvement of the float density, result set = the minimal
// min-cut side of the flow at the last "not equal" alpha.
//
// The max flow at sink capacity c per vertex is F(c) = kT - h(c) with
// h(c) = max over vertex sets S of k T(S) - c |S|: the upper envelope of one
// line per set, convex and piecewise linear, and its minimal maximisers are
// nested.  Before a component's binary search the envelope is built
// exactly: starting from the lines of the empty set and the whole
// component, one flow at the crossing c = k dT / dV of two adjacent known
// lines (capacities scaled by dV, so every capacity is an integer) either
// returns a set strictly above both (inserted between them; its size and
// clique count come from the flow's min-cut side) or proves the envelope
// there is just those two lines.  Only segments whose lines can be maximal
// inside the range of c the search will query are refined, all of them in
// one round of independent flows, so a component costs a few flows (one,
// the certificate, when it is its own densest subgraph) instead of one per
// alpha step; every alpha query is then answered from the envelope with the
// same integer arithmetic as a flow.
class ExactSearch {
 public:
  ExactSearch(uint32_t k, bool ub1, bool ub2, const std::vector<Component> &comps,
              const uint32_t *fcd, double core_density, uint64_t kmax)
      : k_(k), ub1_(ub1), ub2_(ub2), comps_(comps), fcd_(fcd) {
    const uint32_t nc = (uint32_t)comps.size();
    std::vector<double> cdens(nc);
    for (uint32_t c = 0; c < nc; ++c)
      cdens[c] = (double)(int64_t)comps[c].T / ((double)comps[c].V);
    // initial lower bound: the densest component, in float, then the core
    // density if its ceiling is larger (helpers.cpp:1128-1143)
    uint32_t current = 0;
    for (uint32_t c = 0; c < nc; ++c)
      if (lower_ < cdens[c]) {
        lower_ = (float)cdens[c];
        current = c;
      }
    if (std::ceil(lower_) < std::ceil(core_density)) lower_ = (float)core_density;
    upper_ = (float)(int64_t)kmax;
    best_density = lower_;
    best_comp = current;
    best_cliques = nc ? comps[current].T : 0;
    best_size = nc ? comps[current].V : 0;
    for (auto d : cdens)
      if (paper_lower_bound < d) paper_lower_bound = d;
  }

  // Advances the search until it needs flows (true: *reqs holds one round
  // of independent flows, each to be answered with feed_flow / feed_reach
  // by its index) or ends (false).  A failure leaves *err != CDS_OK.
  bool next(std::vector<FlowRequest> *reqs, cds_status *err) {
    *err = CDS_OK;
    reqs->clear();
    for (;;) {
      switch (state_) {
        case kCompStart: {
          if (ci_ >= comps_.size()) {
            state_ = kDone;
            return false;
          }
          const Component &C = comps_[ci_];
          if (ub1_) {
            ub1_val = ub1_value(k_, C.T);
            if (ub1_val < upper_) upper_ = ub1_val;
          }
          if (ub2_) {
            ub2_val = ub2_value(k_, C.T, fcd_ + C.vbase, C.V);
            if (ub2_val < upper_) upper_ = ub2_val;
          }
          // exact(): helpers.cpp:1256-1300
          lb_ = lower_;
          ub_ = upper_;
          alpha_ = (lb_ + ub_) * 0.5f;
          const int32_t vv = (int32_t)((uint32_t)C.V * (uint32_t)(C.V - 1));
          bias_ = 1.0f / (float)vv;
          if (bias_ < 1e-2f) bias_ = 1e-2f;
          target_ = (float)(int64_t)(C.T * k_);
          res.assign(C.V, 1);
          res_cliques_ = C.T;
          if (!((ub_ - lb_) > bias_)) {
            finish_component(false);
            break;
          }
          // envelope: the empty set (tight for every c above the densest
          // density) and the whole component (tight at c = 0)
          chain_.clear();
          chain_.push_back({0, 0, {}, HUGE_VAL});
          chain_.push_back({C.T, C.V, {}, 0.0});
          confirmed_.assign(1, 0);
          // the search only asks for c = fl(alpha k) with lb <= alpha <= ub
          c_lo_ = (double)(lb_ * (float)k_);
          c_hi_ = (double)(ub_ * (float)k_);
          state_ = kEnvelope;
          break;
        }
        case kRound: {
          // insert the sets the last round found (right to left, so the
          // earlier segments keep their indices) and mark the rest final
          for (size_t r = round_.size(); r-- > 0;) {
            Probe &pr = round_[r];
            if (pr.found) {
              chain_.insert(chain_.begin() + pr.seg + 1, std::move(pr.set));
              confirmed_.insert(confirmed_.begin() + pr.seg + 1, 0);
            } else {
              confirmed_[pr.seg] = 1;
            }
          }
          round_.clear();
          state_ = kEnvelope;
          break;
        }
        case kEnvelope: {
          // A segment's unknown lines can only be maximal between the
          // points where its two known lines are tight; segments whose span
          // misses the queried range need no flow.  All others are probed
          // in one round.
          for (size_t i = 0; i < confirmed_.size(); ++i) {
            if (confirmed_[i] || chain_[i].tight < c_lo_ * (1 - 1e-12) ||
                chain_[i + 1].tight > c_hi_ * (1 + 1e-12))
              continue;
            const EnvSet &a = chain_[i], &b = chain_[i + 1];
            FlowRequest rq;
            rq.comp = ci_;
            rq.capT = (long long)(k_ * (b.T - a.T));
            rq.U = (long long)(b.V - a.V);
            reqs->push_back(rq);
            Probe pr;
            pr.seg = i;
            round_.push_back(std::move(pr));
          }
          if (reqs->empty()) {
            if (chain_.size() == 2) certificates++;
            state_ = kLoop;
            break;
          }
          state_ = kRound;
          return true;
        }
        case kLoop: {
          if (!((ub_ - lb_) > bias_)) {
            finish_component(true);
            break;
          }
          const Component &C = comps_[ci_];
          const float cap_t = alpha_ * (float)k_;
          int e2 = 0;
          std::frexp(cap_t, &e2);
          const int sc = std::max(0, 24 - e2);
          if (sc > 40) {
            *err = fail(CDS_ERR_RANGE, "alpha %g too small for exact scaling", alpha_);
            return false;
          }
          // h(c) from the envelope, c = fl(alpha k) scaled by 2^sc; the
          // minimal maximiser is the smallest set attaining it
          const __int128 capT = (__int128)std::ldexp((double)cap_t, sc);
          __int128 h = 0;
          size_t arg = 0;
          for (size_t j = 1; j < chain_.size(); ++j) {
            const __int128 v = ((__int128)(chain_[j].T * k_) << sc) - capT * chain_[j].V;
            if (v > h) {
              h = v;
              arg = j;
            }
          }
          const __int128 F = ((__int128)(C.T * k_) << sc) - h;
          const bool eq = flow_equal(F, target_, sc);
          tr_comp.push_back(ci_);
          tr_lb.push_back(lb_);
          tr_ub.push_back(ub_);
          tr_alpha.push_back(alpha_);
          tr_margin.push_back((double)(C.T * k_) - std::ldexp((double)F, -sc));
          if (eq) {
            ub_ = alpha_;
          } else {
            lb_ = alpha_;
            const EnvSet &s = chain_[arg];
            if (s.reach.empty())  // the whole component
              res.assign(C.V, 1);
            else
              for (uint32_t v = 0; v < C.V; ++v) res[v] = s.reach[v] ? 1 : -1;
            res_cliques_ = s.T;
          }
          alpha_ = (lb_ + ub_) * 0.5f;
          break;
        }
        default:
          *err = fail(CDS_ERR_INTERNAL, "exact search advanced while waiting");
          return false;
      }
    }
  }

  // Result of request i of the current round.  Returns true if the search
  // also needs that flow's min-cut side (feed_reach).
  bool feed_flow(size_t i, long long F) {
    const Component &C = comps_[ci_];
    flows++;
    Probe &pr = round_[i];
    const EnvSet &a = chain_[pr.seg], &b = chain_[pr.seg + 1];
    const __int128 p = (__int128)k_ * (b.T - a.T);
    const __int128 q = (__int128)(b.V - a.V);
    // h at the crossing (scaled by q) versus the two known lines there
    pr.h = (__int128)(C.T * k_) * q - (__int128)F;
    const __int128 known = (__int128)(a.T * k_) * q - p * a.V;
    pr.found = pr.h > known;
    return pr.found;
  }

  // reach[v] != 0 for request i's min-cut side; `cliques` inside it.
  cds_status feed_reach(size_t i, const int *reach, uint64_t cliques) {
    const Component &C = comps_[ci_];
    Probe &pr = round_[i];
    EnvSet &s = pr.set;
    s.T = cliques;
    s.V = 0;
    s.reach.assign(reach, reach + C.V);
    for (uint32_t v = 0; v < C.V; ++v) s.V += reach[v] != 0;
    const EnvSet &a = chain_[pr.seg], &b = chain_[pr.seg + 1];
    const __int128 p = (__int128)k_ * (b.T - a.T);
    const __int128 q = (__int128)(b.V - a.V);
    s.tight = (double)p / (double)q;  // found as the maximiser at this c
    if (!(s.V > a.V && s.V < b.V) || (__int128)(s.T * k_) * q - p * s.V != pr.h)
      return fail(CDS_ERR_INTERNAL,
                  "component %u: envelope set (%llu cliques, %u vertices) inconsistent",
                  ci_, (unsigned long long)s.T, s.V);
    return CDS_OK;
  }

  bool done() const { return state_ == kDone; }

  // results
  float ub1_val = 0, ub2_val = 0;
  double best_density = 0, paper_lower_bound = 0;
  uint32_t best_comp = 0, best_size = 0;
  uint64_t best_cliques = 0;
  std::vector<int> best_res;  // empty = the whole component
  std::vector<int> res;
  std::vector<float> tr_lb, tr_ub, tr_alpha;
  std::vector<uint32_t> tr_comp;
  std::vector<double> tr_margin;
  uint32_t flows_solved = 0, certificates = 0, flows = 0;

 private:
  enum State { kCompStart, kEnvelope, kRound, kLoop, kDone };

  // one line of the envelope: a vertex set, its cliques and size, and its
  // membership (empty for the empty set and the whole component)
  struct EnvSet {
    uint64_t T = 0;
    uint32_t V = 0;
    std::vector<int> reach;
    double tight = 0;  // a c at which the set is a maximiser
  };

  // result accounting of one component (helpers.cpp:1210-1252)
  void finish_component(bool searched) {
    const Component &C = comps_[ci_];
    if (searched) flows_solved++;
    uint32_t vcount = 0;
    for (uint32_t v = 0; v < C.V; ++v) vcount += res[v] != -1;
    if (vcount == 0) vcount = C.V;
    const float density = (float)(int64_t)res_cliques_ / (float)vcount;
    if (density > best_density) {
      best_density = density;
      lower_ = density;
      best_comp = ci_;
      best_res = res;
      best_size = vcount;
      best_cliques = res_cliques_;
    }
    ++ci_;
    state_ = kCompStart;
  }

  uint32_t k_;
  bool ub1_, ub2_;
  const std::vector<Component> &comps_;
  const uint32_t *fcd_;
  float lower_ = 0.0f, upper_ = 0.0f;
  State state_ = kCompStart;
  uint32_t ci_ = 0;
  float lb_ = 0, ub_ = 0, alpha_ = 0, bias_ = 0, target_ = 0;
  uint64_t res_cliques_ = 0;
  std::vector<EnvSet> chain_;  // ascending size
  std::vector<char> confirmed_;  // per adjacent pair of chain_
  // one flow of the current round: the chain segment it probes and what
  // it found there
  struct Probe {
    size_t seg = 0;
    __int128 h = 0;
    bool found = false;
    EnvSet set;
  };
  std::vector<Probe> round_;
  double c_lo_ = 0, c_hi_ = 0;
};

// Bound on how far the reference's float Edmonds-Karp flow sum
// (cpuCode/c++/src/helpers.cpp:1343-1405: float residuals and a float
// running sum over the augmentations) strays from the exact flow on a
// component with T cliques: kFloatDriftScale u (kT)^1.5, u = 2^-24.  The
// decision there is |F - kT| < 1e-2f (helpers.cpp:1279-1290), so a step
// whose exact margin lies within this bound of 1e-2 may go the other way
// in the reference.
double float_ek_drift(uint32_t k, uint64_t T) {
  const double kt = (double)k * (double)T;
  return kFloatDriftScale * std::ldexp(kt * std::sqrt(kt), -24);
}

// Copies the search trace into the result, with each step's drift bound
// and fragility flag.
template <class Comps>
void fill_trace(cds_result *R, const ExactSearch &S, uint32_t k, const Comps &comps) {
  const size_t tl = S.tr_alpha.size();
  R->trace_len = (uint32_t)tl;
  R->trace_lb = host_alloc<float>(tl);
  R->trace_ub = host_alloc<float>(tl);
  R->trace_alpha = host_alloc<float>(tl);
  R->trace_component = host_alloc<uint32_t>(tl);
  R->trace_margin = host_alloc<double>(tl);
  R->trace_fragile = host_alloc<uint8_t>(tl);
  R->trace_drift = host_alloc<double>(tl);
  R->fragile_decisions = 0;
  for (size_t i = 0; i < tl; ++i) {
    R->trace_lb[i] = S.tr_lb[i];
    R->trace_ub[i] = S.tr_ub[i];
    R->trace_alpha[i] = S.tr_alpha[i];
    R->trace_component[i] = S.tr_comp[i];
    R->trace_margin[i] = S.tr_margin[i];
    const double d = float_ek_drift(k, comps[S.tr_comp[i]].T);
    R->trace_drift[i] = d;
    // the float decision |F + e - kT| < 1e-2 with |e| <= d differs from
    // the exact one (margin m = kT - F >= 0) only if |m - 1e-2| <= d
    const bool fr = std::fabs(S.tr_margin[i] - (double)1e-2f) <= d;
    R->trace_fragile[i] = fr ? 1 : 0;
    R->fragile_decisions += fr;
  }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" cds_status cds_engine_create(const cds_opts *opts, cds_engine **out) {
  if (!out) return fail(CDS_ERR_INVALID, "null out");
  int ndev = 0;
  hipError_t he = hipGetDeviceCount(&ndev);
  if (he != hipSuccess || ndev == 0)
    return fail(CDS_ERR_HIP, "no HIP device available (%s)",
                hipGetErrorString(he));
  auto *e = new cds_engine();
  e->device = opts ? opts->device : 0;
  e->budget = opts ? opts->hbm_budget_bytes : 0;
  e->profile = opts ? opts->profile_kernels != 0 : false;
  if (e->device < 0 || e->device >= ndev) {
    delete e;
    return fail(CDS_ERR_INVALID, "device %d out of range (%d devices)",
                opts ? opts->device : 0, ndev);
  }
  hipError_t r = hipSetDevice(e->device);
  if (r == hipSuccess) r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (r == hipSuccess) {
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, e->device) == hipSuccess) {
      uint64_t keep = ~0ull;  // keep freed blocks cached in the pool
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    r = hipHostMalloc(&e->pinned, kPinnedBytes, hipHostMallocDefault);
  }
  // this device's copy of the zero-allocation scratch, looked up once
  if (r == hipSuccess) r = noalloc_attach(e->device);
  if (r != hipSuccess) {
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return fail(CDS_ERR_HIP, "engine init failed: %s", hipGetErrorString(r));
  }
  {
    // large blocks are cached up to 40 % of the device's memory
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) total_b = 0;
    e->arena = new DeviceArena(e->stream, total_b / 10 * 4);
  }
  // Freed stream-ordered memory stays in the device pool between solves:
  // returning ~100 GB to the driver after a large solve and mapping it
  // again for the next one cost up to 1.5 s of wall time per solve.
  {
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, e->device) == hipSuccess && pool) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
  }
  if (const char *fb = getenv("CDS_FLOW_BLOCKS")) e->flow_blocks = (unsigned)atoi(fb);
  if (const char *rw = getenv("CDS_REPLAY_GRID_WORK"))
    e->replay_grid_work = strtoull(rw, nullptr, 10);
  if (const char *pc = getenv("CDS_PEEL_COARSE")) e->peel_coarse = atoi(pc) ? 1 : 0;
  if (const char *ps = getenv("CDS_PEEL_COARSE_SHIFT")) {
    const int v = atoi(ps);
    if (v >= 1 && v <= 16) {
      e->peel_shift = (uint32_t)v;
      e->peel_shift_env = true;
    }
  }
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, e->device) == hipSuccess) {
      std::string a = prop.gcnArchName;
      e->arch = a.substr(0, a.find(':'));
      e->cus = (uint32_t)prop.multiProcessorCount;
    }
    const char *tc = (opts && opts->tune_cache) ? opts->tune_cache : getenv("CDS_TUNE_CACHE");
    if (tc && *tc) {
      std::vector<TuneEntry> all;
      const cds_status ts = read_tune_file(tc, &all, true);
      if (ts != CDS_OK) {
        cds_engine_destroy(e);
        return ts;
      }
      for (const TuneEntry &t : all)
        if (t.arch == e->arch && t.cus == e->cus) e->tuned.push_back(t);
    }
  }
  *out = e;
  return CDS_OK;
}

namespace cds {
void share_device(cds_engine *e, bool shared) {
  std::lock_guard<std::mutex> lk(e->mu);
  e->coop_half = shared;
}
void copy_tuning(cds_engine *dst, cds_engine *src) {
  std::lock_guard<std::mutex> lk(src->mu);
  dst->deflt = src->deflt;
  dst->tuned = src->tuned;
}
}  // namespace cds

extern "C" cds_status cds_engine_set_tuning(cds_engine *e, const cds_tuning *t) {
  if (!e) return fail(CDS_ERR_INVALID, "null engine");
  if (t && !tuning_valid(*t)) return fail(CDS_ERR_INVALID, "tuning out of range");
  std::lock_guard<std::mutex> lk(e->mu);
  e->deflt = t ? *t : cds_tuning{0, 0, 0, -1, 0, 0};
  return CDS_OK;
}

extern "C" cds_status cds_engine_get_tuning(cds_engine *e, uint32_t k, uint64_t m,
                                            cds_tuning *out) {
  if (!e || !out) return fail(CDS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(e->mu);
  *out = tuning_for(e, k, m);
  return CDS_OK;
}

extern "C" cds_status cds_engine_load_tuning(cds_engine *e, const char *path,
                                             uint32_t *loaded) {
  if (!e || !path) return fail(CDS_ERR_INVALID, "null argument");
  std::vector<TuneEntry> all;
  CHK(read_tune_file(path, &all, false));
  std::lock_guard<std::mutex> lk(e->mu);
  uint32_t cnt = 0;
  for (const TuneEntry &t : all) {
    if (t.arch != e->arch || t.cus != e->cus) continue;
    bool replaced = false;
    for (TuneEntry &x : e->tuned)
      if (x.k == t.k && x.cls == t.cls) {
        x = t;
        replaced = true;
      }
    if (!replaced) e->tuned.push_back(t);
    ++cnt;
  }
  if (loaded) *loaded = cnt;
  return CDS_OK;
}

extern "C" int cds_engine_device(const cds_engine *e) { return e ? e->device : -1; }

extern "C" void cds_engine_destroy(cds_engine *e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  delete e->arena;  // frees its cached blocks on the stream
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  {
    // hand the pool's cached memory back (other engines on the device
    // re-grow it on demand)
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, e->device) == hipSuccess && pool)
      (void)hipMemPoolTrimTo(pool, 0);
  }
  flow_batch_graph_free(e->flow_graph);
  if (e->pinned) (void)hipHostFree(e->pinned);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

extern "C" cds_status cds_graph_upload(cds_engine *e, const cds_graph *g,
                                       cds_dgraph **out) {
  if (!e || !g || !out) return fail(CDS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(e->mu);
  HIPCHK(hipSetDevice(e->device));
  auto *d = new cds_dgraph();
  d->n = g->n;
  d->m = g->m;
  d->max_degree = g->max_degree;
  const size_t nn = (size_t)g->n + 1, na = (size_t)g->offsets[g->n];
  if (e->budget && (nn + na) * 4 > e->budget) {
    delete d;
    return fail(CDS_ERR_OOM, "graph exceeds the HBM budget");
  }
  hipError_t r = hipMalloc(&d->off, nn * 4);
  if (r == hipSuccess) r = hipMalloc(&d->adj, (na ? na : 1) * 4);
  if (r == hipSuccess)
    r = hipMemcpyAsync(d->off, g->offsets, nn * 4, hipMemcpyHostToDevice, e->stream);
  if (r == hipSuccess && na)
    r = hipMemcpyAsync(d->adj, g->neighbors, na * 4, hipMemcpyHostToDevice, e->stream);
  if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
  if (r != hipSuccess) {
    (void)hipFree(d->off);
    (void)hipFree(d->adj);
    delete d;
    return fail(CDS_ERR_HIP, "graph upload failed: %s", hipGetErrorString(r));
  }
  d->host_off.assign(g->offsets, g->offsets + nn);
  *out = d;
  return CDS_OK;
}

extern "C" void cds_dgraph_free(cds_engine *e, cds_dgraph *d) {
  if (!d) return;
  if (e) (void)hipSetDevice(e->device);
  (void)hipFree(d->off);
  (void)hipFree(d->adj);
  delete d;
}

extern "C" void cds_result_free(cds_result *r) {
  if (!r) return;
  free(r->clique_degree);
  free(r->core);
  free(r->vertices);
  free(r->trace_lb);
  free(r->trace_ub);
  free(r->trace_alpha);
  free(r->trace_component);
  free(r->trace_margin);
  free(r->trace_fragile);
  free(r->trace_drift);
  memset(r, 0, sizeof(*r));
}

namespace {

// DAG + listing shared by cds_solve and cds_count_cliques.  On return
// (materialise = true) `cliques` holds T*k ids AoS.
struct Listing {
  DBuf<uint32_t> dcnt, doff, dadj, ipv, item_count, cliques;
  DBuf<uint64_t> item_off, item_base, cd;
  DBuf<uint2> items;
  // staged triangles of the count pass (k = 3), moved into place by the
  // write step instead of a second listing pass
  DBuf<uint32_t> stage, sover;
  DBuf<uint64_t> item_tmp;
  DBuf<unsigned long long> scur;
  bool staged = false;
  uint64_t T = 0, items_n = 0, work = 0, tri_bound = 0;
  uint32_t max_out = 0;
  double t_count = 0, t_write = 0;
  explicit Listing(DeviceArena *a)
      : dcnt(a), doff(a), dadj(a), ipv(a), item_count(a), cliques(a),
        item_off(a), item_base(a), cd(a), items(a), stage(a), sover(a), item_tmp(a),
        scur(a) {}
};

cds_status run_dag(cds_engine *e, const cds_dgraph *g, Listing &L) {
  DeviceArena &A = *e->arena;
  const uint32_t n = g->n;
  HIPCHK(L.dcnt.alloc(n));
  HIPCHK(L.doff.alloc((size_t)n + 1));
  HIPCHK(dag_count(n, g->off, g->adj, L.dcnt.p, e->stream));
  HIPCHK(scan_u32_32(L.dcnt.p, L.doff.p, n, A));
  HIPCHK(L.dadj.alloc(g->m));
  HIPCHK(dag_fill(n, g->off, g->adj, L.doff.p, L.dadj.p, e->stream));
  // max d+, W = sum over DAG arcs (u,v) of d+(v) = sum_v indeg(v) * d+(v),
  // and the triangle bound sum over arcs of min(d+(u), d+(v))
  DBuf<unsigned long long> st3(&A);
  HIPCHK(st3.alloc(3));
  HIPCHK(dag_stats(n, g->off, L.dcnt.p, L.doff.p, L.dadj.p, st3.p, e->stream));
  unsigned long long hs[3] = {0, 0, 0};
  CHK(readback(e, hs, st3.p, 3));
  L.max_out = (uint32_t)hs[0];
  L.work = hs[1];
  L.tri_bound = hs[2];
  return CDS_OK;
}

cds_status run_listing_write(cds_engine *e, uint32_t k, Listing &L, Timer *tm);

// Staged triangle listing (on unless CDS_LIST_STAGE=0).
static bool list_stage_on() {
  const char *v = getenv("CDS_LIST_STAGE");
  return !v || !*v || *v != '0';
}

// will_write: the cliques are written later (run_listing_write), so the
// count pass may stage them.
cds_status run_listing(cds_engine *e, const cds_dgraph *g, uint32_t k,
                       bool materialise, Listing &L, Timer *tm, bool will_write = false) {
  DeviceArena &A = *e->arena;
  const uint32_t n = g->n;
  HIPCHK(L.cd.alloc(n));
  HIPCHK(L.cd.zero());
  if (k == 2) {
    L.T = g->m;
    int t0 = tm->mark();
    HIPCHK(list_edges(n, L.doff.p, L.dadj.p, L.cd.p, nullptr, e->stream));
    int t1 = tm->mark();
    if (materialise) {
      HIPCHK(L.cliques.alloc(2 * L.T));
      HIPCHK(list_edges(n, L.doff.p, L.dadj.p, nullptr, L.cliques.p, e->stream));
    }
    int t2 = tm->mark();
    HIPCHK(hipStreamSynchronize(e->stream));
    L.t_count = tm->ms(t0, t1);
    L.t_write = tm->ms(t1, t2);
    return CDS_OK;
  }
  if (k == 4 && L.max_out > kList4LdsCap)
    return fail(CDS_ERR_UNSUPPORTED,
                "k=4 listing needs every DAG out-degree <= %u (max is %u)",
                kList4LdsCap, L.max_out);
  HIPCHK(L.ipv.alloc(n));
  HIPCHK(L.item_off.alloc((size_t)n + 1));
  const ListShape shape = list_shape(e);
  HIPCHK(list_item_counts(n, shape.chunk, L.doff.p, L.ipv.p, e->stream));
  HIPCHK(scan_u32(L.ipv.p, L.item_off.p, n, A));
  CHK(readback(e, &L.items_n, L.item_off.p + n, 1));
  HIPCHK(L.items.alloc(L.items_n));
  HIPCHK(list_item_fill(n, shape.chunk, L.doff.p, L.item_off.p, L.items.p, e->stream));
  HIPCHK(L.item_count.alloc(L.items_n));
  HIPCHK(L.item_base.alloc(L.items_n + 1));
  // Triangles: the count pass also stages every triangle it finds, each
  // item in a reservation of its worst case, when those fit a quarter of
  // the free device memory (CDS_LIST_STAGE=0 turns it off)
  ListStage stg;
  if (k == 3 && (materialise || will_write) && list_stage_on()) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
        (unsigned __int128)L.tri_bound * 12 <= free_b / 4 && L.tri_bound < (1ull << 40)) {
      HIPCHK(L.stage.alloc(3 * std::max<uint64_t>(L.tri_bound, 1)));
      HIPCHK(L.item_tmp.alloc(L.items_n));
      HIPCHK(L.scur.alloc(1));
      HIPCHK(L.scur.zero());
      HIPCHK(L.sover.alloc(1));
      HIPCHK(L.sover.zero());
      stg.buf = L.stage.p;
      stg.cursor = L.scur.p;
      stg.cap = L.tri_bound;
      stg.item_tmp = L.item_tmp.p;
      stg.over = L.sover.p;
    }
  }
  int t0 = tm->mark();
  HIPCHK(list_cliques(k, shape, L.items_n, L.items.p, L.doff.p, L.dadj.p, L.max_out,
                      L.item_count.p, L.cd.p, nullptr, nullptr, e->stream, stg));
  int t1 = tm->mark();
  HIPCHK(scan_u32(L.item_count.p, L.item_base.p, L.items_n, A));
  CHK(readback(e, &L.T, L.item_base.p + L.items_n, 1));
  L.t_count = tm->ms(t0, t1);
  if (stg.buf) {
    uint32_t over = 1;
    CHK(readback(e, &over, L.sover.p, 1));
    L.staged = over == 0;
    if (!L.staged) {  // the bound did not hold (it always should): list again
      L.stage.reset();
      L.item_tmp.reset();
    }
  }
  if (materialise) CHK(run_listing_write(e, k, L, tm));
  return CDS_OK;
}

// The write pass of a listing counted by run_listing (k >= 3): every clique,
// k vertex ids each, at its item's offset.
cds_status run_listing_write(cds_engine *e, uint32_t k, Listing &L, Timer *tm) {
  if (L.T >= 0xFFFFFFFFull || L.T * k >= (1ull << 40))
    return fail(CDS_ERR_RANGE, "%llu cliques exceed the 32-bit clique ids",
                (unsigned long long)L.T);
  const ListShape shape = list_shape(e);
  HIPCHK(L.cliques.alloc(L.T * k));
  int t2 = tm->mark();
  if (L.staged) {
    HIPCHK(list_unstage(L.items_n, L.item_tmp.p, L.item_count.p, L.item_base.p, L.stage.p,
                        L.cliques.p, e->stream));
  } else {
    HIPCHK(list_cliques(k, shape, L.items_n, L.items.p, L.doff.p, L.dadj.p, L.max_out,
                        nullptr, nullptr, L.item_base.p, L.cliques.p, e->stream));
  }
  int t3 = tm->mark();
  HIPCHK(hipStreamSynchronize(e->stream));
  L.t_write = tm->ms(t2, t3);
  L.stage.reset();
  L.item_tmp.reset();
  return CDS_OK;
}

struct Level {
  uint32_t level;  // core value of the step's vertices (an exact level, or
                   // the top of a coarse step)
  uint64_t removed, killed;
  uint32_t lo = kAliveHost;  // coarse step: the minimum degree that opened it
  bool coarse() const { return lo != kAliveHost && lo < level; }
};

// Peel summary of one graph from its level table (ascending levels).
// The state before levels[i] is the clique core C_level; the oracle's
// density rows include each of them (helpers.cpp:125-134), so rho starts
// as the largest boundary density and k' = ceil(rho).  Inside a level l
// every removed vertex has degree <= l, so its intermediate rows stay below
// (T_end + x l) / (N_end + x) for the x <= M-1 vertices still to go; only
// levels above k' can exceed it.  `replay` lists (in descending order) the
// exact levels with at least two vertices whose bound
// ceil((T_end + (M-1) l) / (N_end + M - 1)) exceeds k'.  A coarse step (its
// vertices have core numbers in [lo, level]) obeys the same bound with
// l = level; `refine` lists (descending) the coarse steps that must be
// re-peeled exactly: those whose bound exceeds k', those straddling k'
// (lo < k' <= level: the densest core would split them), the last step
// (kmax is its exact top level), and every coarse step when all core
// numbers are wanted.
struct PeelSummary {
  std::vector<uint64_t> rm_before, kl_before;
  double rho = 0;
  int64_t kp = 0;
  uint64_t kmax = 0;
  std::vector<size_t> replay, refine;
};

cds_status summarize_levels(const std::vector<Level> &levels, uint64_t n,
                            uint64_t T, PeelSummary *S, bool refine_all = false,
                            double rho_replayed = 0) {
  S->rm_before.assign(levels.size(), 0);
  S->kl_before.assign(levels.size(), 0);
  S->kmax = 0;
  uint64_t r = 0, kk = 0;
  for (size_t i = 0; i < levels.size(); ++i) {
    S->rm_before[i] = r;
    S->kl_before[i] = kk;
    r += levels[i].removed;
    kk += levels[i].killed;
    if (levels[i].removed) S->kmax = std::max<uint64_t>(S->kmax, levels[i].level);
  }
  if (kk != T || r != n)
    return fail(CDS_ERR_INTERNAL, "peel removed %llu of %llu vertices, killed %llu of %llu cliques",
                (unsigned long long)r, (unsigned long long)n,
                (unsigned long long)kk, (unsigned long long)T);
  S->rho = 0;
  for (size_t i = 0; i < levels.size(); ++i) {
    const uint64_t N = n - S->rm_before[i], Tr = T - S->kl_before[i];
    const double d = (double)(int64_t)Tr / static_cast<double>(N);
    if (i == 0 || d > S->rho) S->rho = d;
  }
  // densities of rows inside levels replayed so far count before the
  // refine and replay lists are drawn up: a step straddling the raised k'
  // must be re-peeled
  if (rho_replayed > S->rho) S->rho = rho_replayed;
  S->kp = (int64_t)std::ceil(S->rho);
  S->replay.clear();
  S->refine.clear();
  for (size_t ii = levels.size(); ii-- > 0;) {
    const Level &lv = levels[ii];
    if (lv.coarse()) {
      const bool straddle = (int64_t)lv.lo < S->kp && S->kp <= (int64_t)lv.level;
      if (refine_all || straddle || ii + 1 == levels.size()) {
        S->refine.push_back(ii);
        continue;
      }
    }
    if ((int64_t)lv.level <= S->kp) continue;
    const uint64_t M = lv.removed;
    if (M < 2) continue;
    const uint64_t N_end = n - S->rm_before[ii] - M;
    const uint64_t T_end = T - S->kl_before[ii] - lv.killed;
    // f <= (T_end + x*l) / (N_end + x) for x = M-1 remaining phase vertices
    const unsigned __int128 num =
        (unsigned __int128)T_end + (unsigned __int128)(M - 1) * lv.level;
    const unsigned __int128 den = N_end + M - 1;
    const unsigned __int128 ceil_ub = (num + den - 1) / den;
    if ((int64_t)ceil_ub <= S->kp) continue;
    (lv.coarse() ? S->refine : S->replay).push_back(ii);
  }
  return CDS_OK;
}

// Folds one replayed phase's best intermediate density into rho / k'.
void apply_replay(PeelSummary *S, const ReplayOut &ro) {
  if (!ro.found) return;
  const double d = (double)(int64_t)ro.best_num / static_cast<double>(ro.best_den);
  if (d > S->rho) S->rho = d;
  S->kp = std::max(S->kp, (int64_t)std::ceil(d));
}

}  // namespace

extern "C" cds_status cds_count_cliques(cds_engine *e, const cds_dgraph *g,
                                        uint32_t k, uint64_t *total,
                                        uint64_t *clique_degree,
                                        double *kernel_ms) {
  if (!e || !g || !total) return fail(CDS_ERR_INVALID, "null argument");
  if (k < 2 || k > 4)
    return fail(CDS_ERR_UNSUPPORTED, "k=%u: the device lister covers k=2..4", k);
  std::lock_guard<std::mutex> lk(e->mu);
  HIPCHK(hipSetDevice(e->device));
  begin_solve(e, k, g->m);
  Timer tm(e->stream);
  Listing L(e->arena);
  CHK(run_dag(e, g, L));
  CHK(run_listing(e, g, k, false, L, &tm));
  HIPCHK(hipStreamSynchronize(e->stream));
  *total = L.T;
  if (clique_degree) CHK(readback(e, clique_degree, L.cd.p, g->n));
  if (kernel_ms) *kernel_ms = L.t_count;
  return CDS_OK;
}

// No counterpart in the reference: cds_count_cliques for callers that
// cannot allocate.  Every scratch buffer is a slice of this device's static
// scratch array (its address recorded at engine creation); results reach
// the caller through the engine's pinned block.  Nothing here allocates:
// no hipMalloc, no arena, no events.
extern "C" cds_status cds_count_cliques_noalloc(cds_engine *e, const cds_dgraph *g,
                                                uint32_t k, uint64_t *total,
                                                uint64_t *clique_degree,
                                                double *kernel_ms) {
  if (!e || !g || !total) return fail(CDS_ERR_INVALID, "null argument");
  if (k < 2 || k > 4)
    return fail(CDS_ERR_UNSUPPORTED, "k=%u: the device lister covers k=2..4", k);
  std::lock_guard<std::mutex> lk(e->mu);
  HIPCHK(hipSetDevice(e->device));
  char *base = static_cast<char *>(noalloc_scratch(e->device));
  if (!base) return fail(CDS_ERR_INTERNAL, "no scratch recorded for device %d", e->device);
  begin_solve(e, k, g->m);
  const ListShape shape = list_shape(e);
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  hipStream_t st = e->stream;
  // slices, each 256-byte aligned; items_n <= n + m / chunk bounds the items
  const uint64_t items_max = (uint64_t)n + m / shape.chunk + 1;
  const size_t ws = scan_ws_words(std::max<uint64_t>(n, 1));
  size_t used = 0;
  auto carve = [&](size_t bytes) {
    const size_t at = used;
    used += (bytes + 255) & ~(size_t)255;
    return at;
  };
  const size_t o_dcnt = carve(4ull * n), o_doff = carve(4ull * (n + 1)),
               o_dadj = carve(4ull * m), o_maxd = carve(4), o_cd = carve(8ull * n),
               o_ipv = carve(4ull * n), o_ioff = carve(8ull * (n + 1)),
               o_items = carve(8ull * items_max), o_icount = carve(4ull * items_max),
               o_tsum = carve(8), o_ws32 = carve(4 * ws), o_ws64 = carve(8 * ws);
  if (used > kNoallocScratchBytes)
    return fail(CDS_ERR_RANGE,
                "graph needs %zu bytes of scratch (n=%u, m=%llu); the zero-allocation "
                "count holds %zu",
                used, n, (unsigned long long)m, (size_t)kNoallocScratchBytes);
  auto *dcnt = reinterpret_cast<uint32_t *>(base + o_dcnt);
  auto *doff = reinterpret_cast<uint32_t *>(base + o_doff);
  auto *dadj = reinterpret_cast<uint32_t *>(base + o_dadj);
  auto *maxd = reinterpret_cast<uint32_t *>(base + o_maxd);
  auto *cd = reinterpret_cast<uint64_t *>(base + o_cd);
  auto *ipv = reinterpret_cast<uint32_t *>(base + o_ipv);
  auto *ioff = reinterpret_cast<uint64_t *>(base + o_ioff);
  auto *items = reinterpret_cast<uint2 *>(base + o_items);
  auto *icount = reinterpret_cast<uint32_t *>(base + o_icount);
  auto *tsum = reinterpret_cast<unsigned long long *>(base + o_tsum);
  auto *ws32 = reinterpret_cast<uint32_t *>(base + o_ws32);
  auto *ws64 = reinterpret_cast<uint64_t *>(base + o_ws64);
  uint32_t *hp = static_cast<uint32_t *>(e->pinned);  // the caller part of the block
  auto *hp64 = reinterpret_cast<uint64_t *>(static_cast<char *>(e->pinned) + 8);
  // orientation, DAG offsets and lists, largest out-degree
  HIPCHK(dag_count(n, g->off, g->adj, dcnt, st));
  HIPCHK(scan_u32_32_ws(dcnt, doff, n, ws32, ws, st));
  HIPCHK(dag_fill(n, g->off, g->adj, doff, dadj, st));
  HIPCHK(reduce_max_u32(dcnt, n, maxd, st));
  HIPCHK(fill_u64(cd, n, 0, st));
  HIPCHK(hipMemcpyAsync(hp, maxd, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const uint32_t max_out = hp[0];
  uint64_t T = 0;
  const auto t0 = std::chrono::steady_clock::now();
  if (k == 2) {
    HIPCHK(list_edges(n, doff, dadj, cd, nullptr, st));
    T = m;
  } else {
    if (k == 4 && max_out > kList4LdsCap)
      return fail(CDS_ERR_UNSUPPORTED,
                  "k=4 listing needs every DAG out-degree <= %u (max is %u)", kList4LdsCap,
                  max_out);
    HIPCHK(list_item_counts(n, shape.chunk, doff, ipv, st));
    HIPCHK(scan_u32_ws(ipv, ioff, n, ws64, ws, st));
    HIPCHK(hipMemcpyAsync(hp64, ioff + n, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const uint64_t items_n = hp64[0];
    if (items_n > items_max)
      return fail(CDS_ERR_INTERNAL, "%llu listing items above the bound %llu",
                  (unsigned long long)items_n, (unsigned long long)items_max);
    HIPCHK(list_item_fill(n, shape.chunk, doff, ioff, items, st));
    HIPCHK(list_cliques(k, shape, items_n, items, doff, dadj, max_out, icount, cd, nullptr,
                        nullptr, st));
    HIPCHK(sum_u32(icount, items_n, tsum, st));
    HIPCHK(hipMemcpyAsync(hp64, tsum, 8, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  if (k != 2) T = hp64[0];
  if (kernel_ms)
    *kernel_ms = std::chrono::duration<double, std::milli>(
                     std::chrono::steady_clock::now() - t0)
                     .count();
  *total = T;
  if (clique_degree) {
    // through the pinned block, past its first kPinnedScratch bytes
    const size_t per = (kPinnedBytes - kPinnedScratch) / sizeof(uint64_t);
    auto *stage =
        reinterpret_cast<uint64_t *>(static_cast<char *>(e->pinned) + kPinnedScratch);
    for (size_t at = 0; at < n; at += per) {
      const size_t c = std::min<size_t>(per, n - at);
      HIPCHK(hipMemcpyAsync(stage, cd + at, c * 8, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      memcpy(clique_degree + at, stage, c * 8);
    }
  }
  return CDS_OK;
}

namespace {

// Peel stage of a small dense graph (n <= kDenseMaxVertices): every row of
// the oracle's sequential peel (cpuCode/c++/src/helpers.cpp:16-150) from
// the bitset kernels of prune.hip, so k' is exact without replays.  Fills
// the exact level table, its summary (rho and k' over all rows) and core[];
// kill is allocated for T words, to be derived from core[] when read.
cds_status dense_peel_stage(cds_engine *e, const cds_dgraph *g, const Listing &L, uint32_t k,
                            cds_result *R, DBuf<uint32_t> *core, DBuf<uint32_t> *kill,
                            std::vector<Level> *levels, PeelSummary *PS) {
  DeviceArena &A = *e->arena;
  hipStream_t st = e->stream;
  const uint32_t n = g->n;
  const uint64_t T = L.T;
  if (n > kDenseMaxVertices)
    return fail(CDS_ERR_RANGE, "dense peel of %u vertices (at most %u)", n, kDenseMaxVertices);
  const uint32_t W = (n + 63) / 64;
  DBuf<unsigned long long> bits(&A), alive(&A), cur(&A), rval(&A);
  DBuf<uint32_t> stepv(&A), rv(&A);
  HIPCHK(bits.alloc((size_t)n * W));
  HIPCHK(alive.alloc(W));
  HIPCHK(cur.alloc(n));
  HIPCHK(rval.alloc(n));
  HIPCHK(stepv.alloc(1));
  HIPCHK(rv.alloc(n));
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(dense_peel(n, k, g->off, g->adj, L.cd.p, bits.p, alive.p, cur.p, stepv.p, rv.p,
                    rval.p, st));
  std::vector<uint32_t> hrv(n);
  std::vector<unsigned long long> hval(n);
  CHK(readback(e, hrv.data(), rv.p, n));
  CHK(readback(e, hval.data(), rval.p, n));
  R->t_peel_kernel_ms = std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - t0)
                            .count();
  // rows 0..n-1 of the oracle: T_rem / (n - i) after i removals
  std::vector<uint32_t> hcore(n, kAliveHost);
  uint64_t run = 0, trem = T;
  double rho = (double)(int64_t)T / static_cast<double>(n);
  levels->clear();
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t v = hrv[i];
    const uint64_t val = hval[i];
    if (v >= n || hcore[v] != kAliveHost || val > trem)
      return fail(CDS_ERR_INTERNAL, "dense peel row %u: vertex %u, %llu of %llu cliques", i,
                  v, (unsigned long long)val, (unsigned long long)trem);
    if (val > 0xFFFFFF) R->oracle_quirk_flags |= 1;  // the oracle's sentinel regime
    run = std::max<uint64_t>(run, val);
    hcore[v] = (uint32_t)run;
    if (levels->empty() || levels->back().level != run) levels->push_back({(uint32_t)run, 0, 0});
    levels->back().removed++;
    levels->back().killed += val;
    trem -= val;
    if (i + 1 < n) rho = std::max(rho, (double)(int64_t)trem / static_cast<double>(n - i - 1));
  }
  if (trem != 0)
    return fail(CDS_ERR_INTERNAL, "dense peel left %llu cliques", (unsigned long long)trem);
  CHK(summarize_levels(*levels, n, T, PS));
  PS->rho = std::max(PS->rho, rho);
  PS->kp = (int64_t)std::ceil(PS->rho);
  PS->replay.clear();
  PS->refine.clear();
  HIPCHK(core->alloc(n));
  CHK(upload(e, core->p, hcore.data(), n));
  HIPCHK(kill->alloc(T));
  HIPCHK(hipStreamSynchronize(st));  // hcore is read by the copy above
  R->peel_levels = n;
  R->peel_launches = 0;
  return CDS_OK;
}

// How solve_impl treats its graph: the whole input, or the clique core the
// pruned path cut it to (peeled with the persistent peel or, for a dense
// core, on bitsets).  Only the whole input may be pruned.
enum SolveMode { kSolveInput, kSolveCore, kSolveDenseCore };

cds_status solve_pruned(cds_engine *e, const cds_dgraph *g, const cds_params *p,
                        cds_result *R, bool core_only, Listing &L0);

// Whether a graph with T k-cliques is solved on its clique core instead of
// whole: its clique slots pass 32 bits, or its cliques and their
// incidence would not fit the device (CDS_CORE_PRUNE=1/0 forces it).
bool needs_prune(uint32_t k, uint64_t T) {
  if (const char *f = getenv("CDS_CORE_PRUNE")) return atoi(f) != 0;
  if ((unsigned __int128)k * T >= 0xFFFFFFFFull) return true;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
  const unsigned __int128 need = (unsigned __int128)T * (4ull * k + 4ull * k * k + 16);
  return need > (unsigned __int128)free_b / 10 * 8;
}

cds_status solve_impl(cds_engine *e, const cds_dgraph *g, const cds_params *p,
                      cds_result *R, bool core_only, SolveMode mode = kSolveInput) {
  const bool dense = mode == kSolveDenseCore;
  const auto wall0 = std::chrono::steady_clock::now();
  DeviceArena &A = *e->arena;
  hipStream_t st = e->stream;
  const uint32_t n = g->n, k = p->k;
  memset(R, 0, sizeof(*R));
  begin_solve(e, k, g->m);
  R->k = k;
  R->n = n;
  R->m = g->m;
  R->max_degree = g->max_degree;
  if (n == 0) return fail(CDS_ERR_INVALID, "empty graph");
  Timer tm(st);
  const int ev_start = tm.mark();

  // ---- (a) DAG + listing ----
  Listing L(&A);
  CHK(run_dag(e, g, L));
  const int ev_dag = tm.mark();
  CHK(run_listing(e, g, k, k == 2, L, &tm, true));
  if (k > 2) {
    if (mode == kSolveInput && needs_prune(k, L.T)) {
      L.stage.reset();
      L.item_tmp.reset();
      return solve_pruned(e, g, p, R, core_only, L);
    }
    CHK(run_listing_write(e, k, L, &tm));
  }
  const int ev_list = tm.mark();
  const uint64_t T = L.T;
  R->total_cliques = T;
  R->dag_arcs = g->m;
  R->list_work = L.work;
  R->max_out_degree = L.max_out;
  R->list_items = L.items_n;
  R->t_list_count_ms = L.t_count;
  R->t_list_write_ms = L.t_write;
  R->list_staged = L.staged ? 1 : 0;
  std::vector<uint64_t> cd(n);
  CHK(readback(e, cd.data(), L.cd.p, n));
  uint64_t sum_cd = 0, max_cd = 0;
  for (uint64_t x : cd) {
    sum_cd += x;
    max_cd = std::max(max_cd, x);
  }
  if (sum_cd != T * k)
    return fail(CDS_ERR_INTERNAL, "listing invariant: sum cd %llu != k*T %llu",
                (unsigned long long)sum_cd, (unsigned long long)(T * k));
  if (max_cd >= 0xFFFFFFFFull)
    return fail(CDS_ERR_RANGE, "a clique degree exceeds 32 bits");
  if (sum_cd >= (1ull << 31)) R->oracle_quirk_flags |= 4;  // int total wraps

  // ---- (b) peel ----
  DBuf<uint32_t> core(&A), kill(&A);
  // kill words are derived from the core numbers only when a stage reads
  // them (the persistent peel claims cliques by stamps, not kill words)
  bool kill_valid = true;
  auto ensure_kill = [&]() -> cds_status {
    if (!kill_valid) {
      HIPCHK(peel_derive_kill(T, k, L.cliques.p, core.p, kill.p, st));
      kill_valid = true;
    }
    return CDS_OK;
  };
  std::vector<Level> levels;
  PeelSummary PS;
  uint32_t peel_rounds = 0;
  if (dense) {
    // a dense core: the exact sequential peel on adjacency bitsets
    CHK(dense_peel_stage(e, g, L, k, R, &core, &kill, &levels, &PS));
    kill_valid = false;
  } else {
    DBuf<uint64_t> ioff(&A);
    DBuf<uint32_t> icur(&A), inc(&A), deg(&A), fa(&A),
        fb(&A), cnt(&A);
    DBuf<unsigned long long> killed(&A);
    DBuf<uint32_t> sdeg(&A);  // degree at the start of the removing level
    HIPCHK(ioff.alloc((size_t)n + 1));
    HIPCHK(scan_u64(L.cd.p, ioff.p, n, A));
    HIPCHK(icur.alloc(n));
    HIPCHK(icur.zero());
    HIPCHK(inc.alloc(T * k * k));  // entries: clique id + other k-1 members
    const int ev_inc0 = tm.mark();
    HIPCHK(build_incidence(T, k, L.cliques.p, ioff.p, icur.p, inc.p, st));
    const int ev_inc1 = tm.mark();
    icur.reset();
    HIPCHK(deg.alloc(n));
    HIPCHK(core.alloc(n));
    HIPCHK(kill.alloc(T));
    HIPCHK(peel_init(n, T, L.cd.p, deg.p, core.p, kill.p, st));
    uint32_t *hp = (uint32_t *)e->pinned;
    uint64_t removed_total = 0;
    // persistent-peel state, kept for the exact re-peels of coarse steps
    DBuf<uint32_t> s32(&A), lvl(&A), lmin(&A), ralive(&A), rcount(&A);
    DBuf<uint64_t> s64(&A);
    DBuf<uint32_t> stamp(&A);
    uint32_t sbase = 1;  // first sub-round of the next persistent launch
    DBuf<unsigned long long> lrem(&A), lkil(&A);
    DBuf<PeelCtl> ctl(&A);
    // CDS_PEEL_COARSE / CDS_PEEL_COARSE_SHIFT (tests, diagnostics) win over
    // the tuning, which wins over the rule (coarse steps of [m, 1.5 m] from
    // kCoarseMinSlots incidence entries)
    bool coarse_peel = (uint64_t)k * T >= kCoarseMinSlots;
    uint32_t shift = 1;
    if (e->cur.peel_coarse_shift >= 0) {
      coarse_peel = e->cur.peel_coarse_shift > 0;
      shift = coarse_peel ? (uint32_t)e->cur.peel_coarse_shift : 1;
    }
    if (e->peel_coarse != -1) coarse_peel = e->peel_coarse == 1;
    if (e->peel_shift_env) shift = e->peel_shift;
    const uint32_t cshift = (n < (1u << 24) && coarse_peel) ? shift : 0;
    R->peel_coarse_shift = cshift;
    int wall_khz = 0;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, e->device) !=
        hipSuccess)
      wall_khz = 0;
    // One persistent launch over `alive0` (null: every vertex) appending its
    // steps to `out`; returns the launch's control block.
    // every launch is timed with the spare event pair; R->t_peel_kernel_ms
    // sums them
    auto run_peel = [&](const uint32_t *alive0, uint32_t na0, uint32_t shift,
                        uint32_t max_level, unsigned grid, uint64_t want_removed,
                        std::vector<Level> *out, PeelCtl *hc) -> cds_status {
      unsigned launched = grid;
      {
        CoopGate gate(e, grid);
        HIPCHK(hipEventRecord(tm.ev[22], st));
        HIPCHK(peel_persistent(n, k, T, ioff.p, inc.p, L.cliques.p, deg.p, core.p, kill.p,
                               stamp.p, sbase, s32.p, s64.p, ctl.p, lvl.p, lmin.p, lrem.p,
                               lkil.p, sdeg.p,
                               alive0, na0, shift, max_level, &launched, st));
        HIPCHK(hipEventRecord(tm.ev[23], st));
      }
      note_width(R, &R->peel_grid, grid, launched);
      CHK(readback(e, hc, ctl.p, 1));
      if (hc->bar.abort) return barrier_failed("persistent peel", launched);
      kill_valid = false;
      R->t_peel_kernel_ms += tm.ms(22, 23);
      R->peel_launches++;
      if (hc->rounds >= (1ull << 30) || sbase + hc->rounds + 1 >= (1u << 30))
        return fail(CDS_ERR_INTERNAL, "persistent peel: %llu sub-rounds", hc->rounds);
      sbase += (uint32_t)hc->rounds + 1;
      if (hc->removed != want_removed || hc->nlevels == 0 || hc->nlevels > n)
        return fail(CDS_ERR_INTERNAL, "persistent peel removed %llu of %llu (%u levels)",
                    hc->removed, (unsigned long long)want_removed, hc->nlevels);
      std::vector<uint32_t> hl(hc->nlevels), hm(hc->nlevels);
      std::vector<unsigned long long> hr(hc->nlevels), hk(hc->nlevels);
      CHK(readback(e, hl.data(), lvl.p, hc->nlevels));
      CHK(readback(e, hm.data(), lmin.p, hc->nlevels));
      CHK(readback(e, hr.data(), lrem.p, hc->nlevels));
      CHK(readback(e, hk.data(), lkil.p, hc->nlevels));
      for (uint32_t i = 0; i < hc->nlevels; ++i) {
        Level lv{hl[i], hr[i], hk[i]};
        if (hm[i] < hl[i]) lv.lo = hm[i];
        out->push_back(lv);
      }
      peel_rounds += (uint32_t)hc->rounds;
      if (wall_khz > 0) {
        R->peel_split_ms[0] += (double)hc->t_bar / wall_khz;
        R->peel_split_ms[1] += (double)hc->t_scan / wall_khz;
        R->peel_split_ms[2] += (double)hc->t_proc / wall_khz;
        R->peel_split_ms[3] += (double)hc->t_flush / wall_khz;
      }
      return CDS_OK;
    };
    if (n < (1u << 24)) {
      // every level (or coarse step) in one cooperative launch
      HIPCHK(s32.alloc(6ull * n));
      HIPCHK(s64.alloc(3ull * n));
      HIPCHK(lvl.alloc(n));
      HIPCHK(lmin.alloc(n));
      HIPCHK(lrem.alloc(n));
      HIPCHK(lkil.alloc(n));
      HIPCHK(ctl.alloc(1));
      HIPCHK(sdeg.alloc(n));
      HIPCHK(stamp.alloc(n));
      const unsigned grid = coop_width(e, peel_grid_blocks(n, T, k, e->device, e->cur.peel_blocks));
      if (grid == 0) return fail(CDS_ERR_HIP, "the persistent peel cannot be launched");
      PeelCtl hc;
      CHK(run_peel(nullptr, 0, cshift, kAliveHost, grid, n, &levels, &hc));
      removed_total = n;
    }
    HIPCHK(fa.alloc(n));
    HIPCHK(fb.alloc(n));
    HIPCHK(cnt.alloc(4));  // [0] min, [1] count A, [2] count B
    HIPCHK(killed.alloc(1));
    while (removed_total < n) {
      HIPCHK(peel_min_alive(n, deg.p, core.p, cnt.p, st));
      HIPCHK(hipMemsetAsync(cnt.p + 1, 0, 8, st));
      HIPCHK(killed.zero());
      CHK(readback(e, hp, cnt.p, 1));
      const uint32_t level = hp[0];
      if (level == kAliveHost)
        return fail(CDS_ERR_INTERNAL, "peel: no alive vertex but %llu removed",
                    (unsigned long long)removed_total);
      HIPCHK(peel_frontier(n, level, deg.p, core.p, fa.p, cnt.p + 1, st));
      CHK(readback(e, hp + 1, cnt.p + 1, 1));
      uint32_t f = hp[1];
      uint64_t removed = 0;
      uint32_t *cur = fa.p, *nxt = fb.p, *ccur = cnt.p + 1, *cnxt = cnt.p + 2;
      while (f > 0) {
        removed += f;
        HIPCHK(hipMemsetAsync(cnxt, 0, 4, st));
        HIPCHK(peel_process(k, level, cur, ccur, f, ioff.p, inc.p, L.cliques.p,
                            deg.p, core.p, kill.p, nxt, cnxt, killed.p, st));
        CHK(readback(e, hp + 2, cnxt, 1));
        f = hp[2];
        std::swap(cur, nxt);
        std::swap(ccur, cnxt);
        ++peel_rounds;
      }
      unsigned long long kl = 0;
      CHK(readback(e, &kl, killed.p, 1));
      levels.push_back({level, removed, kl});
      removed_total += removed;
    }
    fa.reset();
    fb.reset();
    R->t_incidence_ms = tm.ms(ev_inc0, ev_inc1);

    std::vector<std::pair<uint32_t, ReplayOut>> replayed;  // by level
    // Replays of exact levels taken from a bitset re-peel's removal
    // sequence; each is used when its level becomes a replay candidate.
    std::vector<std::pair<uint32_t, ReplayOut>> seq_replays;
    // Exact re-peel of coarse step i on adjacency bitsets (k <= 3, at most
    // kReplayBitsMaxVertices alive at the step's start): one workgroup
    // removes the step's vertices one at a time in the oracle's order, so
    // the sequence gives the core numbers (running maximum of the removal
    // counts), the exact levels, and every level's replay at once.
    auto refine_step_bits = [&](size_t i) -> cds_status {
      const Level lv = levels[i];
      const uint32_t Nb = (uint32_t)(n - PS.rm_before[i]), M = (uint32_t)lv.removed;
      const uint64_t T_start = T - PS.kl_before[i];
      DBuf<uint32_t> av(&A), pmap(&A), cnt(&A), seq(&A), hvd(&A), hcd(&A);
      DBuf<unsigned long long> bits(&A);
      DBuf<ReplayOut> rod(&A);
      HIPCHK(av.alloc(n));
      HIPCHK(pmap.alloc(n));
      HIPCHK(cnt.alloc(1));
      HIPCHK(seq.alloc(2ull * M));
      HIPCHK(bits.alloc((size_t)Nb * ((Nb + 63) / 64)));
      HIPCHK(rod.alloc(1));
      HIPCHK(replay_phase_bits(k, lv.level, n, L.doff.p, L.dadj.p, core.p, sdeg.p, T_start, Nb,
                               M, av.p, pmap.p, cnt.p, bits.p, rod.p, nullptr, 0, 0, st,
                               seq.p));
      ReplayOut ro;
      CHK(readback(e, &ro, rod.p, 1));
      if (ro.steps != M)
        return fail(CDS_ERR_INTERNAL, "bitset re-peel of step %u ran %u of %u removals",
                    lv.level, ro.steps, M);
      std::vector<uint32_t> hs(2ull * M), hv(M), hc(M);
      CHK(readback(e, hs.data(), seq.p, 2ull * M));
      std::vector<Level> exact;
      uint32_t run = i > 0 ? levels[i - 1].level : 0;
      for (uint32_t j = 0; j < M; ++j) {
        run = std::max(run, hs[2 * j + 1]);
        hv[j] = hs[2 * j];
        hc[j] = run;
        if (hv[j] >= n) return fail(CDS_ERR_INTERNAL, "bitset re-peel removed vertex %u", hv[j]);
        if (exact.empty() || exact.back().level != run) exact.push_back({run, 0, 0});
        exact.back().removed++;
        exact.back().killed += hs[2 * j + 1];
      }
      uint64_t kl = 0;
      for (const Level &x : exact) {
        kl += x.killed;
        if (x.level < lv.lo || x.level > lv.level)
          return fail(CDS_ERR_INTERNAL, "coarse step [%u, %u] re-peeled to level %u", lv.lo,
                      lv.level, x.level);
      }
      if (kl != lv.killed)
        return fail(CDS_ERR_INTERNAL, "coarse step %u killed %llu of %llu cliques", lv.level,
                    (unsigned long long)kl, (unsigned long long)lv.killed);
      // each exact level's replay: its densest row strictly inside it
      uint64_t Tl = T_start, Nl = Nb;
      uint32_t j0 = 0;
      for (const Level &x : exact) {
        if (x.removed >= 2) {
          ReplayOut r{};
          r.best_den = 1;
          uint64_t Rl = 0;
          for (uint64_t q = 0; q < x.removed; ++q) {
            Rl += hs[2 * (j0 + q) + 1];
            if (q + 1 < x.removed) {
              const uint64_t num = Tl - Rl, den = Nl - (q + 1);
              if (!r.found || (unsigned __int128)num * r.best_den >
                                  (unsigned __int128)r.best_num * den) {
                r.best_num = num;
                r.best_den = den;
                r.found = 1;
              }
            }
          }
          r.steps = (uint32_t)x.removed;
          seq_replays.push_back({x.level, r});
        }
        Tl -= x.killed;
        Nl -= x.removed;
        j0 += (uint32_t)x.removed;
      }
      HIPCHK(hvd.alloc(M));
      HIPCHK(hcd.alloc(M));
      CHK(upload(e, hvd.p, hv.data(), M));
      CHK(upload(e, hcd.p, hc.data(), M));
      HIPCHK(scatter_u32(M, hvd.p, hcd.p, core.p, st));
      HIPCHK(hipStreamSynchronize(st));
      kill_valid = false;
      levels.erase(levels.begin() + (ptrdiff_t)i);
      levels.insert(levels.begin() + (ptrdiff_t)i, exact.begin(), exact.end());
      R->peel_refined_steps++;
      R->refine_bits_steps++;
      return CDS_OK;
    };

    // Exact re-peel of coarse step i: its vertices and cliques become alive
    // again (degrees from sdeg), later steps are parked, and the step's
    // entry is replaced by its exact levels.
    auto refine_step = [&](size_t i) -> cds_status {
      const Level lv = levels[i];
      if ((k == 2 || k == 3) && sdeg.p && replay_bits_mode() != 0 &&
          n - PS.rm_before[i] <= kReplayBitsMaxVertices)
        return refine_step_bits(i);
      if (!ralive.p) {
        HIPCHK(ralive.alloc(n));
        HIPCHK(rcount.alloc(1));
      }
      HIPCHK(refine_prepare(n, T, lv.level, core.p, deg.p, sdeg.p, nullptr, ralive.p,
                            rcount.p, st));
      uint32_t na = 0;
      CHK(readback(e, &na, rcount.p, 1));
      if (na != lv.removed)
        return fail(CDS_ERR_INTERNAL, "coarse step %u holds %u of %llu vertices", lv.level,
                    na, (unsigned long long)lv.removed);
      // as wide as the first pass: the re-peel is bound by its cliques'
      // kills (measured 127 ms against 157 ms with a grid sized by the
      // step's own clique count on R-MAT scale 20, k = 3)
      const unsigned grid = coop_width(e, peel_grid_blocks(n, T, k, e->device, e->cur.peel_blocks));
      if (grid == 0) return fail(CDS_ERR_HIP, "the persistent peel cannot be launched");
      // Exact levels at once: re-peeling a step in narrower steps first and
      // refining those re-processes the step's cliques at every depth, which
      // measured slower on R-MAT scale 20 k=3 and scale 17 k=4.
      std::vector<Level> exact;
      PeelCtl hc;
      CHK(run_peel(ralive.p, na, 0, lv.level, grid, na, &exact, &hc));
      HIPCHK(refine_restore(n, core.p, deg.p, st));
      uint64_t kl = 0;
      for (const Level &x : exact) {
        kl += x.killed;
        if (x.level < lv.lo || x.level > lv.level)
          return fail(CDS_ERR_INTERNAL, "coarse step [%u, %u] re-peeled to level %u", lv.lo,
                      lv.level, x.level);
      }
      if (kl != lv.killed)
        return fail(CDS_ERR_INTERNAL, "coarse step %u killed %llu of %llu cliques", lv.level,
                    (unsigned long long)kl, (unsigned long long)lv.killed);
      levels.erase(levels.begin() + (ptrdiff_t)i);
      levels.insert(levels.begin() + (ptrdiff_t)i, exact.begin(), exact.end());
      R->peel_refined_steps++;
      return CDS_OK;
    };

    // Coarse steps are refined and exact levels replayed until neither list
    // has work: each changes the table or rho, so the summary is recomputed.
    const bool all_cores = p->want_vertex_arrays != 0;
    auto summarize = [&]() -> cds_status {
      double rho_r = 0;
      for (const auto &r : replayed)
        if (r.second.found)
          rho_r = std::max(rho_r, (double)(int64_t)r.second.best_num /
                                      static_cast<double>(r.second.best_den));
      return summarize_levels(levels, n, T, &PS, all_cores, rho_r);
    };
    CHK(summarize());
    {
      DBuf<uint32_t> pv(&A), pdeg(&A), posmap(&A), rdead(&A), pc(&A);
      DBuf<uint64_t> pioff(&A);
      DBuf<uint8_t> palive(&A);
      DBuf<ReplayOut> rout(&A);
      bool allocated = false;
      double refine_ms = 0;
      const int ev_r0 = tm.mark();
      for (;;) {
        if (!PS.refine.empty()) {
          const auto t0 = std::chrono::steady_clock::now();
          const uint32_t bits_before = R->refine_bits_steps;
          CHK(refine_step(PS.refine.front()));
          HIPCHK(hipStreamSynchronize(st));
          // a bitset re-peel is also its levels' replay: its time stays in
          // the replay time, a grid re-peel's goes to the peel
          if (R->refine_bits_steps == bits_before)
            refine_ms += std::chrono::duration<double, std::milli>(
                             std::chrono::steady_clock::now() - t0)
                             .count();
          CHK(summarize());
          continue;
        }
        size_t ii = SIZE_MAX;
        for (size_t cand : PS.replay) {
          bool done = false;
          for (const auto &r : replayed) done |= r.first == levels[cand].level;
          if (!done) {
            ii = cand;
            break;
          }
        }
        if (ii == SIZE_MAX) break;
        const Level &lv = levels[ii];
        const uint64_t M = lv.removed;
        {
          // already replayed by a bitset re-peel of its coarse step
          auto it = std::find_if(seq_replays.begin(), seq_replays.end(),
                                 [&](const std::pair<uint32_t, ReplayOut> &r) {
                                   return r.first == lv.level;
                                 });
          if (it != seq_replays.end()) {
            R->replay_phases++;
            R->replay_bits_phases++;
            replayed.push_back(*it);
            CHK(summarize());
            continue;
          }
        }
        if (!allocated) {
          HIPCHK(pv.alloc(n));
          HIPCHK(pdeg.alloc(n));
          HIPCHK(posmap.alloc(n));
          HIPCHK(palive.alloc(n));
          HIPCHK(rdead.alloc(T));
          HIPCHK(rdead.zero());
          HIPCHK(pc.alloc(1));
          HIPCHK(pioff.alloc((size_t)n + 1));
          HIPCHK(rout.alloc(1));
          allocated = true;
        }
        DBuf<uint32_t> pinc(&A), partial(&A);
        DBuf<GridBarrier> gbar(&A);
        const uint64_t N_start = n - PS.rm_before[ii];
        const bool heavy = M <= kReplayGridMaxVertices && lv.killed * k >= e->replay_grid_work * M;
        // Few vertices alive, k <= 3: one workgroup on adjacency bitsets
        // (CDS_REPLAY_BITS=0 never, =1 for every level it can take).
        if ((k == 2 || k == 3) && sdeg.p && N_start <= kReplayBitsMaxVertices &&
            (replay_bits_mode() == 1 || (replay_bits_mode() < 0 && heavy))) {
          const uint32_t Nb = (uint32_t)N_start, Wb = (Nb + 63) / 64;
          DBuf<unsigned long long> bits(&A);
          HIPCHK(bits.alloc((size_t)Nb * Wb));
          HIPCHK(hipEventRecord(tm.ev[22], st));
          HIPCHK(replay_phase_bits(k, lv.level, n, L.doff.p, L.dadj.p, core.p, sdeg.p,
                                   T - PS.kl_before[ii], Nb, (uint32_t)M, pv.p, posmap.p,
                                   pc.p, bits.p, rout.p, tm.ev[23], (uint64_t)PS.kp,
                                   replay_nstar(k, (uint64_t)PS.kp), st));
          const auto tb = std::chrono::steady_clock::now();
          ReplayOut ro;
          CHK(readback(e, &ro, rout.p, 1));
          float prep = 0;
          (void)hipEventElapsedTime(&prep, tm.ev[22], tm.ev[23]);
          R->t_replay_prep_ms += prep;
          if (peel_trace_on())
            fprintf(stderr,
                    "[peel] bitset replay of level %u: %u of %llu removals (%s), %u alive, "
                    "prep %.3f ms, replay+wait %.3f ms\n",
                    lv.level, ro.steps, (unsigned long long)M, ro.stopped ? "stopped" : "all",
                    Nb, prep,
                    std::chrono::duration<double, std::milli>(
                        std::chrono::steady_clock::now() - tb)
                        .count());
          R->replay_phases++;
          R->replay_bits_phases++;
          if (ro.steps > M || (!ro.stopped && ro.steps != M))
            return fail(CDS_ERR_INTERNAL, "bitset replay of level %u ran %u of %llu steps",
                        lv.level, ro.steps, (unsigned long long)M);
          replayed.push_back({lv.level, ro});
          CHK(summarize());
          continue;
        }
        // Few vertices with many cliques each: replay on the whole GPU.
        unsigned rgrid = 0;
        if (heavy)
          rgrid = coop_width(e, replay_grid_blocks(e->device, lv.killed * k / M,
                                                   e->cur.replay_blocks));
        if (rgrid) {
          HIPCHK(partial.alloc((size_t)rgrid * M));
          HIPCHK(gbar.alloc(1));
          R->replay_grid_phases++;
        } else {
          HIPCHK(pinc.alloc(lv.killed * k * k));
        }
        std::optional<CoopGate> gate;
        if (rgrid) gate.emplace(e, rgrid);
        HIPCHK(hipEventRecord(tm.ev[22], st));
        const unsigned rasked = rgrid;
        CHK(ensure_kill());
        HIPCHK(replay_phase(k, lv.level, n, core.p, kill.p, ioff.p, inc.p,
                            T - PS.kl_before[ii], n - PS.rm_before[ii], (uint32_t)M,
                            pv.p, pdeg.p, posmap.p, palive.p, rdead.p, pc.p,
                            pioff.p, pinc.p, rout.p, partial.p, gbar.p, &rgrid,
                            sdeg.p, A, tm.ev[23], (uint64_t)PS.kp,
                            replay_nstar(k, (uint64_t)PS.kp), st));
        gate.reset();
        if (rasked) note_width(R, &R->replay_grid, rasked, rgrid);
        ReplayOut ro;
        CHK(readback(e, &ro, rout.p, 1));
        if (rgrid > 1) {
          unsigned ab = 0;
          CHK(readback(e, &ab, &gbar.p->abort, 1));
          if (ab) return barrier_failed("whole-GPU replay", rgrid);
        }
        {
          float prep = 0;
          (void)hipEventElapsedTime(&prep, tm.ev[22], tm.ev[23]);
          R->t_replay_prep_ms += prep;
        }
        R->replay_phases++;
        if (ro.steps > M || (!ro.stopped && ro.steps != M))
          return fail(CDS_ERR_INTERNAL, "replay of level %u ran %u of %llu steps",
                      lv.level, ro.steps, (unsigned long long)M);
        replayed.push_back({lv.level, ro});
        // a replay can raise k': the refine list (steps straddling k') and
        // the replay candidates are recomputed against the new value
        CHK(summarize());
      }
      const int ev_r1 = tm.mark();
      HIPCHK(hipStreamSynchronize(st));
      R->t_replay_ms = std::max(0.0, tm.ms(ev_r0, ev_r1) - refine_ms);
    }
  }
  const std::vector<uint64_t> &rm_before = PS.rm_before, &kl_before = PS.kl_before;
  R->peel_levels = peel_rounds;
  R->peel_distinct_levels = (uint32_t)levels.size();
  {
    uint64_t alive = n, scanned = 0;
    for (auto &lv : levels) {
      scanned += alive;
      alive -= lv.removed;
    }
    R->peel_alive_scanned = scanned;
  }
  R->peel_max_density = PS.rho;
  const int64_t kp = PS.kp;
  R->k_prime = (uint64_t)kp;
  const uint64_t kmax = PS.kmax;
  R->kmax = kmax;
  const int ev_peel = tm.mark();

  // ---- densest core ----
  size_t core_level = levels.size();
  for (size_t i = 0; i < levels.size(); ++i)
    if ((int64_t)levels[i].level >= kp) {
      core_level = i;
      break;
    }
  if (core_level == levels.size())
    return fail(CDS_ERR_INTERNAL, "empty densest core (k'=%lld)", (long long)kp);
  const uint32_t core_n = (uint32_t)(n - rm_before[core_level]);
  const uint64_t core_T = T - kl_before[core_level];
  R->core_size = core_n;
  R->core_cliques = core_T;
  R->deleted_cliques = T - core_T;
  R->core_density = (double)(int64_t)core_T / static_cast<double>(core_n);
  const uint32_t kpu = (uint32_t)kp;
  DBuf<unsigned long long> acc(&A);
  HIPCHK(acc.alloc(2));
  // the core's vertices (cv, numbered by cidx) and its cliques (cc) as
  // lists, so the core stages below visit them instead of every vertex
  // and clique of the graph
  DBuf<uint32_t> cv(&A), cidx(&A), ccount1(&A), cc(&A);
  HIPCHK(cv.alloc(core_n));
  HIPCHK(cidx.alloc(n));
  HIPCHK(ccount1.alloc(1));
  HIPCHK(core_list(n, core.p, kpu, cidx.p, cv.p, core_n, ccount1.p, st));
  {
    uint32_t listed = 0;
    CHK(readback(e, &listed, ccount1.p, 1));
    if (listed != core_n)
      return fail(CDS_ERR_INTERNAL, "core list holds %u of %u vertices", listed, core_n);
  }
  HIPCHK(core_edge_count(core_n, g->off, g->adj, core.p, kpu, acc.p, st, cv.p));
  unsigned long long ce = 0;
  CHK(readback(e, &ce, acc.p, 1));
  R->core_edges = ce / 2;
  // every clique in the densest core (the graph is its own densest core,
  // as a pruned core often is): the core stages take all cliques, with no
  // list and no kill words
  const bool all_core = core_T == T;
  const bool list_cc = !core_only && core_T < 0xFFFFFFFFull && !all_core;
  const uint32_t *ckill = kill.p;
  if (all_core) {
    ckill = nullptr;
  } else if (list_cc) {
    HIPCHK(cc.alloc(core_T));
    HIPCHK(core_cliques_ranged(core_n, cv.p, k, L.doff.p, L.item_off.p, L.item_base.p,
                               L.cliques.p, core.p, kpu, cc.p, acc.p, A));
  } else {
    CHK(ensure_kill());  // the stages below read kill words
    HIPCHK(core_cliques(T, k, L.cliques.p, core.p, kill.p, kpu, nullptr, acc.p, st));
  }
  unsigned long long ct = T;
  if (!all_core) CHK(readback(e, &ct, acc.p, 1));
  if (ct != core_T)
    return fail(CDS_ERR_INTERNAL, "core clique count %llu != %llu", ct,
                (unsigned long long)core_T);
  const int ev_core = tm.mark();
  if (core_only) {
    if (p->want_vertex_arrays) {
      R->clique_degree = host_alloc<uint64_t>(n);
      memcpy(R->clique_degree, cd.data(), n * 8);
      R->core = host_alloc<uint32_t>(n);
      CHK(readback(e, R->core, core.p, n));
    }
    HIPCHK(hipStreamSynchronize(st));
    R->times.dag = tm.ms(ev_start, ev_dag);
    R->times.listing = tm.ms(ev_dag, ev_list);
    R->times.core_decompose = tm.ms(ev_list, ev_peel);
    R->times.densest_core = tm.ms(ev_peel, ev_core);
    R->times.total = std::chrono::duration<double, std::milli>(
                         std::chrono::steady_clock::now() - wall0)
                         .count();
    R->peak_device_bytes = A.peak_bytes();
    return CDS_OK;
  }

  // Flow networks of 2^32 clique-member slots or more run the wide max flow
  // (clique-id incidence, 32-bit residuals); clique ids stay 32-bit, so a
  // core beyond that is refused here, before pruning and bucketing it.
  // CDS_FLOW_WIDE=1 sends every network to the wide kernel (tests).
  const char *fwide = getenv("CDS_FLOW_WIDE");
  const bool wide = (fwide && atoi(fwide) != 0) || core_T * k >= 0xFFFFFFFFull;
  if (core_T >= 0xFFFFFFFFull)
    return fail(CDS_ERR_RANGE, "%llu core cliques exceed 32-bit ids",
                (unsigned long long)core_T);
  // ---- prune ----
  DBuf<uint8_t> mark(&A);
  {
    const size_t na = g->host_off[n];
    HIPCHK(mark.alloc(na));
    HIPCHK(mark.zero());
    if (core_n <= kDenseCoreMax && k <= 4) {
      DBuf<unsigned long long> cbits(&A);
      HIPCHK(cbits.alloc((size_t)core_n * ((core_n + 63) / 64)));
      HIPCHK(prune_mark_bits(k, cv.p, core_n, cidx.p, g->off, g->adj, cbits.p, mark.p, st));
    } else if (core_n <= kDenseCoreMax) {
      DBuf<uint32_t> cmat(&A);
      HIPCHK(cmat.alloc((size_t)core_n * core_n));
      HIPCHK(prune_mark_dense(core_T, k, L.cliques.p, cc.p, cv.p, core_n, cidx.p, g->off,
                              g->adj, cmat.p, mark.p, st));
    } else {
      HIPCHK(prune_mark(core_T, k, L.cliques.p, ckill, kpu, g->off, g->adj, mark.p, st,
                        cc.p));
    }
    HIPCHK(count_nonzero_u8(na, mark.p, acc.p, st));
    unsigned long long pe = 0;
    CHK(readback(e, &pe, acc.p, 1));
    R->pruned_edges = pe / 2;
  }
  const int ev_prune = tm.mark();

  // ---- components ----
  DBuf<uint32_t> parent(&A);
  HIPCHK(parent.alloc(n));
  HIPCHK(uf_init(n, parent.p, st));
  // components of the pruned graph = classes of "shares a core clique";
  // only the core vertices are visited, read back and numbered
  HIPCHK(uf_union_edges(core_n, g->off, g->adj, mark.p, core.p, kpu, parent.p, st, cv.p));
  HIPCHK(uf_flatten(n, parent.p, st));
  std::vector<uint32_t> hcv(core_n), hpar(core_n);
  {
    DBuf<uint32_t> cpar(&A);
    HIPCHK(cpar.alloc(core_n));
    HIPCHK(gather_u32(core_n, cv.p, parent.p, cpar.p, st));
    CHK(readback(e, hcv.data(), cv.p, core_n));
    CHK(readback(e, hpar.data(), cpar.p, core_n));
  }
  // core vertices ascending: components in order of their smallest vertex
  // (every root is its component's smallest vertex), flow ids ascending
  // within each component
  std::vector<uint32_t> ord(core_n);
  for (uint32_t i = 0; i < core_n; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return hcv[a] < hcv[b]; });
  std::vector<uint32_t> comp_of(core_n, 0), fv_of(core_n, 0), fv_orig(core_n);
  std::vector<Component> comps;
  {
    std::unordered_map<uint32_t, uint32_t> root_comp;
    root_comp.reserve(core_n);
    for (uint32_t r = 0; r < core_n; ++r) {
      const uint32_t i = ord[r], v = hcv[i];
      if (hpar[i] == v) {
        root_comp[v] = (uint32_t)comps.size();
        Component c;
        c.first = v;
        comps.push_back(c);
      }
      const auto it = root_comp.find(hpar[i]);
      if (it == root_comp.end())
        return fail(CDS_ERR_INTERNAL, "core vertex %u has root %u outside the core", v,
                    hpar[i]);
      comp_of[i] = it->second;
      comps[it->second].V++;
    }
    uint32_t base = 0;
    for (auto &c : comps) {
      c.vbase = base;
      base += c.V;
    }
    if (base != core_n)
      return fail(CDS_ERR_INTERNAL, "component sizes %u != core %u", base, core_n);
    std::vector<uint32_t> fill(comps.size(), 0);
    for (uint32_t r = 0; r < core_n; ++r) {
      const uint32_t i = ord[r];
      const uint32_t ci = comp_of[i];
      fv_of[i] = comps[ci].vbase + fill[ci]++;
      fv_orig[fv_of[i]] = hcv[i];
    }
  }
  const uint32_t ncomp = (uint32_t)comps.size();
  R->num_components = ncomp;
  DBuf<uint32_t> d_comp_of(&A), d_fv_of(&A), ccount(&A), fcd(&A), ccur(&A),
      fmem(&A), ficur(&A), finc(&A);
  DBuf<uint64_t> coff(&A), fioff(&A);
  HIPCHK(d_comp_of.alloc(n));
  HIPCHK(d_fv_of.alloc(n));
  {
    // by vertex id, set for the core vertices only (the only ones read)
    DBuf<uint32_t> lcomp(&A), lfv(&A);
    HIPCHK(lcomp.alloc(core_n));
    HIPCHK(lfv.alloc(core_n));
    CHK(upload(e, lcomp.p, comp_of.data(), core_n));
    CHK(upload(e, lfv.p, fv_of.data(), core_n));
    HIPCHK(scatter_u32x2(core_n, cv.p, lcomp.p, lfv.p, d_comp_of.p, d_fv_of.p, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  HIPCHK(fcd.alloc(core_n));
  std::vector<uint32_t> hcc(ncomp), hfcd(core_n, 0);
  std::vector<uint64_t> hfioff(core_n + 1, 0);
  if (all_core) {
    // every clique is a core clique: a flow vertex's slots are its clique
    // degree, and a component's cliques a k-th of their sum
    HIPCHK(core_flow_degrees(core_n, cv.p, L.cd.p, d_fv_of.p, fcd.p, st));
    CHK(readback(e, hfcd.data(), fcd.p, core_n));
    for (uint32_t c = 0; c < ncomp; ++c) {
      uint64_t s = 0;
      for (uint32_t f = comps[c].vbase; f < comps[c].vbase + comps[c].V; ++f) s += hfcd[f];
      if (s % k)
        return fail(CDS_ERR_INTERNAL, "component %u holds %llu clique slots", c,
                    (unsigned long long)s);
      hcc[c] = (uint32_t)(s / k);
    }
  } else {
    HIPCHK(ccount.alloc(ncomp));
    HIPCHK(ccount.zero());
    HIPCHK(comp_clique_totals(core_T, k, L.cliques.p, ckill, kpu, d_comp_of.p, ncomp,
                              ccount.p, st, cc.p));
    CHK(readback(e, hcc.data(), ccount.p, ncomp));
  }
  std::vector<uint64_t> hcoff(ncomp + 1, 0);
  for (uint32_t c = 0; c < ncomp; ++c) {
    comps[c].T = hcc[c];
    comps[c].cbase = hcoff[c];
    hcoff[c + 1] = hcoff[c] + hcc[c];
  }
  if (hcoff[ncomp] != core_T)
    return fail(CDS_ERR_INTERNAL, "bucketed %llu of %llu core cliques",
                (unsigned long long)hcoff[ncomp], (unsigned long long)core_T);
  HIPCHK(coff.alloc(ncomp + 1));
  CHK(upload(e, coff.p, hcoff.data(), ncomp + 1));
  HIPCHK(ccur.alloc(ncomp));
  HIPCHK(ccur.zero());
  // The core's cliques as flow records: per clique ms words, its k member
  // flow-vertex ids then the k residual words of the implicit-clique max
  // flow (the explicit kernels read the members with the same stride).
  const uint32_t ms = imp_record_words(k);
  HIPCHK(fmem.alloc(core_T * ms));
  HIPCHK(comp_cliques_build(core_T, k, L.cliques.p, ckill, kpu, d_comp_of.p, d_fv_of.p,
                            ncomp, coff.p, ccur.p, fmem.p, nullptr, st, ms, cc.p));
  cc.reset();
  // flow incidence: per core vertex, its (clique slot, position) pairs
  if (!all_core) {
    HIPCHK(flow_incidence_degrees(core_T, k, fmem.p, core_n, fcd.p, st, ms));
    CHK(readback(e, hfcd.data(), fcd.p, core_n));
  }
  for (uint32_t f = 0; f < core_n; ++f) hfioff[f + 1] = hfioff[f] + hfcd[f];
  if (hfioff[core_n] != core_T * k)
    return fail(CDS_ERR_INTERNAL, "flow incidence holds %llu of %llu pairs",
                (unsigned long long)hfioff[core_n], (unsigned long long)(core_T * k));
  HIPCHK(fioff.alloc((size_t)core_n + 1));
  CHK(upload(e, fioff.p, hfioff.data(), core_n + 1));
  HIPCHK(finc.alloc(core_T * k));
  HIPCHK(ficur.alloc(core_n));
  HIPCHK(flow_incidence_fill(core_T, k, fmem.p, fioff.p, core_n, ficur.p, finc.p, wide, st,
                             ms));
  HIPCHK(hipStreamSynchronize(st));
  ficur.reset();
  if (wide) {
    // room for the flow state: the full clique list and kill levels are
    // not read again (the answer's cliques are counted over fmem)
    L.cliques.reset();
    kill.reset();
  }
  const int ev_cc = tm.mark();

  // ---- (c) dynamicExact (helpers.cpp:1120-1254) ----
  ExactSearch search(k, p->ub1 != 0, p->ub2 != 0, comps, hfcd.data(),
                     R->core_density, kmax);
  R->paper_lower_bound = search.paper_lower_bound;
  uint64_t sweeps = 0;
  // flow state, sized for the largest component
  uint32_t maxV = 0;
  uint64_t maxT = 0;
  for (auto &c : comps) {
    maxV = std::max(maxV, c.V);
    maxT = std::max(maxT, c.T);
  }
  DBuf<int> freach(&A);
  bool flow_alloc = false;
  DBuf<long long> gx(&A), grvs(&A), gexv(&A), grem(&A), gexc(&A);
  DBuf<int> gx32(&A), gexc32(&A);
  DBuf<int> ghv(&A), ghmin(&A), ghc(&A);
  // Implicit clique nodes (the core's clique records hold member ids and
  // residuals; no clique heights, excess or work lists) unless
  // CDS_FLOW_IMPLICIT=0; a flow whose unit is too large for 32-bit
  // residuals runs the explicit network (int64 residual, height and excess
  // arrays, clique work lists).
  const char *fimp = getenv("CDS_FLOW_IMPLICIT");
  const bool imp_ok = !(fimp && atoi(fimp) == 0);
  DBuf<long long> ginflow(&A);
  DBuf<int> gpend(&A);
  bool state_alloc = false;
  DBuf<uint32_t> gl0(&A), gl1(&A), gq0(&A), gq1(&A);
  DBuf<uint64_t> glc0(&A), glc1(&A), garc(&A);
  DBuf<uint32_t> gastop(&A);
  DBuf<GridFlowCtl> gctl(&A);
  std::vector<int> reach_h;
  // Exact max flow of component ci with sink capacity capT (all
  // capacities in units of 1/U); leaves the min-cut side in freach.
  auto run_flow = [&](uint32_t ci, long long capT, long long U,
                      long long *F) -> cds_status {
    const Component &C = comps[ci];
    uint32_t maxcd = 0;
    for (uint32_t v = 0; v < C.V; ++v) maxcd = std::max(maxcd, hfcd[C.vbase + v]);
    if ((__int128)maxcd * U >= ((__int128)1 << 62) ||
        (__int128)capT * C.V >= ((__int128)1 << 62) ||
        (__int128)C.T * k * U >= ((__int128)1 << 62))
      return fail(CDS_ERR_RANGE, "component %u exceeds the int64 flow range", ci);
    // wide kernel: a slot's residual stays in [-U, (k-1)U] and a clique's
    // excess below k(k-1)U, both kept in 32 bits
    if (wide && (__int128)k * k * U >= ((__int128)1 << 31))
      return fail(CDS_ERR_RANGE, "component %u: unit %lld too large for 32-bit residuals",
                  ci, U);
    FlowProblem P;
    P.k = k;
    P.V = C.V;
    P.T = (uint32_t)C.T;
    P.vbase = C.vbase;
    P.sbase = C.cbase * k;
    P.mem = fmem.p + C.cbase * ms;
    P.mstride = ms;
    P.ioff = fioff.p + C.vbase;
    P.inc = finc.p;
    P.cd = fcd.p + C.vbase;
    // push-relabel on as many workgroups as the network needs (one for a
    // small component); state sized for the largest component, allocated
    // at the first flow of the solve that needs it.  An implicit-clique
    // flow that reaches its round limit is solved again on the explicit
    // network (the result counts the flows each one finished).
    for (int attempt = 0;; ++attempt) {
    const bool use_imp =
        imp_ok && attempt == 0 && (__int128)k * k * U < ((__int128)1 << 31);
    P.by_clique = wide ? 1u : 0u;
    if (use_imp && !ginflow.p) {
      HIPCHK(ginflow.alloc(maxV));
      HIPCHK(gpend.alloc(maxV));
    }
    if (!use_imp && !wide && !gx.p) {
      HIPCHK(gx.alloc(maxT * k));
      HIPCHK(gexc.alloc(maxT));
      HIPCHK(ghc.alloc(maxT));
    }
    if (!use_imp && wide && !gx32.p) {
      HIPCHK(gx32.alloc(maxT * k));
      HIPCHK(gexc32.alloc(maxT));
      HIPCHK(ghc.alloc(maxT));
    }
    if (!use_imp && !gq0.p) {
      HIPCHK(gq0.alloc(maxT));
      HIPCHK(gq1.alloc(maxT));
    }
    if (!state_alloc) {
      state_alloc = true;
      HIPCHK(grvs.alloc(maxV));
      HIPCHK(gexv.alloc(maxV));
      HIPCHK(grem.alloc(maxV));
      HIPCHK(ghv.alloc(maxV));
      HIPCHK(ghmin.alloc(maxV));
      HIPCHK(gl0.alloc(maxV));
      HIPCHK(gl1.alloc(maxV));
      HIPCHK(glc0.alloc(maxV));
      HIPCHK(glc1.alloc(maxV));
      HIPCHK(garc.alloc(maxV));
      HIPCHK(gastop.alloc(maxV));
      HIPCHK(gctl.alloc(1));
    }
    GridFlowState G;
    G.x = gx.p;
    G.x32 = gx32.p;
    G.exc32 = gexc32.p;
    G.rec32 = use_imp ? fmem.p + C.cbase * ms : nullptr;
    G.inflow = ginflow.p;
    G.pend = gpend.p;
    G.rvs = grvs.p;
    G.exv = gexv.p;
    G.rem = grem.p;
    G.exc = gexc.p;
    G.hv = ghv.p;
    G.hmin = ghmin.p;
    G.hc = ghc.p;
    G.list0 = gl0.p;
    G.list1 = gl1.p;
    G.lchunk0 = glc0.p;
    G.lchunk1 = glc1.p;
    G.cq0 = gq0.p;
    G.cq1 = gq1.p;
    G.ctl = gctl.p;
    G.reach = freach.p;
    G.arc = garc.p;
    G.astop = gastop.p;
    // CDS_FLOW_TRACE=1 prints one line per push-relabel round to stderr
    DBuf<unsigned long long> ftrace(&A);
    if (flow_trace_on()) {
      HIPCHK(ftrace.alloc(8ull * kFlowTraceRounds));
      G.trace = ftrace.p;
      G.trace_cap = kFlowTraceRounds;
    }
    unsigned launched = 0;
    {
      const unsigned fgrid =
          coop_width(e, flow_grid_blocks(P, e->device, flow_blocks_of(e), wide, use_imp));
      if (fgrid == 0) return fail(CDS_ERR_HIP, "the max flow cannot be launched");
      CoopGate gate(e, fgrid);
      launched = fgrid;
      HIPCHK(hipEventRecord(tm.ev[22], st));
      uint32_t imp_rounds = kImpFlowRounds;
      if (const char *ir = getenv("CDS_IMP_FLOW_ROUNDS")) imp_rounds = (uint32_t)atoi(ir);
      HIPCHK(grid_max_flow(P, G, capT, U, use_imp ? imp_rounds : kGridFlowRounds,
                           &launched, st));
      HIPCHK(hipEventRecord(tm.ev[23], st));
      note_width(R, &R->flow_grid, fgrid, launched);
    }
    GridFlowCtl hc;
    CHK(readback(e, &hc, gctl.p, 1));
    if (hc.bar.abort) return barrier_failed("whole-network max flow", launched);
    R->t_flow_kernel_ms += tm.ms(22, 23);
    R->flow_clique_visits += hc.clique_visits;
    R->flow_arc_visits += hc.chunk_visits * kFlowChunkArcs;
    R->flow_relabels += hc.relabels;
    if (hc.limit) {
      if (flow_trace_on())
        fprintf(stderr, "[flow] component %u V=%u T=%u: %s flow hit %u rounds\n", ci, P.V,
                P.T, use_imp ? "implicit" : "explicit", hc.rounds);
      if (use_imp) continue;
      return fail(CDS_ERR_NOT_CONVERGED,
                  "grid max flow of component %u hit %u rounds", ci, hc.rounds);
    }
    sweeps += hc.rounds;
    R->flow_launches++;
    R->flow_implicit += use_imp ? 1 : 0;
    *F = (long long)hc.sink;
    if (G.trace) {
      const uint32_t nr = std::min(hc.rounds, kFlowTraceRounds);
      std::vector<unsigned long long> tr(8ull * nr);
      CHK(readback(e, tr.data(), G.trace, tr.size()));
      fprintf(stderr, "[flow] component %u V=%u T=%u rounds=%u relabels=%u\n", ci,
              P.V, P.T, hc.rounds, hc.relabels);
      int khz = 100000;
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->device);
      const double us = 1e3 / (khz > 0 ? khz : 100000);  // microseconds per tick
      for (uint32_t r = 0; r < nr; ++r) {
        const unsigned long long *t = &tr[8ull * r];
        fprintf(stderr,
                "[flow] round %u cliques %llu vertices %llu chunks %llu active %llu "
                "us %.1f %.1f %.1f %.1f%s\n",
                r, t[0], t[1], t[2], t[3] & ~(1ull << 63), t[4] * us, t[5] * us,
                t[6] * us, t[7] * us, (t[3] >> 63) ? " (after relabel)" : "");
      }
    }
    return CDS_OK;
    }  // attempts
  };
  {
    std::vector<FlowRequest> round;
    cds_status serr = CDS_OK;
    while (search.next(&round, &serr)) {
      if (!flow_alloc) {
        HIPCHK(freach.alloc(maxV));
        flow_alloc = true;
      }
      for (size_t i = 0; i < round.size(); ++i) {
        const FlowRequest &rq = round[i];
        long long F = 0;
        CHK(run_flow(rq.comp, rq.capT, rq.U, &F));
        if (search.feed_flow(i, F)) {
          const Component &C = comps[rq.comp];
          reach_h.resize(C.V);
          CHK(readback(e, reach_h.data(), freach.p, C.V));
          HIPCHK(count_cliques_in_reach((uint32_t)C.T, k, fmem.p + C.cbase * ms,
                                        C.vbase, freach.p, acc.p, st, ms));
          unsigned long long cc = 0;
          CHK(readback(e, &cc, acc.p, 1));
          CHK(search.feed_reach(i, reach_h.data(), cc));
        }
      }
    }
    if (serr != CDS_OK) return serr;
  }
  R->ub1_val = search.ub1_val;
  R->ub2_val = search.ub2_val;
  R->flow_networks_solved = search.flows_solved;
  R->density_certificates = search.certificates;
  R->flow_sweeps = sweeps;
  const int ev_exact = tm.mark();

  // ---- result ----
  R->density = (float)search.best_density;
  R->size = search.best_size;
  std::vector<uint32_t> verts;
  // the set lies in one component of the densest core: its flow-vertex ids
  std::vector<uint8_t> hf(core_n, 0);
  {
    const Component &C = comps[search.best_comp];
    for (uint32_t v = 0; v < C.V; ++v)
      if (search.best_res.empty() || search.best_res[v] != -1) {
        verts.push_back(fv_orig[C.vbase + v]);
        hf[C.vbase + v] = 1;
      }
  }
  R->set_size = (uint32_t)verts.size();
  R->vertices = host_alloc<uint32_t>(verts.size());
  memcpy(R->vertices, verts.data(), verts.size() * 4);
  {
    // k-cliques inside the reported set (exact numerator): they are core
    // cliques, so they are counted over the core's clique records
    // (a whole component holds exactly its own cliques)
    const Component &C = comps[search.best_comp];
    unsigned long long sc = C.T;
    if (verts.size() != C.V) {
      DBuf<uint8_t> mask(&A);
      HIPCHK(mask.alloc(core_n));
      CHK(upload(e, mask.p, hf.data(), core_n));
      HIPCHK(count_cliques_in_mask(core_T, k, fmem.p, mask.p, acc.p, st, ms));
      CHK(readback(e, &sc, acc.p, 1));
    }
    R->set_cliques = sc;
  }
  fill_trace(R, search, k, comps);
  if (p->want_vertex_arrays) {
    R->clique_degree = host_alloc<uint64_t>(n);
    memcpy(R->clique_degree, cd.data(), n * 8);
    R->core = host_alloc<uint32_t>(n);
    CHK(readback(e, R->core, core.p, n));
  }
  HIPCHK(hipStreamSynchronize(st));
  R->times.dag = tm.ms(ev_start, ev_dag);
  R->times.listing = tm.ms(ev_dag, ev_list);
  R->times.core_decompose = tm.ms(ev_list, ev_peel);
  R->times.densest_core = tm.ms(ev_peel, ev_core);
  R->times.prune = tm.ms(ev_core, ev_prune);
  R->times.components = tm.ms(ev_prune, ev_cc);
  R->times.exact = tm.ms(ev_cc, ev_exact);
  R->times.total = std::chrono::duration<double, std::milli>(
                       std::chrono::steady_clock::now() - wall0)
                       .count();
  R->peak_device_bytes = A.peak_bytes();
  return CDS_OK;
}

// ---------------------------------------------------------------------------
// Batched solve of B graphs with the same (k, ub1, ub2) as one disjoint
// union: one DAG, listing, incidence and persistent peel over the union;
// per-graph level tables from a (graph, core) sort; every replay phase of
// every graph in one launch; per-graph k' applied through 0/1 flags so the
// prune, union-find and flow-network kernels run once with threshold 1;
// then one ExactSearch per graph, advanced in lockstep, with all pending
// flows of a step batched (one workgroup per small component in one plain
// launch, larger components on several workgroups each in cooperative
// launches).  Results equal cds_solve on each graph.
// ---------------------------------------------------------------------------

// A device graph built here (an induced subgraph); frees its CSR.
struct OwnedDgraph {
  cds_dgraph g;
  DBuf<uint32_t> off, adj;
  explicit OwnedDgraph(DeviceArena *a) : off(a), adj(a) {}
};

// The pruned path: the graph is cut to its tau-clique-core S by clique-count
// passes over induced subgraphs (prune.hip: every vertex with fewer than
// tau cliques in the current S goes, until none does; then tau rises to
// ceil(T(S) / |S|) while that grows), and the whole solve runs on G[S].
// Every S at a point where nothing drops is a boundary state of the
// oracle's peel (its tau-clique-core), so tau <= k' throughout, and the
// peel rows above S are at most max(rho(S), tau - 1): k', kmax, the densest
// core and everything after it are those of G[S].  The count L0 of the
// whole graph gives T and the clique degrees.
cds_status solve_pruned(cds_engine *e, const cds_dgraph *g, const cds_params *p,
                        cds_result *R, bool core_only, Listing &L0) {
  const auto wall0 = std::chrono::steady_clock::now();
  DeviceArena &A = *e->arena;
  hipStream_t st = e->stream;
  const uint32_t n = g->n, k = p->k;
  const uint64_t T = L0.T;
  DBuf<uint8_t> keep(&A), state(&A);
  DBuf<uint32_t> newid(&A), cnt(&A), sel(&A), dropped(&A), dlist(&A), ddeg(&A), dpoff(&A);
  DBuf<unsigned long long> dstats(&A);
  DBuf<uint64_t> cd_first(&A);  // the whole graph's counts, when dropped from in place
  DBuf<uint64_t> cd_bits(&A);   // counts of G[S] from its adjacency bitsets
  uint64_t *cdw = nullptr;      // the counts of G[S] that drops in place update
  HIPCHK(keep.alloc(n));
  HIPCHK(hipMemsetAsync(keep.p, 1, n, st));
  HIPCHK(newid.alloc((size_t)n + 1));
  HIPCHK(cnt.alloc(n));
  HIPCHK(sel.alloc(n));
  HIPCHK(dropped.alloc(1));
  HIPCHK(state.alloc(n));
  HIPCHK(dlist.alloc(n));
  HIPCHK(ddeg.alloc(n));
  HIPCHK(dpoff.alloc((size_t)n + 1));
  HIPCHK(dstats.alloc(3));
  auto gs = std::make_unique<OwnedDgraph>(&A);
  std::unique_ptr<Listing> Ls;
  uint64_t tau = (T + n - 1) / n, Ts = T;
  uint32_t ns = n, alive = n, passes = 0, unlinks = 0;
  const uint64_t *cds = L0.cd.p;
  // Drops in place (CDS_PRUNE_UNLINK=1 whenever possible, 0 never): the
  // cliques through the dropped vertices are taken from the counts of the
  // rest instead of counting the rebuilt subgraph again, when at most
  // 1/16 of the cliques of S pass through them.
  int unlink_mode = -1;
  if (const char *f = getenv("CDS_PRUNE_UNLINK")) unlink_mode = atoi(f) != 0 ? 1 : 0;
  // a G[S] of at most kBitsCountMax vertices (k = 3, 4) is counted on its
  // adjacency bitsets instead of by listing (CDS_PRUNE_BITS=0: never)
  bool count_bits = k == 3 || k == 4;
  if (const char *f = getenv("CDS_PRUNE_BITS")) count_bits = count_bits && atoi(f) != 0;
  // G[S] from keep[]
  auto build_subgraph = [&]() -> cds_status {
    HIPCHK(induced_subgraph(n, g->off, g->adj, keep.p, newid.p, cnt.p, nullptr, nullptr,
                            nullptr, false, A));
    CHK(readback(e, &ns, newid.p + n, 1));
    if (ns == 0) return fail(CDS_ERR_INTERNAL, "core pruning left no vertex (tau %llu)",
                             (unsigned long long)tau);
    auto ng = std::make_unique<OwnedDgraph>(&A);
    HIPCHK(ng->off.alloc((size_t)ns + 1));
    HIPCHK(scan_u32_32(cnt.p, ng->off.p, ns, A));
    ng->g.host_off.resize((size_t)ns + 1);
    CHK(readback(e, ng->g.host_off.data(), ng->off.p, (size_t)ns + 1));
    const uint64_t arcs = ng->g.host_off[ns];
    HIPCHK(ng->adj.alloc(arcs));
    HIPCHK(induced_subgraph(n, g->off, g->adj, keep.p, newid.p, nullptr, ng->off.p,
                            ng->adj.p, sel.p, true, A));
    ng->g.n = ns;
    ng->g.m = arcs / 2;
    ng->g.off = ng->off.p;
    ng->g.adj = ng->adj.p;
    uint32_t md = 0;
    for (uint32_t v = 0; v < ns; ++v)
      md = std::max(md, ng->g.host_off[v + 1] - ng->g.host_off[v]);
    ng->g.max_degree = md;
    gs = std::move(ng);
    alive = ns;
    HIPCHK(hipMemsetAsync(state.p, 1, ns, st));
    return CDS_OK;
  };
  bool first = true;
  Timer tm(st);
  for (bool rebuild = true;;) {
    if (rebuild) {
      CHK(build_subgraph());
      if (!first) {
        // clique counts inside S
        const auto tp0 = std::chrono::steady_clock::now();
        Ls.reset();
        const uint64_t arcs = gs->g.host_off[ns];
        const bool bits = count_bits && ns <= kBitsCountMax;
        if (bits) {
          DBuf<unsigned long long> rows(&A), tot(&A);
          DBuf<uint32_t> src(&A);
          HIPCHK(rows.alloc((size_t)ns * ((ns + 63) / 64)));
          HIPCHK(tot.alloc(1));
          HIPCHK(src.alloc(arcs));
          HIPCHK(cd_bits.alloc(ns));
          HIPCHK(bits_clique_count(k, ns, gs->g.off, gs->g.adj, (uint32_t)arcs, rows.p, src.p,
                                   reinterpret_cast<unsigned long long *>(cd_bits.p), tot.p,
                                   st));
          unsigned long long t = 0;
          CHK(readback(e, &t, tot.p, 1));
          Ts = t;
          cdw = cd_bits.p;
        } else {
          cd_bits.reset();
          Ls = std::make_unique<Listing>(&A);
          CHK(run_dag(e, &gs->g, *Ls));
          CHK(run_listing(e, &gs->g, k, false, *Ls, &tm));
          Ts = Ls->T;
          cdw = Ls->cd.p;
        }
        cds = cdw;
        ++passes;
        if (peel_trace_on())
          fprintf(stderr, "[prune] pass %u tau %llu: %u vertices %llu arcs %llu cliques, "
                  "count%s %.2f ms\n", passes, (unsigned long long)tau, ns,
                  (unsigned long long)arcs, (unsigned long long)Ts, bits ? " (bitsets)" : "",
                  std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - tp0).count());
      }
      first = false;
      rebuild = false;
    }
    HIPCHK(hipMemsetAsync(dropped.p, 0, sizeof(uint32_t), st));
    HIPCHK(hipMemsetAsync(dstats.p, 0, 3 * sizeof(unsigned long long), st));
    HIPCHK(prune_select(ns, cds, tau, gs->g.off, state.p, dlist.p, dropped.p, dstats.p, st));
    uint32_t nd = 0;
    unsigned long long stats[2] = {0, 0};
    CHK(readback(e, &nd, dropped.p, 1));
    if (nd) {
      CHK(readback(e, stats, dstats.p, 2));
      const bool in_place = (k == 3 || k == 4) && unlink_mode != 0 &&
                            stats[1] <= kUnlinkMaxDegree &&
                            (unlink_mode == 1 || stats[0] <= Ts / 16);
      if (in_place) {
        const auto tp0 = std::chrono::steady_clock::now();
        if (!cdw) {
          // the whole graph's counts are reported: drop from a copy
          HIPCHK(cd_first.alloc(n));
          HIPCHK(hipMemcpyAsync(cd_first.p, L0.cd.p, (size_t)n * sizeof(uint64_t),
                                hipMemcpyDeviceToDevice, st));
          cds = cdw = cd_first.p;
        }
        HIPCHK(prune_unlink(k, nd, dlist.p, (uint32_t)stats[1], gs->g.off, gs->g.adj, state.p,
                            reinterpret_cast<unsigned long long *>(cdw), dstats.p + 2,
                            ddeg.p, dpoff.p, A));
        HIPCHK(prune_finish(nd, dlist.p, sel.p, state.p, keep.p, st));
        unsigned long long removed = 0;
        CHK(readback(e, &removed, dstats.p + 2, 1));
        if (removed > Ts)
          return fail(CDS_ERR_INTERNAL, "pruning removed %llu of %llu cliques", removed,
                      (unsigned long long)Ts);
        Ts -= removed;
        alive -= nd;
        ++unlinks;
        if (peel_trace_on())
          fprintf(stderr, "[prune] in place tau %llu: %u dropped (max degree %llu), %llu "
                  "cliques removed, %u vertices %llu cliques left, %.2f ms\n",
                  (unsigned long long)tau, nd, stats[1], removed, alive,
                  (unsigned long long)Ts,
                  std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - tp0).count());
        continue;
      }
      HIPCHK(prune_finish(nd, dlist.p, sel.p, state.p, keep.p, st));
      rebuild = true;
      continue;
    }
    // S is the tau-clique-core; its density is a peel row
    const uint64_t t2 = (Ts + alive - 1) / alive;
    if (t2 <= tau) break;
    tau = t2;
  }
  // S lost vertices in place since G[S] was built: build it again
  if (alive != ns) CHK(build_subgraph());
  Ls.reset();
  // the solve on G[S]: persistent peel if its cliques fit the incidence,
  // else the bitset peel of a dense core
  const char *fd = getenv("CDS_DENSE_PEEL");
  bool dense = (unsigned __int128)k * Ts >= 0xFFFFFFFFull || needs_prune(k, Ts);
  if (fd) dense = atoi(fd) != 0;
  if (dense && ns > kDenseMaxVertices)
    return fail(CDS_ERR_RANGE,
                "the %llu-clique-core has %u vertices and %llu cliques: too many to "
                "materialise, too many vertices for the dense peel (at most %u)",
                (unsigned long long)tau, ns, (unsigned long long)Ts, kDenseMaxVertices);
  cds_result RS;
  CHK(solve_impl(e, &gs->g, p, &RS, core_only, dense ? kSolveDenseCore : kSolveCore));
  std::unique_ptr<cds_result, void (*)(cds_result *)> hold(&RS, cds_result_free);
  if (RS.total_cliques != Ts)
    return fail(CDS_ERR_INTERNAL, "the pruned core holds %llu cliques, %llu were kept",
                (unsigned long long)RS.total_cliques, (unsigned long long)Ts);
  // the answer in the input's ids
  std::vector<uint32_t> hsel(ns);
  CHK(readback(e, hsel.data(), sel.p, ns));
  *R = RS;
  memset(&RS, 0, sizeof(RS));  // arrays now owned by R
  R->n = n;
  R->m = g->m;
  R->max_degree = g->max_degree;
  R->total_cliques = T;
  R->deleted_cliques = T - R->core_cliques;
  R->dag_arcs = g->m;
  R->list_work = L0.work;
  R->max_out_degree = L0.max_out;
  R->list_items = L0.items_n;
  R->t_list_count_ms = L0.t_count;
  if (T * k >= (1ull << 31)) R->oracle_quirk_flags |= 4;  // the oracle's int total wraps
  for (uint32_t i = 0; i < R->set_size; ++i) R->vertices[i] = hsel[R->vertices[i]];
  if (p->want_vertex_arrays) {
    std::vector<uint32_t> core(n, kAliveHost);
    if (R->core)
      for (uint32_t i = 0; i < ns; ++i) core[hsel[i]] = R->core[i];
    free(R->core);
    free(R->clique_degree);
    R->core = host_alloc<uint32_t>(n);
    memcpy(R->core, core.data(), (size_t)n * 4);
    R->clique_degree = host_alloc<uint64_t>(n);
    CHK(readback(e, R->clique_degree, L0.cd.p, n));
  }
  R->prune_tau = tau;
  R->prune_vertices = ns;
  R->prune_passes = passes;
  R->prune_unlinks = unlinks;
  R->dense_peel = dense ? 1 : 0;
  R->times.total = std::chrono::duration<double, std::milli>(
                       std::chrono::steady_clock::now() - wall0)
                       .count();
  return CDS_OK;
}

cds_status solve_union(cds_engine *e, const std::vector<const cds_dgraph *> &gs,
                       const std::vector<const cds_params *> &ps,
                       const std::vector<cds_result *> &rs) {
  const auto wall0 = std::chrono::steady_clock::now();
  DeviceArena &A = *e->arena;
  hipStream_t st = e->stream;
  const uint32_t B = (uint32_t)gs.size();
  const uint32_t k = ps[0]->k;
  const bool ub1 = ps[0]->ub1 != 0, ub2 = ps[0]->ub2 != 0;
  for (uint32_t b = 0; b < B; ++b) {
    cds_result *R = rs[b];
    memset(R, 0, sizeof(*R));
    R->k = k;
    R->n = gs[b]->n;
    R->m = gs[b]->m;
    R->max_degree = gs[b]->max_degree;
    if (gs[b]->n == 0) return fail(CDS_ERR_INVALID, "batch item %u: empty graph", b);
  }
  Timer tm(st);
  const int ev_start = tm.mark();

  // ---- the union CSR ----
  std::vector<uint32_t> vbase(B + 1, 0);
  std::vector<uint64_t> abase(B + 1, 0);
  for (uint32_t b = 0; b < B; ++b) {
    vbase[b + 1] = vbase[b] + gs[b]->n;
    abase[b + 1] = abase[b] + gs[b]->host_off[gs[b]->n];
  }
  const uint32_t n = vbase[B];
  if (abase[B] >= 0xFFFFFFFFull || n >= (1u << 24))
    return fail(CDS_ERR_RANGE, "batch union too large (%u vertices)", n);
  cds_dgraph U;
  U.n = n;
  U.m = abase[B] / 2;
  U.host_off.resize((size_t)n + 1);
  for (uint32_t b = 0; b < B; ++b) {
    U.max_degree = std::max(U.max_degree, gs[b]->max_degree);
    for (uint32_t i = 0; i < gs[b]->n; ++i)
      U.host_off[vbase[b] + i] = (uint32_t)abase[b] + gs[b]->host_off[i];
  }
  U.host_off[n] = (uint32_t)abase[B];
  begin_solve(e, k, U.m);
  DBuf<uint32_t> uoff(&A), uadj(&A), gid(&A);
  {
    std::vector<UnionPart> parts(B);
    for (uint32_t b = 0; b < B; ++b)
      parts[b] = UnionPart{gs[b]->off, gs[b]->adj, gs[b]->n, vbase[b],
                           (uint32_t)abase[b], 0};
    DBuf<UnionPart> dparts(&A);
    HIPCHK(dparts.alloc(B));
    HIPCHK(hipMemcpyAsync(dparts.p, parts.data(), B * sizeof(UnionPart),
                          hipMemcpyHostToDevice, st));
    HIPCHK(uoff.alloc((size_t)n + 1));
    HIPCHK(uadj.alloc(abase[B]));
    HIPCHK(gid.alloc(n));
    HIPCHK(union_build(B, dparts.p, n, (uint32_t)abase[B], uoff.p, uadj.p, gid.p, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  U.off = uoff.p;
  U.adj = uadj.p;
  std::vector<uint32_t> hgid(n);
  for (uint32_t b = 0; b < B; ++b)
    std::fill(hgid.begin() + vbase[b], hgid.begin() + vbase[b + 1], b);

  // ---- (a) DAG + listing ----
  Listing L(&A);
  CHK(run_dag(e, &U, L));
  const int ev_dag = tm.mark();
  CHK(run_listing(e, &U, k, true, L, &tm));
  const int ev_list = tm.mark();
  const uint64_t T = L.T;
  std::vector<uint64_t> cd(n);
  CHK(readback(e, cd.data(), L.cd.p, n));
  std::vector<uint64_t> Tb(B, 0);
  {
    uint64_t sum_cd = 0, max_cd = 0;
    for (uint32_t b = 0; b < B; ++b) {
      uint64_t s = 0;
      for (uint32_t v = vbase[b]; v < vbase[b + 1]; ++v) {
        s += cd[v];
        max_cd = std::max(max_cd, cd[v]);
      }
      if (s % k)
        return fail(CDS_ERR_INTERNAL, "batch item %u: clique degrees sum %llu", b,
                    (unsigned long long)s);
      Tb[b] = s / k;
      sum_cd += s;
      if (s >= (1ull << 31)) rs[b]->oracle_quirk_flags |= 4;
    }
    if (sum_cd != T * k)
      return fail(CDS_ERR_INTERNAL, "listing invariant: sum cd %llu != k*T %llu",
                  (unsigned long long)sum_cd, (unsigned long long)(T * k));
    if (max_cd >= 0xFFFFFFFFull)
      return fail(CDS_ERR_RANGE, "a clique degree exceeds 32 bits");
  }
  for (uint32_t b = 0; b < B; ++b) {
    rs[b]->total_cliques = Tb[b];
    rs[b]->dag_arcs = gs[b]->m;
    rs[b]->list_work = L.work;
    rs[b]->max_out_degree = L.max_out;
    rs[b]->list_items = L.items_n;
    rs[b]->t_list_count_ms = L.t_count;
    rs[b]->t_list_write_ms = L.t_write;
  }

  // ---- (b) incidence + peel over the union ----
  DBuf<uint64_t> ioff(&A);
  DBuf<uint32_t> icur(&A), inc(&A), deg(&A), core(&A), kill(&A), sdeg(&A);
  HIPCHK(ioff.alloc((size_t)n + 1));
  HIPCHK(scan_u64(L.cd.p, ioff.p, n, A));
  HIPCHK(icur.alloc(n));
  HIPCHK(icur.zero());
  HIPCHK(inc.alloc(T * k * k));
  HIPCHK(hipEventRecord(tm.ev[22], st));
  HIPCHK(build_incidence(T, k, L.cliques.p, ioff.p, icur.p, inc.p, st));
  HIPCHK(hipEventRecord(tm.ev[23], st));
  // sub-stage kernel times of the union (the same for every graph in it)
  double t_inc = tm.ms(22, 23), t_peel = 0, t_replay = 0;
  icur.reset();
  HIPCHK(deg.alloc(n));
  HIPCHK(core.alloc(n));
  HIPCHK(kill.alloc(T));
  HIPCHK(sdeg.alloc(n));
  HIPCHK(peel_init(n, T, L.cd.p, deg.p, core.p, kill.p, st));
  uint32_t peel_grid = 0, peel_rounds = 0, coop_shrinks = 0, replay_grid = 0;
  {
    DBuf<uint32_t> s32(&A), lvl(&A), lmin(&A);
    DBuf<uint64_t> s64(&A);
    DBuf<unsigned long long> lrem(&A), lkil(&A);
    DBuf<PeelCtl> ctl(&A);
    HIPCHK(s32.alloc(6ull * n));
    HIPCHK(s64.alloc(3ull * n));
    HIPCHK(lvl.alloc(n));
    HIPCHK(lmin.alloc(n));
    HIPCHK(lrem.alloc(n));
    HIPCHK(lkil.alloc(n));
    HIPCHK(ctl.alloc(1));
    DBuf<uint32_t> stamp(&A);
    HIPCHK(stamp.alloc(n));
    // A union peels exact levels of every graph at once, thousands of
    // barrier-ended sub-rounds: the width tuned for one graph of the same
    // edge class does not transfer.  The built-in rule capped at 64
    // workgroups measured best (4096-graph mix: k=3 unions 79 / 55 ms with
    // 64 workgroups, 112 / 85 ms with 256).
    peel_grid = coop_width(e, std::min(peel_grid_blocks(n, T, k, e->device, 0), 64u));
    if (peel_grid == 0) return fail(CDS_ERR_HIP, "the persistent peel cannot be launched");
    {
      CoopGate gate(e, peel_grid);
      HIPCHK(hipEventRecord(tm.ev[22], st));
      const unsigned asked = peel_grid;
      HIPCHK(peel_persistent(n, k, T, ioff.p, inc.p, L.cliques.p, deg.p, core.p,
                             kill.p, stamp.p, 1, s32.p, s64.p, ctl.p, lvl.p, lmin.p, lrem.p,
                             lkil.p,
                             sdeg.p, nullptr, 0, 0, kAliveHost, &peel_grid, st));
      HIPCHK(hipEventRecord(tm.ev[23], st));
      if (peel_grid < asked) coop_shrinks++;
    }
    PeelCtl hc;
    CHK(readback(e, &hc, ctl.p, 1));
    if (hc.bar.abort) return barrier_failed("persistent peel", peel_grid);
    t_peel = tm.ms(22, 23);
    HIPCHK(peel_derive_kill(T, k, L.cliques.p, core.p, kill.p, st));
    if (hc.removed != n)
      return fail(CDS_ERR_INTERNAL, "persistent peel removed %llu of %u", hc.removed, n);
    peel_rounds = (uint32_t)hc.rounds;
  }
  deg.reset();

  // ---- per-graph level tables: runs of (graph, core) ----
  DBuf<uint64_t> lkey(&A), skey(&A), ukey(&A);
  DBuf<uint32_t> lval(&A), sval(&A), rcount(&A), nruns_d(&A), rkilled(&A);
  HIPCHK(lkey.alloc(n));
  HIPCHK(lval.alloc(n));
  HIPCHK(skey.alloc(n));
  HIPCHK(sval.alloc(n));
  HIPCHK(ukey.alloc(n));
  HIPCHK(rcount.alloc(n));
  HIPCHK(nruns_d.alloc(1));
  HIPCHK(level_runs(n, gid.p, core.p, lkey.p, lval.p, skey.p, sval.p, ukey.p,
                    rcount.p, nruns_d.p, A, st));
  lkey.reset();
  lval.reset();
  uint32_t nruns = 0;
  CHK(readback(e, &nruns, nruns_d.p, 1));
  std::vector<uint64_t> hukey(nruns);
  std::vector<uint32_t> hcount(nruns), hkilled(nruns);
  CHK(readback(e, hukey.data(), ukey.p, nruns));
  CHK(readback(e, hcount.data(), rcount.p, nruns));
  HIPCHK(rkilled.alloc(nruns));
  HIPCHK(run_killed(T, k, L.cliques.p, kill.p, gid.p, ukey.p, nruns, rkilled.p, st));
  CHK(readback(e, hkilled.data(), rkilled.p, nruns));
  std::vector<std::vector<Level>> levels(B);
  std::vector<std::vector<uint32_t>> run_pos(B);  // first sorted position
  {
    uint32_t pos = 0;
    for (uint32_t r = 0; r < nruns; ++r) {
      const uint32_t b = (uint32_t)(hukey[r] >> 32);
      if (b >= B) return fail(CDS_ERR_INTERNAL, "level run of graph %u", b);
      levels[b].push_back({(uint32_t)hukey[r], hcount[r], hkilled[r]});
      run_pos[b].push_back(pos);
      pos += hcount[r];
    }
  }
  std::vector<PeelSummary> PS(B);
  for (uint32_t b = 0; b < B; ++b) {
    cds_status s = summarize_levels(levels[b], gs[b]->n, Tb[b], &PS[b]);
    if (s != CDS_OK) return s;
  }

  // ---- exact replays: every candidate phase of every graph at once; a
  // phase with few vertices and many cliques each (the plain path's rule)
  // runs on the whole GPU instead of one workgroup ----
  {
    std::vector<ReplayPhase> phases, heavy;
    std::vector<uint32_t> phase_graph, heavy_graph;
    uint64_t pinc_words = 0;
    for (uint32_t b = 0; b < B; ++b)
      for (size_t ii : PS[b].replay) {
        const Level &lv = levels[b][ii];
        ReplayPhase ph;
        ph.level = lv.level;
        ph.base = run_pos[b][ii];
        ph.M = (uint32_t)lv.removed;
        ph.pad = 0;
        ph.T_start = Tb[b] - PS[b].kl_before[ii];
        ph.N_start = gs[b]->n - PS[b].rm_before[ii];
        if (union_grid_replay() && ph.M <= kReplayGridMaxVertices &&
            lv.killed * k >= e->replay_grid_work * ph.M) {
          heavy.push_back(ph);
          heavy_graph.push_back(b);
          continue;
        }
        phases.push_back(ph);
        phase_graph.push_back(b);
        pinc_words += lv.killed * k * k;
      }
    DBuf<uint32_t> pdeg(&A), posmap(&A), rdead(&A);
    DBuf<uint8_t> palive(&A);
    if (!phases.empty() || !heavy.empty()) {
      HIPCHK(pdeg.alloc(n));
      HIPCHK(posmap.alloc(n));
      HIPCHK(palive.alloc(n));
      HIPCHK(rdead.alloc(T));
      HIPCHK(rdead.zero());
    }
    if (!phases.empty()) {
      const uint32_t np = (uint32_t)phases.size();
      DBuf<ReplayPhase> dph(&A);
      DBuf<uint32_t> pinc(&A);
      DBuf<uint64_t> pioff(&A);
      DBuf<ReplayOut> rout(&A);
      HIPCHK(dph.alloc(np));
      HIPCHK(hipMemcpyAsync(dph.p, phases.data(), np * sizeof(ReplayPhase),
                            hipMemcpyHostToDevice, st));
      HIPCHK(pioff.alloc((size_t)n + 1));
      HIPCHK(pinc.alloc(pinc_words));
      HIPCHK(rout.alloc(np));
      HIPCHK(positions_of(n, sval.p, posmap.p, st));
      HIPCHK(hipEventRecord(tm.ev[22], st));
      HIPCHK(replay_phases_multi(k, np, dph.p, n, sval.p, sdeg.p, pdeg.p, posmap.p,
                                 palive.p, core.p, kill.p, ioff.p, inc.p, pioff.p,
                                 pinc.p, rdead.p, rout.p, A, st));
      HIPCHK(hipEventRecord(tm.ev[23], st));
      t_replay += tm.ms(22, 23);
      std::vector<ReplayOut> ro(np);
      CHK(readback(e, ro.data(), rout.p, np));
      for (uint32_t i = 0; i < np; ++i) {
        if (ro[i].steps != phases[i].M)
          return fail(CDS_ERR_INTERNAL, "batch item %u: replay of level %u ran %u of %u steps",
                      phase_graph[i], phases[i].level, ro[i].steps, phases[i].M);
        apply_replay(&PS[phase_graph[i]], ro[i]);
        rs[phase_graph[i]]->replay_phases++;
      }
    }
    for (size_t i = 0; i < heavy.size(); ++i) {
      // one phase at a time on the whole GPU; its vertices are a run of the
      // (graph, core)-sorted list, its cliques its own graph's
      const ReplayPhase &ph = heavy[i];
      const uint32_t b = heavy_graph[i];
      uint64_t killed = 0;
      for (const Level &lv : levels[b])
        if (lv.level == ph.level) killed = lv.killed;
      unsigned grid = coop_width(
          e, replay_grid_blocks(e->device, killed * k / ph.M, e->cur.replay_blocks));
      DBuf<uint32_t> partial(&A), pcount(&A);
      DBuf<GridBarrier> gbar(&A);
      DBuf<ReplayOut> hout(&A);
      HIPCHK(partial.alloc((size_t)grid * ph.M));
      HIPCHK(pcount.alloc(1));
      HIPCHK(gbar.alloc(1));
      HIPCHK(hout.alloc(1));
      const uint64_t kp = (uint64_t)PS[b].kp;
      {
        CoopGate gate(e, grid);
        HIPCHK(hipEventRecord(tm.ev[22], st));
        const unsigned asked = grid;
        HIPCHK(replay_phase_list(k, ph.level, sval.p + ph.base, ph.M, core.p, kill.p, ioff.p,
                                 inc.p, ph.T_start, ph.N_start, pdeg.p, posmap.p, palive.p,
                                 rdead.p, pcount.p, hout.p, partial.p, gbar.p, &grid, sdeg.p,
                                 kp, replay_nstar(k, kp), st));
        HIPCHK(hipEventRecord(tm.ev[23], st));
        if (grid < asked) coop_shrinks++;
        if (replay_grid == 0 || grid < replay_grid) replay_grid = grid;
      }
      t_replay += tm.ms(22, 23);
      ReplayOut ro;
      CHK(readback(e, &ro, hout.p, 1));
      if (grid > 1) {
        unsigned ab = 0;
        CHK(readback(e, &ab, &gbar.p->abort, 1));
        if (ab) return barrier_failed("whole-GPU replay", grid);
      }
      if (ro.steps > ph.M || (!ro.stopped && ro.steps != ph.M))
        return fail(CDS_ERR_INTERNAL, "batch item %u: replay of level %u ran %u of %u steps",
                    b, ph.level, ro.steps, ph.M);
      apply_replay(&PS[b], ro);
      rs[b]->replay_phases++;
      rs[b]->replay_grid_phases++;
    }
  }
  const int ev_peel = tm.mark();
  for (uint32_t b = 0; b < B; ++b) {
    rs[b]->t_incidence_ms = t_inc;
    rs[b]->t_peel_kernel_ms = t_peel;
    rs[b]->t_replay_ms = t_replay;
  }

  // ---- densest cores: per-graph k' turned into 0/1 flags ----
  std::vector<uint32_t> kpb(B);
  std::vector<uint32_t> core_n(B);
  std::vector<uint64_t> core_T(B);
  for (uint32_t b = 0; b < B; ++b) {
    cds_result *R = rs[b];
    const PeelSummary &S = PS[b];
    R->peel_max_density = S.rho;
    R->k_prime = (uint64_t)S.kp;
    R->kmax = S.kmax;
    R->peel_levels = peel_rounds;
    R->peel_grid = peel_grid;
    R->replay_grid = replay_grid;
    R->coop_shrinks = coop_shrinks;
    R->peel_distinct_levels = (uint32_t)levels[b].size();
    size_t cl = levels[b].size();
    for (size_t i = 0; i < levels[b].size(); ++i)
      if ((int64_t)levels[b][i].level >= S.kp) {
        cl = i;
        break;
      }
    if (cl == levels[b].size())
      return fail(CDS_ERR_INTERNAL, "batch item %u: empty densest core", b);
    core_n[b] = (uint32_t)(gs[b]->n - S.rm_before[cl]);
    core_T[b] = Tb[b] - S.kl_before[cl];
    R->core_size = core_n[b];
    R->core_cliques = core_T[b];
    R->deleted_cliques = Tb[b] - core_T[b];
    R->core_density = (double)(int64_t)core_T[b] / static_cast<double>(core_n[b]);
    if (S.kp > 0xFFFFFFFFll) return fail(CDS_ERR_RANGE, "k' exceeds 32 bits");
    kpb[b] = (uint32_t)S.kp;
  }
  DBuf<uint32_t> dkp(&A), core01(&A), kill01(&A);
  HIPCHK(dkp.alloc(B));
  HIPCHK(hipMemcpyAsync(dkp.p, kpb.data(), B * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  HIPCHK(core01.alloc(n));
  HIPCHK(kill01.alloc(T));
  HIPCHK(densest_flags(n, T, k, L.cliques.p, gid.p, core.p, kill.p, dkp.p, core01.p,
                       kill01.p, st));
  const int ev_core = tm.mark();

  // ---- prune + per-graph edge counts ----
  DBuf<uint8_t> mark(&A);
  HIPCHK(mark.alloc(abase[B]));
  HIPCHK(mark.zero());
  HIPCHK(prune_mark(T, k, L.cliques.p, kill01.p, 1, uoff.p, uadj.p, mark.p, st));
  {
    DBuf<unsigned long long> ce(&A), ke(&A);
    HIPCHK(ce.alloc(B));
    HIPCHK(ke.alloc(B));
    HIPCHK(ce.zero());
    HIPCHK(ke.zero());
    HIPCHK(graph_edge_counts(n, gid.p, uoff.p, uadj.p, core01.p, mark.p, ce.p, ke.p, st));
    std::vector<unsigned long long> hce(B), hke(B);
    CHK(readback(e, hce.data(), ce.p, B));
    CHK(readback(e, hke.data(), ke.p, B));
    for (uint32_t b = 0; b < B; ++b) {
      rs[b]->core_edges = hce[b] / 2;
      rs[b]->pruned_edges = hke[b] / 2;
    }
  }
  const int ev_prune = tm.mark();

  // ---- components (oracle order: graph, then smallest vertex) ----
  DBuf<uint32_t> parent(&A);
  HIPCHK(parent.alloc(n));
  HIPCHK(uf_init(n, parent.p, st));
  HIPCHK(uf_union_edges(n, uoff.p, uadj.p, mark.p, core01.p, 1, parent.p, st));
  HIPCHK(uf_flatten(n, parent.p, st));
  std::vector<uint32_t> hpar(n), hflag(n);
  CHK(readback(e, hpar.data(), parent.p, n));
  CHK(readback(e, hflag.data(), core01.p, n));
  uint32_t core_total = 0;
  for (uint32_t b = 0; b < B; ++b) core_total += core_n[b];
  std::vector<uint32_t> comp_of(n, 0), fv_of(n, 0), fv_orig(core_total);
  std::vector<Component> comps;
  std::vector<uint32_t> comp_first(B + 1, 0);
  {
    std::vector<uint32_t> cidx(n, 0);
    for (uint32_t b = 0; b < B; ++b) {
      comp_first[b] = (uint32_t)comps.size();
      for (uint32_t v = vbase[b]; v < vbase[b + 1]; ++v) {
        if (!hflag[v]) continue;
        if (hpar[v] == v) {
          cidx[v] = (uint32_t)comps.size();
          Component c;
          c.first = v;
          comps.push_back(c);
        }
        const uint32_t ci = cidx[hpar[v]];
        comp_of[v] = ci;
        comps[ci].V++;
      }
    }
    comp_first[B] = (uint32_t)comps.size();
    uint32_t base = 0;
    for (auto &c : comps) {
      c.vbase = base;
      base += c.V;
    }
    if (base != core_total)
      return fail(CDS_ERR_INTERNAL, "component sizes %u != cores %u", base, core_total);
    std::vector<uint32_t> fill(comps.size(), 0);
    for (uint32_t v = 0; v < n; ++v) {
      if (!hflag[v]) continue;
      const uint32_t ci = comp_of[v];
      fv_of[v] = comps[ci].vbase + fill[ci]++;
      fv_orig[fv_of[v]] = v;
    }
  }
  const uint32_t ncomp = (uint32_t)comps.size();
  uint64_t core_T_total = 0;
  for (uint32_t b = 0; b < B; ++b) {
    rs[b]->num_components = comp_first[b + 1] - comp_first[b];
    core_T_total += core_T[b];
  }
  DBuf<uint32_t> d_comp_of(&A), d_fv_of(&A), ccount(&A), fcd(&A), ccur(&A), fmem(&A),
      finc(&A);
  DBuf<uint64_t> coff(&A), fioff(&A);
  HIPCHK(d_comp_of.alloc(n));
  HIPCHK(d_fv_of.alloc(n));
  CHK(upload(e, d_comp_of.p, comp_of.data(), n));
  CHK(upload(e, d_fv_of.p, fv_of.data(), n));
  HIPCHK(ccount.alloc(ncomp));
  HIPCHK(ccount.zero());
  HIPCHK(fcd.alloc(core_total));
  HIPCHK(comp_clique_totals(T, k, L.cliques.p, kill01.p, 1, d_comp_of.p, ncomp, ccount.p,
                            st));
  std::vector<uint32_t> hcc(ncomp), hfcd(core_total, 0);
  std::vector<uint64_t> hfioff((size_t)core_total + 1, 0);
  CHK(readback(e, hcc.data(), ccount.p, ncomp));
  std::vector<uint64_t> hcoff((size_t)ncomp + 1, 0);
  for (uint32_t c = 0; c < ncomp; ++c) {
    comps[c].T = hcc[c];
    comps[c].cbase = hcoff[c];
    hcoff[c + 1] = hcoff[c] + hcc[c];
  }
  if (hcoff[ncomp] != core_T_total)
    return fail(CDS_ERR_INTERNAL, "bucketed %llu of %llu core cliques",
                (unsigned long long)hcoff[ncomp], (unsigned long long)core_T_total);
  if (core_T_total * k >= 0xFFFFFFFFull)
    return fail(CDS_ERR_RANGE, "core clique slots exceed 32 bits");
  HIPCHK(coff.alloc((size_t)ncomp + 1));
  CHK(upload(e, coff.p, hcoff.data(), ncomp + 1));
  HIPCHK(ccur.alloc(ncomp));
  HIPCHK(ccur.zero());
  HIPCHK(fmem.alloc(core_T_total * k));
  {
    HIPCHK(comp_cliques_build(T, k, L.cliques.p, kill01.p, 1, d_comp_of.p, d_fv_of.p,
                              ncomp, coff.p, ccur.p, fmem.p, nullptr, st));
  }
  HIPCHK(flow_incidence_degrees(core_T_total, k, fmem.p, core_total, fcd.p, st));
  CHK(readback(e, hfcd.data(), fcd.p, core_total));
  for (uint32_t f = 0; f < core_total; ++f) hfioff[f + 1] = hfioff[f] + hfcd[f];
  if (hfioff[core_total] != core_T_total * k)
    return fail(CDS_ERR_INTERNAL, "flow incidence holds %llu of %llu pairs",
                (unsigned long long)hfioff[core_total],
                (unsigned long long)(core_T_total * k));
  HIPCHK(fioff.alloc((size_t)core_total + 1));
  CHK(upload(e, fioff.p, hfioff.data(), core_total + 1));
  HIPCHK(finc.alloc(core_T_total * k));
  {
    DBuf<uint32_t> ficur(&A);
    HIPCHK(ficur.alloc(core_total));
    HIPCHK(flow_incidence_fill(core_T_total, k, fmem.p, fioff.p, core_total, ficur.p,
                               finc.p, false, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  inc.reset();
  const int ev_cc = tm.mark();

  // ---- (c) one ExactSearch per graph, flows batched per step ----
  std::vector<std::vector<Component>> gcomps(B);
  for (uint32_t b = 0; b < B; ++b)
    gcomps[b].assign(comps.begin() + comp_first[b], comps.begin() + comp_first[b + 1]);
  std::vector<std::unique_ptr<ExactSearch>> search(B);
  for (uint32_t b = 0; b < B; ++b) {
    search[b].reset(new ExactSearch(k, ub1, ub2, gcomps[b], hfcd.data(),
                                    rs[b]->core_density, PS[b].kmax));
    rs[b]->paper_lower_bound = search[b]->paper_lower_bound;
  }
  std::vector<uint64_t> sweeps(B, 0);
  std::vector<char> active(B, 1);
  DBuf<unsigned long long> acc(&A);
  HIPCHK(acc.alloc(2));
  const char *bg = getenv("CDS_BATCH_GRAPH");
  const bool use_graph = !(bg && atoi(bg) == 0);
  uint32_t graph_steps = 0;
  for (;;) {
    struct Pending {
      uint32_t b;
      size_t i;  // index in the graph's round
      FlowRequest rq;
      const Component *C;
    };
    std::vector<Pending> pend;
    std::vector<FlowRequest> round;
    for (uint32_t b = 0; b < B; ++b) {
      if (!active[b]) continue;
      cds_status serr = CDS_OK;
      if (search[b]->next(&round, &serr)) {
        for (size_t i = 0; i < round.size(); ++i)
          pend.push_back({b, i, round[i], &gcomps[b][round[i].comp]});
      } else {
        if (serr != CDS_OK) return serr;
        active[b] = 0;
      }
    }
    if (pend.empty()) break;
    // state slices for every pending flow
    const size_t np = pend.size();
    std::vector<uint64_t> xo(np + 1, 0), vo(np + 1, 0), to(np + 1, 0);
    uint32_t max_v = 0;
    for (size_t i = 0; i < np; ++i) {
      const Component &C = *pend[i].C;
      uint32_t maxcd = 0;
      for (uint32_t v = 0; v < C.V; ++v) maxcd = std::max(maxcd, hfcd[C.vbase + v]);
      if ((__int128)maxcd * pend[i].rq.U >= ((__int128)1 << 62) ||
          (__int128)pend[i].rq.capT * C.V >= ((__int128)1 << 62) ||
          (__int128)C.T * k * pend[i].rq.U >= ((__int128)1 << 62))
        return fail(CDS_ERR_RANGE, "batch item %u: component exceeds the int64 flow range",
                    pend[i].b);
      xo[i + 1] = xo[i] + C.T * k;
      vo[i + 1] = vo[i] + C.V;
      to[i + 1] = to[i] + C.T;
      max_v = std::max(max_v, C.V);
    }
    DBuf<long long> fx(&A), frvs(&A), fexv(&A), frem(&A), fexc(&A);
    DBuf<int> fhv(&A), fhmin(&A), fhc(&A), freach(&A);
    DBuf<uint32_t> fl0(&A), fl1(&A), fq0(&A), fq1(&A), fastop(&A);
    DBuf<uint64_t> flc0(&A), flc1(&A), farc(&A);
    DBuf<GridFlowCtl> fctl(&A);
    DBuf<FlowBatchItem> fitems(&A);
    DBuf<unsigned long long> fcnt(&A);
    HIPCHK(fx.alloc(xo[np]));
    HIPCHK(frvs.alloc(vo[np]));
    HIPCHK(fexv.alloc(vo[np]));
    HIPCHK(frem.alloc(vo[np]));
    HIPCHK(fexc.alloc(to[np]));
    HIPCHK(fhv.alloc(vo[np]));
    HIPCHK(fhmin.alloc(vo[np]));
    HIPCHK(fhc.alloc(to[np]));
    HIPCHK(freach.alloc(vo[np]));
    HIPCHK(fl0.alloc(vo[np]));
    HIPCHK(fl1.alloc(vo[np]));
    HIPCHK(flc0.alloc(vo[np]));
    HIPCHK(flc1.alloc(vo[np]));
    HIPCHK(farc.alloc(vo[np]));
    HIPCHK(fastop.alloc(vo[np]));
    HIPCHK(fq0.alloc(to[np]));
    HIPCHK(fq1.alloc(to[np]));
    HIPCHK(fctl.alloc(np));
    HIPCHK(fcnt.alloc(np));
    std::vector<FlowBatchItem> all(np);
    const unsigned cap = flow_batch_capacity(max_v, e->device);
    if (cap == 0) return fail(CDS_ERR_HIP, "the batched max flow cannot be launched");
    // cooperative launches stay within half the CUs, like every other
    // cooperative kernel here
    const unsigned coop_cap = std::max(1u, cap / 2);
    for (size_t i = 0; i < np; ++i) {
      const Component &C = *pend[i].C;
      FlowBatchItem &it = all[i];
      it.P.k = k;
      it.P.V = C.V;
      it.P.T = (uint32_t)C.T;
      it.P.vbase = C.vbase;
      it.P.sbase = C.cbase * k;
      it.P.mem = fmem.p + C.cbase * k;
      it.P.ioff = fioff.p + C.vbase;
      it.P.inc = finc.p;
      it.P.cd = fcd.p + C.vbase;
      it.S.x = fx.p + xo[i];
      it.S.rvs = frvs.p + vo[i];
      it.S.exv = fexv.p + vo[i];
      it.S.rem = frem.p + vo[i];
      it.S.exc = fexc.p + to[i];
      it.S.hv = fhv.p + vo[i];
      it.S.hmin = fhmin.p + vo[i];
      it.S.hc = fhc.p + to[i];
      it.S.list0 = fl0.p + vo[i];
      it.S.list1 = fl1.p + vo[i];
      it.S.lchunk0 = flc0.p + vo[i];
      it.S.lchunk1 = flc1.p + vo[i];
      it.S.cq0 = fq0.p + to[i];
      it.S.cq1 = fq1.p + to[i];
      it.S.ctl = fctl.p + i;
      it.S.reach = freach.p + vo[i];
      it.S.arc = farc.p + vo[i];
      it.S.astop = fastop.p + vo[i];
      it.capT = pend[i].rq.capT;
      it.U = pend[i].rq.U;
      // one workgroup per 64k arcs, at most a quarter of the CUs (the
      // tuned flow width is the whole-network flow's of one graph of the
      // union's size; only CDS_FLOW_BLOCKS overrides the per-problem size)
      const uint64_t want = (C.T * k + 65535) / 65536;
      it.nblocks = (uint32_t)std::max<uint64_t>(
          1, std::min<uint64_t>(e->flow_blocks ? e->flow_blocks : want,
                                std::max(1u, coop_cap / 2)));
      it.first_block = 0;
    }
    HIPCHK(fctl.zero());
    // launches: every single-workgroup problem together in one plain
    // launch, then the multi-workgroup ones packed into cooperative
    // launches of at most coop_cap workgroups
    std::vector<std::vector<size_t>> groups;
    {
      std::vector<size_t> singles, multis;
      for (size_t i = 0; i < np; ++i) (all[i].nblocks == 1 ? singles : multis).push_back(i);
      if (!singles.empty()) groups.push_back(singles);
      std::vector<size_t> cur;
      uint32_t used = 0;
      for (size_t i : multis) {
        if (used + all[i].nblocks > coop_cap) {
          groups.push_back(cur);
          cur.clear();
          used = 0;
        }
        cur.push_back(i);
        used += all[i].nblocks;
      }
      if (!cur.empty()) groups.push_back(cur);
    }
    HIPCHK(fitems.alloc(np));
    std::vector<unsigned long long> hcnt(np, 0);
    size_t item_at = 0;
    for (const auto &grp : groups) {
      std::vector<FlowBatchItem> li;
      uint32_t blocks = 0, mv = 0;
      bool coop = false;
      for (size_t i : grp) {
        FlowBatchItem it = all[i];
        it.first_block = blocks;
        blocks += it.nblocks;
        coop |= it.nblocks > 1;
        mv = std::max(mv, it.P.V);
        li.push_back(it);
      }
      coop = coop && blocks > 1;
      if (!coop && use_graph) {
        // single-workgroup problems: the step is one graph launch
        HIPCHK(flow_batch_graph(&e->flow_graph, li.data(), (uint32_t)li.size(), blocks,
                                kGridFlowRounds, mv, fcnt.p + item_at, st));
        ++graph_steps;
      } else {
        FlowBatchItem *dli = fitems.p + item_at;
        HIPCHK(hipMemcpyAsync(dli, li.data(), li.size() * sizeof(FlowBatchItem),
                              hipMemcpyHostToDevice, st));
        CoopGate gate(e, coop ? blocks : 1);
        HIPCHK(flow_batch(dli, (uint32_t)li.size(), blocks, coop, kGridFlowRounds, mv,
                          fcnt.p + item_at, st));
      }
      std::vector<unsigned long long> sc(li.size());
      CHK(readback(e, sc.data(), fcnt.p + item_at, li.size()));
      for (size_t j = 0; j < grp.size(); ++j) hcnt[grp[j]] = sc[j];
      item_at += li.size();
    }
    std::vector<GridFlowCtl> hctl(np);
    std::vector<int> hreach(vo[np]);
    CHK(readback(e, hctl.data(), fctl.p, np));
    CHK(readback(e, hreach.data(), freach.p, vo[np]));
    for (size_t i = 0; i < np; ++i) {
      const uint32_t b = pend[i].b;
      if (hctl[i].bar.abort) return barrier_failed("batched max flow", all[i].nblocks);
      if (hctl[i].limit)
        return fail(CDS_ERR_NOT_CONVERGED, "batch item %u: max flow hit %u rounds", b,
                    hctl[i].rounds);
      sweeps[b] += hctl[i].rounds;
      rs[b]->flow_launches++;
      if (search[b]->feed_flow(pend[i].i, (long long)hctl[i].sink))
        CHK(search[b]->feed_reach(pend[i].i, hreach.data() + vo[i], hcnt[i]));
    }
  }
  const int ev_exact = tm.mark();

  // ---- results ----
  std::vector<std::vector<uint32_t>> verts(B);
  std::vector<uint8_t> hm(n, 0);
  for (uint32_t b = 0; b < B; ++b) {
    ExactSearch &S = *search[b];
    cds_result *R = rs[b];
    R->flow_graph_launches = graph_steps;
    R->ub1_val = S.ub1_val;
    R->ub2_val = S.ub2_val;
    R->flow_networks_solved = S.flows_solved;
    R->density_certificates = S.certificates;
    R->flow_sweeps = sweeps[b];
    R->density = (float)S.best_density;
    R->size = S.best_size;
    if (!gcomps[b].empty()) {
      const Component &C = gcomps[b][S.best_comp];
      for (uint32_t v = 0; v < C.V; ++v)
        if (S.best_res.empty() || S.best_res[v] != -1) {
          const uint32_t u = fv_orig[C.vbase + v];
          hm[u] = 1;
          verts[b].push_back(u - vbase[b]);
        }
    }
  }
  {
    DBuf<uint8_t> mask(&A);
    DBuf<uint32_t> inset(&A);
    HIPCHK(mask.alloc(n));
    CHK(upload(e, mask.p, hm.data(), n));
    HIPCHK(inset.alloc(B));
    HIPCHK(inset.zero());
    HIPCHK(graph_cliques_in_mask(T, k, L.cliques.p, gid.p, mask.p, inset.p, st));
    std::vector<uint32_t> hin(B);
    CHK(readback(e, hin.data(), inset.p, B));
    for (uint32_t b = 0; b < B; ++b) rs[b]->set_cliques = hin[b];
  }
  std::vector<uint32_t> hcore;
  bool any_arrays = false;
  for (uint32_t b = 0; b < B; ++b) any_arrays |= ps[b]->want_vertex_arrays != 0;
  if (any_arrays) {
    hcore.resize(n);
    CHK(readback(e, hcore.data(), core.p, n));
  }
  HIPCHK(hipStreamSynchronize(st));
  const double total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0)
          .count();
  for (uint32_t b = 0; b < B; ++b) {
    ExactSearch &S = *search[b];
    cds_result *R = rs[b];
    R->set_size = (uint32_t)verts[b].size();
    R->vertices = host_alloc<uint32_t>(verts[b].size());
    memcpy(R->vertices, verts[b].data(), verts[b].size() * 4);
    fill_trace(R, S, k, gcomps[b]);
    if (ps[b]->want_vertex_arrays) {
      const uint32_t nb = gs[b]->n;
      R->clique_degree = host_alloc<uint64_t>(nb);
      memcpy(R->clique_degree, cd.data() + vbase[b], (size_t)nb * 8);
      R->core = host_alloc<uint32_t>(nb);
      memcpy(R->core, hcore.data() + vbase[b], (size_t)nb * 4);
    }
    R->times.dag = tm.ms(ev_start, ev_dag);
    R->times.listing = tm.ms(ev_dag, ev_list);
    R->times.core_decompose = tm.ms(ev_list, ev_peel);
    R->times.densest_core = tm.ms(ev_peel, ev_core);
    R->times.prune = tm.ms(ev_core, ev_prune);
    R->times.components = tm.ms(ev_prune, ev_cc);
    R->times.exact = tm.ms(ev_cc, ev_exact);
    R->times.total = total_ms;
    R->peak_device_bytes = A.peak_bytes();
  }
  return CDS_OK;
}

}  // namespace

namespace cds {
bool solves_on_core(uint32_t k, uint64_t T) { return needs_prune(k, T); }
}  // namespace cds

extern "C" cds_status cds_solve(cds_engine *e, const cds_dgraph *g,
                                const cds_params *p, cds_result *out) {
  if (!e || !g || !p || !out) return fail(CDS_ERR_INVALID, "null argument");
  if (p->k < 2 || p->k > 4)
    return fail(CDS_ERR_UNSUPPORTED, "k=%u: the device lister covers k=2..4",
                p->k);
  std::lock_guard<std::mutex> lk(e->mu);
  HIPCHK(hipSetDevice(e->device));
  cds_status s = solve_impl(e, g, p, out, false);
  if (s != CDS_OK) {
    (void)hipStreamSynchronize(e->stream);
    cds_result_free(out);
  }
  return s;
}

extern "C" cds_status cds_core_decompose(cds_engine *e, const cds_dgraph *g,
                                         const cds_params *p, cds_result *out) {
  if (!e || !g || !p || !out) return fail(CDS_ERR_INVALID, "null argument");
  if (p->k < 2 || p->k > 4)
    return fail(CDS_ERR_UNSUPPORTED, "k=%u: the device lister covers k=2..4",
                p->k);
  std::lock_guard<std::mutex> lk(e->mu);
  HIPCHK(hipSetDevice(e->device));
  cds_status s = solve_impl(e, g, p, out, true);
  if (s != CDS_OK) {
    (void)hipStreamSynchronize(e->stream);
    cds_result_free(out);
  }
  return s;
}

extern "C" cds_status cds_solve_batch(cds_engine *e,
                                      const cds_dgraph *const *graphs,
                                      const cds_params *params, size_t count,
                                      cds_result *results) {
  if (!e || (!graphs && count) || (!params && count) || (!results && count))
    return fail(CDS_ERR_INVALID, "null argument");
  for (size_t i = 0; i < count; ++i) {
    memset(&results[i], 0, sizeof(cds_result));
    if (!graphs[i]) return fail(CDS_ERR_INVALID, "batch item %zu: null graph", i);
    if (params[i].k < 2 || params[i].k > 4)
      return fail(CDS_ERR_UNSUPPORTED, "batch item %zu: k=%u: the device lister covers k=2..4",
                  i, params[i].k);
    if (params[i].early_stop != -1)
      return fail(CDS_ERR_UNSUPPORTED, "batch item %zu: only earlyStop = -1 is supported", i);
  }
  if (count == 0) return CDS_OK;
  std::lock_guard<std::mutex> lk(e->mu);
  HIPCHK(hipSetDevice(e->device));
  // Group by (k, ub1, ub2); cut each group into unions of bounded size
  // (vertices below the persistent peel's 2^24, memory bounded by edges).
  constexpr uint64_t kUnionVertices = 1u << 22, kUnionEdges = 8u << 20;
  std::vector<size_t> order(count);
  for (size_t i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    const cds_params &x = params[a], &y = params[b];
    if (x.k != y.k) return x.k < y.k;
    if ((x.ub1 != 0) != (y.ub1 != 0)) return (x.ub1 != 0) < (y.ub1 != 0);
    return (x.ub2 != 0) < (y.ub2 != 0);
  });
  auto same_group = [&](size_t a, size_t b) {
    return params[a].k == params[b].k && (params[a].ub1 != 0) == (params[b].ub1 != 0) &&
           (params[a].ub2 != 0) == (params[b].ub2 != 0);
  };
  cds_status status = CDS_OK;
  std::vector<char> done(count, 0);
  size_t i = 0;
  while (i < count && status == CDS_OK) {
    std::vector<const cds_dgraph *> gs;
    std::vector<const cds_params *> ps;
    std::vector<cds_result *> rs;
    std::vector<size_t> idx;
    uint64_t nv = 0, ne = 0;
    size_t j = i;
    while (j < count && same_group(order[i], order[j])) {
      const cds_dgraph *g = graphs[order[j]];
      if (!gs.empty() && (nv + g->n > kUnionVertices || ne + g->m > kUnionEdges)) break;
      if (g->n >= kUnionVertices) {
        if (!gs.empty()) break;  // a huge graph is solved on its own
        status = solve_impl(e, g, &params[order[j]], &results[order[j]], false);
        if (status != CDS_OK) status = fail(status, "batch item %zu: %s", order[j], cds_last_error());
        done[order[j]] = status == CDS_OK;
        ++j;
        break;
      }
      gs.push_back(g);
      ps.push_back(&params[order[j]]);
      rs.push_back(&results[order[j]]);
      idx.push_back(order[j]);
      nv += g->n;
      ne += g->m;
      ++j;
    }
    if (status == CDS_OK && !gs.empty()) {
      status = solve_union(e, gs, ps, rs);
      if (status == CDS_OK) {
        for (size_t x : idx) done[x] = 1;
      } else {
        for (cds_result *r : rs) cds_result_free(r);
      }
    }
    i = j;
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  if (status != CDS_OK) {
    for (size_t x = 0; x < count; ++x)
      if (done[x]) cds_result_free(&results[x]);
    return status;
  }
  return CDS_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: coordinate descent over the launch shapes on one graph.  Each
// field in turn takes each candidate value with the others fixed; a
// candidate is kept if its best-of-reps solve time beats the incumbent's by
// 2 %.  Every candidate solve must give the first solve's answer.
// ---------------------------------------------------------------------------
namespace {

bool same_answer(const cds_result &a, const cds_result &b) {
  if (a.total_cliques != b.total_cliques || a.kmax != b.kmax || a.k_prime != b.k_prime ||
      a.core_size != b.core_size || a.core_cliques != b.core_cliques ||
      a.num_components != b.num_components || a.size != b.size ||
      a.set_cliques != b.set_cliques || a.set_size != b.set_size ||
      memcmp(&a.density, &b.density, sizeof(float)) != 0)
    return false;
  for (uint32_t i = 0; i < a.set_size; ++i)
    if (a.vertices[i] != b.vertices[i]) return false;
  return true;
}

}  // namespace

extern "C" cds_status cds_autotune(cds_engine *e, const cds_dgraph *g, const cds_params *p,
                                   const char *path, cds_tuning *best, double *best_ms,
                                   double *base_ms) {
  if (!e || !g || !p) return fail(CDS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(e->mu);
  HIPCHK(hipSetDevice(e->device));
  cds_params q = *p;
  q.want_vertex_arrays = 0;
  const uint32_t k = q.k;
  struct Forced {
    cds_engine *e;
    ~Forced() { e->forced = false; }
  } forced{e};
  e->forced = true;
  cds_tuning cur = tuning_for(e, k, g->m);
  cds_result ref;
  memset(&ref, 0, sizeof ref);
  // first solve: warms the arena and fixes the answer every candidate must give
  e->cur = cur;
  CHK(solve_impl(e, g, &q, &ref, false));
  struct RefFree {
    cds_result *r;
    ~RefFree() { cds_result_free(r); }
  } ref_free{&ref};
  const double first_ms = std::max(ref.times.total, 1e-3);
  // best of `reps` solves, about 0.3 s of work per candidate (at least 2)
  const int reps = (int)std::min(8.0, std::max(2.0, std::ceil(300.0 / first_ms)));
  auto timed = [&](const cds_tuning &t, double *ms) -> cds_status {
    e->cur = t;
    double bestt = 1e300;
    for (int r = 0; r < reps; ++r) {
      cds_result R;
      memset(&R, 0, sizeof R);
      const cds_status st = solve_impl(e, g, &q, &R, false);
      if (st != CDS_OK) {
        cds_result_free(&R);
        return st;
      }
      const bool same = same_answer(R, ref);
      bestt = std::min(bestt, R.times.total);
      cds_result_free(&R);
      if (!same)
        return fail(CDS_ERR_INTERNAL,
                    "autotune: shape (%u,%u) peel %u/%d replay %u flow %u changed the answer",
                    t.list_threads, t.list_chunk, t.peel_blocks, t.peel_coarse_shift,
                    t.replay_blocks, t.flow_blocks);
    }
    *ms = bestt;
    return CDS_OK;
  };
  // CDS_TUNE_VERBOSE=1: one stderr line per candidate (progress of long sweeps)
  const bool verbose = [] {
    const char *v = getenv("CDS_TUNE_VERBOSE");
    return v && atoi(v) != 0;
  }();
  auto report = [&](const cds_tuning &t, double ms) {
    if (verbose)
      fprintf(stderr, "[cds_autotune] k=%u m=%llu (%u,%u) peel %u/%d replay %u flow %u: %.3f ms\n",
              k, (unsigned long long)g->m, t.list_threads, t.list_chunk, t.peel_blocks,
              t.peel_coarse_shift, t.replay_blocks, t.flow_blocks, ms);
  };
  double cur_ms = 0;
  CHK(timed(cur, &cur_ms));
  report(cur, cur_ms);
  const double start_ms = cur_ms;
  const unsigned cap = std::max(1u, e->cus / 2);
  const unsigned rule_peel = peel_grid_blocks(g->n, ref.total_cliques, k, e->device, 0);
  auto try_values = [&](auto set, const std::vector<cds_tuning> &cands) -> cds_status {
    (void)set;
    for (const cds_tuning &t : cands) {
      if (memcmp(&t, &cur, sizeof t) == 0 || !tuning_valid(t)) continue;
      double ms = 0;
      CHK(timed(t, &ms));
      report(t, ms);
      if (ms < cur_ms * 0.98) {
        cur = t;
        cur_ms = ms;
      }
    }
    return CDS_OK;
  };
  std::vector<cds_tuning> c;
  // listing tile shape (k = 3, 4)
  if (k >= 3) {
    c.clear();
    const uint32_t shapes[5][2] = {{0, 0}, {256, 32}, {512, 64}, {512, 32}, {128, 32}};
    for (auto &sh : shapes) {
      cds_tuning t = cur;
      t.list_threads = sh[0];
      t.list_chunk = sh[1];
      c.push_back(t);
    }
    CHK(try_values(0, c));
  }
  // peel: coarse steps, then width (narrow widths only where the rule is
  // narrow: a large graph on a few workgroups takes seconds per solve)
  c.clear();
  for (int sh : {-1, 0, 1, 2, 3}) {
    cds_tuning t = cur;
    t.peel_coarse_shift = sh;
    c.push_back(t);
  }
  CHK(try_values(0, c));
  c.clear();
  for (unsigned b : {0u, 1u, 4u, 16u, cap / 4, cap / 2, cap, 2 * cap}) {
    if (b > 2 * cap || (b != 0 && b < rule_peel / 4)) continue;
    cds_tuning t = cur;
    t.peel_blocks = b;
    c.push_back(t);
  }
  CHK(try_values(0, c));
  c.clear();
  for (unsigned b : {0u, 16u, cap / 4, cap / 2, cap}) {
    if (b > cap) continue;
    cds_tuning t = cur;
    t.replay_blocks = b;
    c.push_back(t);
  }
  CHK(try_values(0, c));
  c.clear();
  for (unsigned b : {0u, 1u, 8u, cap / 4, cap / 2}) {
    if (b > cap) continue;
    cds_tuning t = cur;
    t.flow_blocks = b;
    c.push_back(t);
  }
  CHK(try_values(0, c));
  // keep the winner for this (k, class) in the engine and the cache file
  TuneEntry ent;
  ent.arch = e->arch;
  ent.cus = e->cus;
  ent.k = k;
  ent.cls = edge_class(g->m);
  ent.t = cur;
  ent.best_ms = cur_ms;
  bool replaced = false;
  for (TuneEntry &x : e->tuned)
    if (x.k == ent.k && x.cls == ent.cls) {
      x = ent;
      replaced = true;
    }
  if (!replaced) e->tuned.push_back(ent);
  if (path) CHK(merge_tune_file(path, ent));
  if (best) *best = cur;
  if (best_ms) *best_ms = cur_ms;
  if (base_ms) *base_ms = start_ms;
  return CDS_OK;
}
