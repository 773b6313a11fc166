// Serving: request loops with HBM-budget admission and continuous
// batching.
//
// The reference has no server (one graph per process, cudaCode/main.cu:
// 865-1104).  Here callers submit DSD queries from any thread.  Two lanes,
// each a loop thread with its own engine (own HIP stream) on the device:
//   - the batch lane takes graphs below `large_vertices`: whenever its
//     engine is free it admits every waiting request that fits (FIFO, by an
//     estimate of the device footprint against the HBM budget and a cap on
//     the batch size), uploads them, solves them together with
//     cds_solve_batch (shared launches over their disjoint union) and
//     completes their tickets; requests that arrive meanwhile join the next
//     batch;
//   - the large lane solves bigger graphs one at a time, so a long solve
//     does not hold up the small requests behind it.
// Both lanes draw on one HBM budget.  A request's answer does not depend on
// its lane or batch (cds_solve_batch equals cds_solve on each graph).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"

using cds::copy_tuning;
using cds::share_device;
using cds::fail;

namespace {

struct Request {
  cds_ticket id = 0;
  std::vector<uint32_t> off, nbr;
  uint32_t n = 0;
  cds_params params;
  uint64_t bytes = 0;  // admission estimate
  std::chrono::steady_clock::time_point submitted;
  // completion
  bool done = false, cancelled = false, taken = false;
  cds_status status = CDS_OK;
  std::string error;
  cds_result result;
  double queue_ms = 0, service_ms = 0;
};

// Device bytes a request may hold while it is solved: CSR and DAG, the
// listing work items, and clique storage (cliques, incidence, kill levels)
// for T cliques -- at submission an assumed T = k * m.  An estimate, used
// only for admission.
uint64_t estimate_bytes(uint32_t n, uint64_t m, uint32_t k, uint64_t T = 0) {
  if (T == 0) T = (uint64_t)k * m;
  return 16ull * n + 24ull * m + T * (4ull * k + 4ull * k * k + 8ull) + (1ull << 20);
}

// Once the clique count is known (large lane): the same plus the flow
// network of a densest core of up to T cliques (int64 residual, member and
// incidence slot per member; excess, height, two list entries per clique).
uint64_t estimate_counted_bytes(uint32_t n, uint64_t m, uint32_t k, uint64_t T) {
  return estimate_bytes(n, m, k, T) + T * (16ull * k + 20ull);
}

}  // namespace

struct Lane {
  cds_engine *engine = nullptr;
  bool owned = false;  // created by the server
  bool batched = true;
  std::deque<std::shared_ptr<Request>> queue;
  std::condition_variable cv_work;
  std::thread loop;
};

struct cds_server {
  cds_server_opts opts;
  std::mutex mu;
  std::condition_variable cv_done, cv_budget;
  Lane lanes[2];  // 0: batched small graphs, 1: large graphs one at a time
  std::unordered_map<cds_ticket, std::shared_ptr<Request>> all;
  cds_ticket next_id = 1;
  bool stop = false;
  uint64_t in_flight_bytes = 0;  // admitted and not yet completed
  uint64_t reserve_bytes = 0;    // large lane waiting to grow its admission
  cds_server_stats stats;
};

namespace {

void complete(cds_server *s, const std::shared_ptr<Request> &r, cds_status st,
              const char *err) {
  std::lock_guard<std::mutex> lk(s->mu);
  r->status = st;
  if (err) r->error = err;
  r->done = true;
  const auto now = std::chrono::steady_clock::now();
  r->service_ms =
      std::chrono::duration<double, std::milli>(now - r->submitted).count() - r->queue_ms;
  if (st == CDS_OK)
    s->stats.completed++;
  else
    s->stats.failed++;
  s->cv_done.notify_all();
}

void serve(cds_server *s, Lane *L) {
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch;
    uint64_t bytes = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      L->cv_work.wait(lk, [&] { return s->stop || !L->queue.empty(); });
      if (s->stop) break;
      if (L->batched && s->opts.max_wait_us) {
        // wait briefly for more arrivals to share the launches
        const auto until = std::chrono::steady_clock::now() +
                           std::chrono::microseconds(s->opts.max_wait_us);
        L->cv_work.wait_until(lk, until, [&] {
          return s->stop || L->queue.size() >= s->opts.max_batch;
        });
        if (s->stop) break;
      }
      const size_t cap = L->batched ? s->opts.max_batch : 1;
      for (;;) {
        while (!L->queue.empty() && L->queue.front()->cancelled) L->queue.pop_front();
        if (L->queue.empty() || batch.size() >= cap) break;
        std::shared_ptr<Request> r = L->queue.front();  // may be cancelled while we wait
        if (s->opts.hbm_budget_bytes &&
            s->in_flight_bytes + s->reserve_bytes + bytes + r->bytes >
                s->opts.hbm_budget_bytes) {
          if (!batch.empty()) break;  // the rest waits for this batch
          // nothing admitted yet: wait until the other lane frees HBM
          s->cv_budget.wait(lk, [&] {
            return s->stop || s->in_flight_bytes + s->reserve_bytes + r->bytes <=
                                  s->opts.hbm_budget_bytes;
          });
          if (s->stop) break;
          continue;
        }
        bytes += r->bytes;
        r->queue_ms = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - r->submitted)
                          .count();
        batch.push_back(r);
        L->queue.pop_front();
      }
      if (s->stop) break;
      if (batch.empty()) continue;
      s->in_flight_bytes += bytes;
      s->stats.batches++;
      s->stats.admitted += batch.size();
      if (s->in_flight_bytes > s->stats.peak_admitted_bytes)
        s->stats.peak_admitted_bytes = s->in_flight_bytes;
      if (batch.size() > s->stats.max_batch_seen) s->stats.max_batch_seen = batch.size();
    }
    cds_engine *e = L->engine;
    // upload (device-resident until the batch completes)
    std::vector<cds_dgraph *> dgs(batch.size(), nullptr);
    std::vector<std::shared_ptr<Request>> ok;
    std::vector<const cds_dgraph *> gptr;
    std::vector<cds_params> params;
    for (size_t i = 0; i < batch.size(); ++i) {
      auto &r = batch[i];
      cds_graph g;
      memset(&g, 0, sizeof g);
      g.n = r->n;
      g.offsets = r->off.data();
      g.neighbors = r->nbr.data();
      g.m = r->off[r->n] / 2;
      uint32_t maxd = 0;
      for (uint32_t v = 0; v < r->n; ++v) maxd = std::max(maxd, r->off[v + 1] - r->off[v]);
      g.max_degree = maxd;
      const cds_status st = cds_graph_upload(e, &g, &dgs[i]);
      if (st != CDS_OK) {
        complete(s, r, st, cds_last_error());
        continue;
      }
      if (!L->batched && s->opts.hbm_budget_bytes) {
        // A large graph's clique count can exceed the submission estimate
        // by orders of magnitude (dense cores): count exactly, then refuse
        // it or grow its admission before solving.
        uint64_t T = 0;
        cds_status cs = cds_count_cliques(e, dgs[i], r->params.k, &T, nullptr, nullptr);
        uint64_t need = 0;
        if (cs == CDS_OK) {
          need = estimate_counted_bytes(r->n, g.m, r->params.k, T);
          // A graph the engine solves on its clique core (too many cliques
          // to materialise) needs far less than its whole clique list; it
          // runs with the whole budget to itself instead of being refused.
          if (need > s->opts.hbm_budget_bytes && cds::solves_on_core(r->params.k, T))
            need = s->opts.hbm_budget_bytes;
          if (need > s->opts.hbm_budget_bytes)
            cs = fail(CDS_ERR_OOM, "request needs ~%llu bytes for %llu cliques, budget %llu",
                      (unsigned long long)need, (unsigned long long)T,
                      (unsigned long long)s->opts.hbm_budget_bytes);
        }
        if (cs != CDS_OK) {
          complete(s, r, cs, cds_last_error());
          continue;
        }
        if (need > r->bytes) {
          std::unique_lock<std::mutex> lk(s->mu);
          const uint64_t extra = need - r->bytes;
          s->reserve_bytes += extra;
          s->cv_budget.wait(lk, [&] {
            return s->stop ||
                   s->in_flight_bytes + extra <= s->opts.hbm_budget_bytes;
          });
          s->reserve_bytes -= extra;
          if (s->stop) {
            lk.unlock();
            complete(s, r, CDS_ERR_CANCELLED, "server is stopping");
            continue;
          }
          s->in_flight_bytes += extra;
          bytes += extra;
          if (s->in_flight_bytes > s->stats.peak_admitted_bytes)
            s->stats.peak_admitted_bytes = s->in_flight_bytes;
        }
      }
      ok.push_back(r);
      gptr.push_back(dgs[i]);
      params.push_back(r->params);
    }
    if (!ok.empty()) {
      std::vector<cds_result> res(ok.size());
      const cds_status st =
          ok.size() == 1 ? cds_solve(e, gptr[0], &params[0], &res[0])
                         : cds_solve_batch(e, gptr.data(), params.data(), ok.size(),
                                           res.data());
      if (st == CDS_OK) {
        for (size_t i = 0; i < ok.size(); ++i) {
          ok[i]->result = res[i];
          complete(s, ok[i], CDS_OK, nullptr);
        }
      } else if (ok.size() == 1) {
        complete(s, ok[0], st, cds_last_error());
      } else {
        // isolate the failure: every request on its own
        for (size_t i = 0; i < ok.size(); ++i) {
          cds_result r1;
          const cds_status s1 = cds_solve(e, gptr[i], &params[i], &r1);
          if (s1 == CDS_OK) ok[i]->result = r1;
          complete(s, ok[i], s1, s1 == CDS_OK ? nullptr : cds_last_error());
        }
      }
    }
    for (cds_dgraph *d : dgs)
      if (d) cds_dgraph_free(e, d);
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &r : batch) {  // the host copy of the graph is no longer needed
      std::vector<uint32_t>().swap(r->off);
      std::vector<uint32_t>().swap(r->nbr);
    }
    s->in_flight_bytes -= bytes;
    s->cv_budget.notify_all();
  }
  // stop: fail what is still queued
  std::lock_guard<std::mutex> lk(s->mu);
  for (auto &r : L->queue) {
    r->status = CDS_ERR_CANCELLED;
    r->done = true;
  }
  L->queue.clear();
  s->cv_done.notify_all();
}

}  // namespace

extern "C" cds_status cds_server_create(cds_engine *e, const cds_server_opts *o,
                                        cds_server **out) {
  if (!e || !out) return fail(CDS_ERR_INVALID, "null argument");
  cds_opts eo;
  memset(&eo, 0, sizeof eo);
  eo.device = cds_engine_device(e);
  cds_engine *large = nullptr;
  const cds_status st = cds_engine_create(&eo, &large);
  if (st != CDS_OK) return st;
  copy_tuning(large, e);
  share_device(large, true);
  share_device(e, true);
  auto *s = new cds_server();
  memset(&s->opts, 0, sizeof s->opts);
  if (o) s->opts = *o;
  if (s->opts.max_batch == 0) s->opts.max_batch = 1024;
  if (s->opts.large_vertices == 0) s->opts.large_vertices = 65536;
  memset(&s->stats, 0, sizeof s->stats);
  s->lanes[0].engine = e;
  s->lanes[0].batched = true;
  s->lanes[1].engine = large;
  s->lanes[1].owned = true;
  s->lanes[1].batched = false;
  for (Lane &L : s->lanes) L.loop = std::thread(serve, s, &L);
  *out = s;
  return CDS_OK;
}

extern "C" void cds_server_destroy(cds_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  for (Lane &L : s->lanes) L.cv_work.notify_all();
  s->cv_budget.notify_all();
  for (Lane &L : s->lanes)
    if (L.loop.joinable()) L.loop.join();
  for (Lane &L : s->lanes) {
    if (L.owned)
      cds_engine_destroy(L.engine);
    else
      share_device(L.engine, false);  // the caller's engine: full width again
  }
  for (auto &kv : s->all)
    if (kv.second->done && kv.second->status == CDS_OK && !kv.second->taken)
      cds_result_free(&kv.second->result);
  delete s;
}

extern "C" cds_status cds_submit(cds_server *s, const cds_graph *g, const cds_params *p,
                                 cds_ticket *out) {
  if (!s || !g || !p || !out) return fail(CDS_ERR_INVALID, "null argument");
  if (!g->offsets || (!g->neighbors && g->offsets[g->n]))
    return fail(CDS_ERR_INVALID, "graph without CSR arrays");
  if (p->k < 2 || p->k > 4)
    return fail(CDS_ERR_UNSUPPORTED, "k=%u: the device lister covers k=2..4", p->k);
  auto r = std::make_shared<Request>();
  r->n = g->n;
  r->off.assign(g->offsets, g->offsets + g->n + 1);
  r->nbr.assign(g->neighbors, g->neighbors + g->offsets[g->n]);
  r->params = *p;
  memset(&r->result, 0, sizeof r->result);
  r->bytes = estimate_bytes(g->n, g->offsets[g->n] / 2, p->k);
  r->submitted = std::chrono::steady_clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(CDS_ERR_CANCELLED, "server is stopping");
    if (s->opts.hbm_budget_bytes && r->bytes > s->opts.hbm_budget_bytes)
      return fail(CDS_ERR_OOM, "request needs ~%llu bytes, budget %llu",
                  (unsigned long long)r->bytes,
                  (unsigned long long)s->opts.hbm_budget_bytes);
    r->id = s->next_id++;
    s->all[r->id] = r;
    Lane &L = s->lanes[g->n >= s->opts.large_vertices ? 1 : 0];
    L.queue.push_back(r);
    s->stats.submitted++;
    *out = r->id;
    L.cv_work.notify_one();
  }
  return CDS_OK;
}

extern "C" cds_status cds_poll(cds_server *s, cds_ticket t, int *done) {
  if (!s || !done) return fail(CDS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(t);
  if (it == s->all.end()) return fail(CDS_ERR_INVALID, "unknown ticket %llu",
                                      (unsigned long long)t);
  *done = it->second->done ? 1 : 0;
  return CDS_OK;
}

extern "C" cds_status cds_wait(cds_server *s, cds_ticket t, int64_t timeout_us,
                               cds_result *out, cds_request_times *times) {
  if (!s || !out) return fail(CDS_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(t);
  if (it == s->all.end())
    return fail(CDS_ERR_INVALID, "unknown ticket %llu", (unsigned long long)t);
  std::shared_ptr<Request> r = it->second;
  if (timeout_us < 0) {
    s->cv_done.wait(lk, [&] { return r->done; });
  } else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us),
                                  [&] { return r->done; })) {
    return fail(CDS_ERR_BUSY, "ticket %llu not done yet", (unsigned long long)t);
  }
  s->all.erase(it);
  if (times) {
    times->queue_ms = r->queue_ms;
    times->service_ms = r->service_ms;
  }
  if (r->status != CDS_OK)
    return fail(r->status, "ticket %llu: %s", (unsigned long long)t, r->error.c_str());
  *out = r->result;  // ownership passes to the caller (cds_result_free)
  r->taken = true;
  return CDS_OK;
}

extern "C" cds_status cds_cancel(cds_server *s, cds_ticket t) {
  if (!s) return fail(CDS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(t);
  if (it == s->all.end())
    return fail(CDS_ERR_INVALID, "unknown ticket %llu", (unsigned long long)t);
  auto &r = it->second;
  if (r->done) return fail(CDS_ERR_BUSY, "ticket %llu already done", (unsigned long long)t);
  for (Lane &L : s->lanes)
    for (auto q = L.queue.begin(); q != L.queue.end(); ++q)
      if ((*q)->id == t) {
        L.queue.erase(q);
        r->cancelled = true;
        r->status = CDS_ERR_CANCELLED;
        r->done = true;
        s->cv_done.notify_all();
        return CDS_OK;
      }
  return fail(CDS_ERR_BUSY, "ticket %llu is already on the GPU", (unsigned long long)t);
}

extern "C" cds_status cds_server_get_stats(cds_server *s, cds_server_stats *out) {
  if (!s || !out) return fail(CDS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->queued = s->lanes[0].queue.size() + s->lanes[1].queue.size();
  return CDS_OK;
}
