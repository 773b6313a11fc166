// Machine-readable result record (host C++).
//
// The reference reports only stdout lines (cudaCode/main.cu:1069-1104 and
// the oracle block of cpuCode/c++/src/helpers.cpp:1580-1588); cds_main
// prints those unchanged.  This file renders the same result as one JSON
// object so scripts need not scrape text: the answer (float density as the
// oracle prints it plus the exact fraction), the densest-core and component
// statistics, stage times, kernel times and the density-search trace.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "internal.h"

namespace {

// Minimal JSON writer: every value is preceded by sep(), which emits the
// comma between siblings and nothing right after a key.
class Json {
 public:
  void key(const char *k) {
    sep();
    out_ += '"';
    out_ += k;
    out_ += "\":";
    after_key_ = true;
  }
  void u64(const char *k, unsigned long long v) {
    key(k);
    item_u64(v);
  }
  void f64(const char *k, double v) {
    key(k);
    item_f64(v);
  }
  void str(const char *k, const char *v) {
    key(k);
    sep();
    out_ += '"';
    out_ += v;
    out_ += '"';
  }
  void raw(const char *k, const char *v) {
    key(k);
    sep();
    out_ += v;
  }
  void open(const char *k, char bracket) {
    if (k) key(k);
    sep();
    out_ += bracket;
    first_ = true;
  }
  void close(char bracket) {
    out_ += bracket;
    first_ = false;
  }
  void item_u64(unsigned long long v) {
    sep();
    fmt("%llu", v);
  }
  // JSON has no NaN or infinity: those become null.  %.17g round-trips a
  // double exactly.
  void item_f64(double v) {
    sep();
    if (!std::isfinite(v))
      out_ += "null";
    else
      fmt("%.17g", v);
  }
  const std::string &text() const { return out_; }

 private:
  void sep() {
    if (after_key_) {
      after_key_ = false;
      return;
    }
    if (!first_) out_ += ',';
    first_ = false;
  }
  void fmt(const char *f, ...) {
    char buf[64];
    va_list ap;
    va_start(ap, f);
    const int len = vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    if (len > 0) out_.append(buf, (size_t)len < sizeof buf ? (size_t)len : sizeof buf - 1);
  }
  std::string out_;
  bool first_ = true;
  bool after_key_ = false;
};

std::string render(const cds_result *r) {
  Json j;
  j.open(nullptr, '{');
  j.str("engine", cds_version());
  j.u64("k", r->k);
  j.u64("n", r->n);
  j.u64("m", r->m);
  j.u64("max_degree", r->max_degree);
  j.u64("total_cliques", r->total_cliques);
  j.open("answer", '{');
  // the float the oracle prints (%.9g round-trips a float), and the exact
  // fraction of the set
  char dens[32];
  snprintf(dens, sizeof dens, "%.9g", (double)r->density);
  j.raw("density", std::isfinite(r->density) ? dens : "null");
  j.u64("size", r->size);
  j.u64("set_cliques", r->set_cliques);
  j.u64("set_size", r->set_size);
  char frac[48];
  snprintf(frac, sizeof frac, "%llu/%u", (unsigned long long)r->set_cliques, r->set_size);
  j.str("density_exact", frac);
  j.open("vertices", '[');
  for (uint32_t i = 0; i < r->set_size && r->vertices; ++i) j.item_u64(r->vertices[i]);
  j.close(']');
  j.close('}');
  j.open("core", '{');
  j.u64("kmax", r->kmax);
  j.u64("k_prime", r->k_prime);
  j.f64("peel_max_density", r->peel_max_density);
  j.f64("core_density", r->core_density);
  j.u64("core_size", r->core_size);
  j.u64("core_edges", r->core_edges);
  j.u64("core_cliques", r->core_cliques);
  j.u64("deleted_cliques", r->deleted_cliques);
  j.u64("pruned_edges", r->pruned_edges);
  j.u64("peel_levels", r->peel_levels);
  j.u64("peel_distinct_levels", r->peel_distinct_levels);
  j.u64("replay_phases", r->replay_phases);
  // solved on the prune_tau-clique-core (0: on the whole graph)
  j.u64("prune_tau", r->prune_tau);
  j.u64("prune_vertices", r->prune_vertices);
  j.u64("prune_passes", r->prune_passes);
  j.u64("dense_peel", r->dense_peel);
  j.close('}');
  j.open("search", '{');
  j.u64("components", r->num_components);
  j.f64("paper_lower_bound", r->paper_lower_bound);
  j.f64("ub1", r->ub1_val);
  j.f64("ub2", r->ub2_val);
  j.u64("flow_networks_solved", r->flow_networks_solved);
  j.u64("flow_launches", r->flow_launches);
  j.u64("density_certificates", r->density_certificates);
  j.u64("flow_sweeps", r->flow_sweeps);
  j.u64("flow_clique_visits", r->flow_clique_visits);
  j.u64("flow_arc_visits", r->flow_arc_visits);
  j.u64("flow_relabels", r->flow_relabels);
  j.u64("flow_implicit", r->flow_implicit);
  j.u64("flow_graph_launches", r->flow_graph_launches);
  j.u64("replay_bits_phases", r->replay_bits_phases);
  j.u64("refine_bits_steps", r->refine_bits_steps);
  j.u64("list_staged", r->list_staged);
  j.u64("prune_unlinks", r->prune_unlinks);
  // steps whose exact decision the reference's float Edmonds-Karp may take
  // the other way (cds.h: trace_fragile)
  j.u64("fragile_decisions", r->fragile_decisions);
  j.open("trace", '[');
  for (uint32_t i = 0; i < r->trace_len; ++i) {
    j.open(nullptr, '{');
    j.u64("component", r->trace_component ? r->trace_component[i] : 0);
    j.f64("lb", r->trace_lb ? r->trace_lb[i] : NAN);
    j.f64("ub", r->trace_ub ? r->trace_ub[i] : NAN);
    j.f64("alpha", r->trace_alpha ? r->trace_alpha[i] : NAN);
    j.f64("margin", r->trace_margin ? r->trace_margin[i] : NAN);
    j.f64("drift_bound", r->trace_drift ? r->trace_drift[i] : NAN);
    j.u64("fragile", r->trace_fragile ? r->trace_fragile[i] : 0);
    j.close('}');
  }
  j.close(']');
  j.close('}');
  // widths the cooperative kernels ran at; shrinks: launches the runtime
  // refused at the width asked and that ran narrower
  j.open("cooperative", '{');
  j.u64("peel_workgroups", r->peel_grid);
  j.u64("flow_workgroups", r->flow_grid);
  j.u64("replay_workgroups", r->replay_grid);
  j.u64("shrinks", r->coop_shrinks);
  j.close('}');
  j.open("times_ms", '{');
  j.f64("upload", r->times.upload);
  j.f64("dag", r->times.dag);
  j.f64("listing", r->times.listing);
  j.f64("core_decompose", r->times.core_decompose);
  j.f64("densest_core", r->times.densest_core);
  j.f64("prune", r->times.prune);
  j.f64("components", r->times.components);
  j.f64("exact", r->times.exact);
  j.f64("total", r->times.total);
  j.close('}');
  j.open("kernels_ms", '{');
  j.f64("list_count", r->t_list_count_ms);
  j.f64("list_write", r->t_list_write_ms);
  j.f64("incidence", r->t_incidence_ms);
  j.f64("peel", r->t_peel_kernel_ms);
  j.f64("replay", r->t_replay_ms);
  j.close('}');
  j.open("listing", '{');
  j.u64("dag_arcs", r->dag_arcs);
  j.u64("work", r->list_work);
  j.u64("max_out_degree", r->max_out_degree);
  j.close('}');
  j.u64("peak_device_bytes", r->peak_device_bytes);
  j.close('}');
  return j.text();
}

}  // namespace

// snprintf convention: writes at most cap-1 characters and a NUL, returns the
// full length; a call with cap 0 sizes the buffer.
extern "C" size_t cds_result_json(const cds_result *r, char *buf, size_t cap) {
  if (!r) return 0;
  const std::string s = render(r);
  if (buf && cap) {
    const size_t w = s.size() < cap - 1 ? s.size() : cap - 1;
    memcpy(buf, s.data(), w);
    buf[w] = '\0';
  }
  return s.size();
}
