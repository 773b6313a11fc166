// Graph input/output and seeded synthetic generators (host C++).
//
// Input format (README.md:72-90, loader cudaCode/src/graph.cpp:7-40 and
// cpuCode/c++/src/graph.cpp:9-40): a header line `n m`, then one line per
// vertex `v n1 n2 ...`.  Both reference loaders drop self loops and keep
// duplicates; the GPU loader additionally requires the vertex lines in id
// order.  This loader accepts lines in any order, drops self loops, and
// rejects ids >= n.  It symmetrises nothing: the reference treats the file as
// the adjacency structure, so an asymmetric file is rejected with an error
// instead of silently producing a different graph.  Each adjacency list is
// sorted ascending and de-duplicated (the reference's cliques are sets, so a
// repeated neighbour cannot change any result; it would only double-count
// degrees, which the checker reports).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "internal.h"

namespace cds {

static thread_local std::string g_last_error;

cds_status fail(cds_status st, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

// Builds a sorted, de-duplicated symmetric CSR from an undirected edge list
// whose entries are packed as (u << 32 | v) with u != v, both directions
// present.
static cds_status csr_from_directed_pairs(uint32_t n, std::vector<uint64_t> &pairs,
                                          cds_graph *out) {
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  if (pairs.size() >= (1ull << 32))
    return fail(CDS_ERR_RANGE, "graph has %zu adjacency entries (>= 2^32)",
                pairs.size());
  uint32_t *off = host_alloc<uint32_t>((size_t)n + 1);
  uint32_t *nb = host_alloc<uint32_t>(pairs.size());
  if (!off || !nb) {
    free(off);
    free(nb);
    return fail(CDS_ERR_OOM, "host allocation failed");
  }
  std::fill(off, off + n + 1, 0u);
  for (uint64_t p : pairs) off[(p >> 32) + 1]++;
  uint32_t maxd = 0;
  for (uint32_t v = 0; v < n; ++v) {
    maxd = std::max(maxd, off[v + 1]);
    off[v + 1] += off[v];
  }
  for (size_t i = 0; i < pairs.size(); ++i) nb[i] = (uint32_t)pairs[i];
  out->n = n;
  out->m = pairs.size() / 2;
  out->m_header = 0;
  out->offsets = off;
  out->neighbors = nb;
  out->max_degree = maxd;
  return CDS_OK;
}

}  // namespace cds

using namespace cds;

extern "C" const char *cds_last_error(void) { return g_last_error.c_str(); }
extern "C" const char *cds_version(void) { return "cds-mi355x 0.1 (gfx950)"; }

extern "C" size_t cds_struct_size(int which) {
  switch (which) {
    case 0: return sizeof(cds_graph);
    case 1: return sizeof(cds_params);
    case 2: return sizeof(cds_result);
    case 3: return sizeof(cds_opts);
    default: return 0;
  }
}

extern "C" void cds_graph_free(cds_graph *g) {
  if (!g) return;
  free(g->offsets);
  free(g->neighbors);
  g->offsets = nullptr;
  g->neighbors = nullptr;
  g->n = 0;
  g->m = 0;
}

extern "C" cds_status cds_graph_from_csr(uint32_t n, const uint32_t *offsets,
                                         const uint32_t *neighbors,
                                         cds_graph *out) {
  if (!out || (!offsets && n)) return fail(CDS_ERR_INVALID, "null argument");
  std::vector<uint64_t> pairs;
  const uint32_t total = n ? offsets[n] : 0;
  if (n && offsets[0] != 0)
    return fail(CDS_ERR_INVALID, "offsets[0] must be 0");
  pairs.reserve(total);
  for (uint32_t v = 0; v < n; ++v) {
    if (offsets[v + 1] < offsets[v])
      return fail(CDS_ERR_INVALID, "offsets not monotone at %u", v);
    for (uint32_t j = offsets[v]; j < offsets[v + 1]; ++j) {
      uint32_t w = neighbors[j];
      if (w >= n) return fail(CDS_ERR_INVALID, "neighbor id %u >= n=%u", w, n);
      if (w == v) continue;
      pairs.push_back((uint64_t)v << 32 | w);
    }
  }
  // Symmetry check: the reverse of every pair must be present.
  std::vector<uint64_t> sorted = pairs;
  std::sort(sorted.begin(), sorted.end());
  sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
  for (uint64_t p : sorted) {
    uint64_t r = (p << 32) | (p >> 32);
    if (!std::binary_search(sorted.begin(), sorted.end(), r))
      return fail(CDS_ERR_INVALID,
                  "adjacency is not symmetric: %u lists %u but not back",
                  (uint32_t)(p >> 32), (uint32_t)p);
  }
  return csr_from_directed_pairs(n, sorted, out);
}

// Parses the adjacency-list format with a hand-written tokenizer (the
// reference uses istringstream per line, cudaCode/src/graph.cpp:24-35).
extern "C" cds_status cds_graph_load_adjlist(const char *path, cds_graph *out) {
  if (!path || !out) return fail(CDS_ERR_INVALID, "null argument");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(CDS_ERR_IO, "Graph file Open Failed: %s", path);
  std::vector<char> buf;
  {
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0) {
      fclose(f);
      return fail(CDS_ERR_IO, "cannot size %s", path);
    }
    buf.resize((size_t)sz + 1);
    size_t got = fread(buf.data(), 1, (size_t)sz, f);
    buf[got] = '\0';
    buf.resize(got + 1);
  }
  fclose(f);
  const char *p = buf.data();
  const char *end = p + buf.size() - 1;
  auto skip_space = [&](bool stop_at_newline) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' ||
                       (!stop_at_newline && *p == '\n')))
      ++p;
  };
  auto read_uint = [&](uint64_t &v) -> bool {
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t x = 0;
    while (p < end && *p >= '0' && *p <= '9') {
      x = x * 10 + (uint64_t)(*p - '0');
      if (x > 0xFFFFFFFFull) return false;
      ++p;
    }
    v = x;
    return true;
  };
  uint64_t n64 = 0, m64 = 0;
  skip_space(false);
  if (!read_uint(n64)) return fail(CDS_ERR_IO, "%s: missing header n", path);
  skip_space(true);
  if (!read_uint(m64)) m64 = 0;  // a header with n only is tolerated
  while (p < end && *p != '\n') ++p;
  const uint32_t n = (uint32_t)n64;
  std::vector<uint64_t> pairs;
  pairs.reserve(m64 ? 2 * m64 : 1024);
  uint64_t line_no = 1;
  while (p < end) {
    if (*p == '\n') {
      ++p;
      ++line_no;
      continue;
    }
    skip_space(true);
    if (p >= end || *p == '\n') continue;
    uint64_t v;
    if (!read_uint(v))
      return fail(CDS_ERR_IO, "%s:%llu: expected a vertex id", path,
                  (unsigned long long)line_no);
    if (v >= n)
      return fail(CDS_ERR_INVALID, "%s:%llu: vertex %llu >= n=%u", path,
                  (unsigned long long)line_no, (unsigned long long)v, n);
    for (;;) {
      skip_space(true);
      if (p >= end || *p == '\n') break;
      uint64_t w;
      if (!read_uint(w))
        return fail(CDS_ERR_IO, "%s:%llu: bad neighbor token", path,
                    (unsigned long long)line_no);
      if (w >= n)
        return fail(CDS_ERR_INVALID, "%s:%llu: neighbor %llu >= n=%u", path,
                    (unsigned long long)line_no, (unsigned long long)w, n);
      if (w == v) continue;  // self loops are skipped like the reference
      pairs.push_back(v << 32 | w);
    }
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  for (uint64_t q : pairs) {
    uint64_t r = (q << 32) | (q >> 32);
    if (!std::binary_search(pairs.begin(), pairs.end(), r))
      return fail(CDS_ERR_INVALID,
                  "%s: adjacency is not symmetric: %u lists %u but not back",
                  path, (uint32_t)(q >> 32), (uint32_t)q);
  }
  cds_status st = csr_from_directed_pairs(n, pairs, out);
  if (st == CDS_OK) out->m_header = m64;
  return st;
}

extern "C" cds_status cds_graph_write_adjlist(const cds_graph *g,
                                              const char *path) {
  if (!g || !path) return fail(CDS_ERR_INVALID, "null argument");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(CDS_ERR_IO, "cannot open %s for writing", path);
  std::string line;
  line.reserve(1 << 16);
  fprintf(f, "%u %llu\n", g->n, (unsigned long long)g->m);
  char tmp[16];
  for (uint32_t v = 0; v < g->n; ++v) {
    line.clear();
    int len = snprintf(tmp, sizeof tmp, "%u", v);
    line.append(tmp, len);
    for (uint32_t j = g->offsets[v]; j < g->offsets[v + 1]; ++j) {
      len = snprintf(tmp, sizeof tmp, " %u", g->neighbors[j]);
      line.append(tmp, len);
    }
    line.push_back('\n');
    fwrite(line.data(), 1, line.size(), f);
  }
  if (fclose(f) != 0) return fail(CDS_ERR_IO, "write failed: %s", path);
  return CDS_OK;
}

// ---------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------

// Adds G(s, p) on s distinct random vertex ids (partial Fisher-Yates).
static void plant_community(Rng &rng, uint32_t n, uint32_t s, double p,
                            std::vector<uint64_t> &pairs) {
  if (s > n) s = n;
  std::vector<uint32_t> ids(n);
  std::iota(ids.begin(), ids.end(), 0u);
  for (uint32_t i = 0; i < s; ++i) {
    uint32_t j = i + (uint32_t)rng.below(n - i);
    std::swap(ids[i], ids[j]);
  }
  for (uint32_t i = 0; i < s; ++i)
    for (uint32_t j = i + 1; j < s; ++j)
      if (rng.uniform() < p) {
        uint32_t u = ids[i], v = ids[j];
        pairs.push_back((uint64_t)u << 32 | v);
        pairs.push_back((uint64_t)v << 32 | u);
      }
}

// Chung-Lu graph with expected degrees w_i = c (i+1)^(-1/(exponent-1)),
// c chosen so the mean expected degree is `mean_degree`; edge {i,j} present
// with probability min(1, w_i w_j / sum(w)).  Sampled in O(n + m) with the
// skipping method of Miller & Hagberg (2011), vertices in descending weight.
extern "C" cds_status cds_gen_chung_lu(uint32_t n, double exponent,
                                       double mean_degree,
                                       uint32_t planted_size, double planted_p,
                                       uint64_t seed, cds_graph *out) {
  if (!out || n < 2 || exponent <= 1.0 || mean_degree <= 0)
    return fail(CDS_ERR_INVALID, "bad Chung-Lu parameters");
  Rng rng(seed);
  const double gamma = 1.0 / (exponent - 1.0);
  std::vector<double> w(n);
  double sum = 0;
  for (uint32_t i = 0; i < n; ++i) {
    w[i] = std::pow((double)i + 1.0, -gamma);
    sum += w[i];
  }
  const double scale = mean_degree * n / sum;
  for (auto &x : w) x *= scale;
  const double S = mean_degree * n;
  std::vector<uint64_t> pairs;
  pairs.reserve((size_t)(mean_degree * n * 1.1) + 16);
  for (uint32_t u = 0; u + 1 < n; ++u) {
    uint64_t v = u + 1;
    double p = std::min(w[u] * w[v] / S, 1.0);
    while (v < n && p > 0) {
      if (p != 1.0) {
        double r = rng.uniform();
        double skip = std::floor(std::log1p(-r) / std::log1p(-p));
        if (!(skip < (double)n)) break;
        v += (uint64_t)skip;
      }
      if (v < n) {
        double q = std::min(w[u] * w[v] / S, 1.0);
        if (rng.uniform() < q / p) {
          pairs.push_back((uint64_t)u << 32 | v);
          pairs.push_back(v << 32 | u);
        }
        p = q;
        ++v;
      }
    }
  }
  if (planted_size >= 2) plant_community(rng, n, planted_size, planted_p, pairs);
  return csr_from_directed_pairs(n, pairs, out);
}

// R-MAT (Chakrabarti, Zhan & Faloutsos 2004): edge_factor * 2^scale samples,
// each choosing one quadrant per bit with probabilities (a, b, c, 1-a-b-c),
// no noise.  Symmetrised, de-duplicated, self loops removed, isolated
// vertices kept.  Optionally plants `planted_count` G(s, planted_p)
// communities with s uniform in [planted_min, planted_max].
extern "C" cds_status cds_gen_rmat(uint32_t scale, uint32_t edge_factor,
                                   double a, double b, double c,
                                   uint32_t planted_count, uint32_t planted_min,
                                   uint32_t planted_max, double planted_p,
                                   uint64_t seed, cds_graph *out) {
  if (!out || scale < 1 || scale > 30 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0)
    return fail(CDS_ERR_INVALID, "bad R-MAT parameters");
  const uint32_t n = 1u << scale;
  const uint64_t samples = (uint64_t)edge_factor << scale;
  const double ab = a + b, abc = a + b + c;
  // Samples are drawn in fixed-size chunks, each from its own seeded stream,
  // so the result does not depend on the number of host threads.
  const uint64_t chunk = 1u << 20;
  const uint64_t nchunks = (samples + chunk - 1) / chunk;
  std::vector<uint64_t> pairs(2 * samples);
  unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::atomic<uint64_t> next_chunk{0};
  for (unsigned t = 0; t < nthreads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        uint64_t ci = next_chunk.fetch_add(1);
        if (ci >= nchunks) break;
        uint64_t sx = seed ^ (0xA0761D6478BD642Full * (ci + 1));
        Rng r(splitmix64(sx));
        const uint64_t lo = ci * chunk, hi = std::min(samples, lo + chunk);
        for (uint64_t e = lo; e < hi; ++e) {
          uint32_t u = 0, v = 0;
          for (uint32_t bit = 0; bit < scale; ++bit) {
            double x = r.uniform();
            uint32_t du = 0, dv = 0;
            if (x < a) {
            } else if (x < ab) {
              dv = 1;
            } else if (x < abc) {
              du = 1;
            } else {
              du = 1;
              dv = 1;
            }
            u = (u << 1) | du;
            v = (v << 1) | dv;
          }
          if (u == v) {
            pairs[2 * e] = pairs[2 * e + 1] = ~0ull;  // dropped below
          } else {
            pairs[2 * e] = (uint64_t)u << 32 | v;
            pairs[2 * e + 1] = (uint64_t)v << 32 | u;
          }
        }
      }
    });
  for (auto &th : pool) th.join();
  pairs.erase(std::remove(pairs.begin(), pairs.end(), ~0ull), pairs.end());
  if (planted_count) {
    Rng rng(seed ^ 0x5DEECE66Dull);
    for (uint32_t i = 0; i < planted_count; ++i) {
      uint32_t s = planted_min +
                   (uint32_t)rng.below((uint64_t)planted_max - planted_min + 1);
      plant_community(rng, n, s, planted_p, pairs);
    }
  }
  return csr_from_directed_pairs(n, pairs, out);
}
