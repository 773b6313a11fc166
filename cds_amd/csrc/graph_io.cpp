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
    case 4: return sizeof(cds_tuning);
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

// ---------------------------------------------------------------------------
// Parallel text loader
// ---------------------------------------------------------------------------

// Host threads for parsing and CSR assembly: the hardware threads, at most
// 16, at least 1 MiB of text or 64K entries per thread; CDS_IO_THREADS
// overrides.  The result never depends on the thread count.
static unsigned io_threads(size_t units, size_t per_thread) {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 1;
  if (t > 16) t = 16;
  if (const char *e = getenv("CDS_IO_THREADS")) {
    const long x = strtol(e, nullptr, 10);
    if (x >= 1 && x <= 256) t = (unsigned)x;
  }
  const size_t cap = units / per_thread + 1;
  return t > cap ? (unsigned)cap : t;
}

template <class F>
static void run_threads(unsigned nt, F f) {
  if (nt <= 1) {
    f(0u);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(f, t);
  f(0u);
  for (auto &x : th) x.join();
}

namespace {

// The first malformed token a parser thread met: its byte offset decides
// which error is reported (the earliest in the file, as a serial scan would).
struct ParseError {
  size_t at = SIZE_MAX;
  cds_status st = CDS_OK;
  uint64_t a = 0;  // offending id, if any
  int kind = 0;    // 1 expected vertex id, 2 vertex >= n, 3 bad token, 4 neighbour >= n
};

// Lines of one text chunk, as records [v, count, w_1 .. w_count].
struct ChunkLists {
  std::vector<uint32_t> rec;
  ParseError err;
};

}  // namespace

// Parses the vertex lines in [b, e); b is at a line start, e at a line start
// or the end of the text.  Tokens are separated by spaces, tabs or '\r'.
static void parse_chunk(const char *text, const char *b, const char *e,
                        uint32_t n, ChunkLists *out) {
  const char *p = b;
  auto skip_blank = [&]() {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  };
  auto read_uint = [&](uint64_t &v) -> bool {
    if (p >= e || *p < '0' || *p > '9') return false;
    uint64_t x = 0;
    while (p < e && *p >= '0' && *p <= '9') {
      x = x * 10 + (uint64_t)(*p - '0');
      if (x > 0xFFFFFFFFull) return false;
      ++p;
    }
    v = x;
    return true;
  };
  auto error = [&](int kind, uint64_t a, const char *at) {
    out->err.at = (size_t)(at - text);
    out->err.kind = kind;
    out->err.a = a;
    out->err.st = (kind == 2 || kind == 4) ? CDS_ERR_INVALID : CDS_ERR_IO;
  };
  std::vector<uint32_t> &rec = out->rec;
  while (p < e) {
    if (*p == '\n') {
      ++p;
      continue;
    }
    skip_blank();
    if (p >= e || *p == '\n') continue;
    const char *tok = p;
    uint64_t v;
    if (!read_uint(v)) return error(1, 0, tok);
    if (v >= n) return error(2, v, tok);
    const size_t head = rec.size();
    rec.push_back((uint32_t)v);
    rec.push_back(0);
    for (;;) {
      skip_blank();
      if (p >= e || *p == '\n') break;
      tok = p;
      uint64_t w;
      if (!read_uint(w)) return error(3, 0, tok);
      if (w >= n) return error(4, w, tok);
      if (w == v) continue;  // self loops are skipped like the reference
      rec.push_back((uint32_t)w);
    }
    rec[head + 1] = (uint32_t)(rec.size() - head - 2);
  }
}

// Assembles the CSR from the parsed line records: per-vertex lists are
// gathered (several lines of one vertex are merged), sorted and
// de-duplicated, and every entry's reverse is checked.  Errors name the
// smallest offending (v, w) so the report is the same for any thread count.
static cds_status assemble_csr(uint32_t n, std::vector<ChunkLists> &chunks,
                               const char *path, cds_graph *out) {
  // raw counts per vertex, then list starts
  std::vector<uint64_t> start((size_t)n + 1, 0);
  uint64_t total = 0;
  const unsigned nc = (unsigned)chunks.size();
  run_threads(nc, [&](unsigned t) {
    const std::vector<uint32_t> &r = chunks[t].rec;
    for (size_t q = 0; q < r.size(); q += 2 + r[q + 1])
      __atomic_fetch_add(&start[(size_t)r[q] + 1], (uint64_t)r[q + 1], __ATOMIC_RELAXED);
  });
  for (uint32_t v = 0; v < n; ++v) start[v + 1] += start[v];
  total = start[n];
  std::vector<uint32_t> raw(total);
  {
    std::vector<uint64_t> cursor(start.begin(), start.end() - 1);
    run_threads(nc, [&](unsigned t) {
      const std::vector<uint32_t> &r = chunks[t].rec;
      for (size_t q = 0; q < r.size(); q += 2 + r[q + 1]) {
        const uint32_t c = r[q + 1];
        if (!c) continue;
        const uint64_t at = __atomic_fetch_add(&cursor[r[q]], (uint64_t)c, __ATOMIC_RELAXED);
        memcpy(raw.data() + at, r.data() + q + 2, (size_t)c * sizeof(uint32_t));
      }
    });
  }
  for (auto &c : chunks) std::vector<uint32_t>().swap(c.rec);
  // vertex ranges of about equal entry counts
  const unsigned nt = io_threads(total + n, 1 << 16);
  std::vector<uint32_t> vcut(nt + 1, n);
  vcut[0] = 0;
  for (unsigned t = 1; t < nt; ++t) {
    const uint64_t want = (total + n) * t / nt;
    uint32_t lo = vcut[t - 1], hi = n;  // first v with start[v] + v >= want
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (start[mid] + mid >= want)
        hi = mid;
      else
        lo = mid + 1;
    }
    vcut[t] = lo;
  }
  std::vector<uint32_t> len(n);
  run_threads(nt, [&](unsigned t) {
    for (uint32_t v = vcut[t]; v < vcut[t + 1]; ++v) {
      uint32_t *b = raw.data() + start[v], *e = raw.data() + start[v + 1];
      std::sort(b, e);
      len[v] = (uint32_t)(std::unique(b, e) - b);
    }
  });
  // symmetry: the first violation of each thread's range, smallest wins
  std::vector<uint64_t> bad(nt, UINT64_MAX);
  run_threads(nt, [&](unsigned t) {
    for (uint32_t v = vcut[t]; v < vcut[t + 1]; ++v) {
      const uint32_t *a = raw.data() + start[v];
      for (uint32_t j = 0; j < len[v]; ++j) {
        const uint32_t w = a[j];
        const uint32_t *bw = raw.data() + start[w];
        if (!std::binary_search(bw, bw + len[w], v)) {
          bad[t] = (uint64_t)v << 32 | w;
          return;
        }
      }
    }
  });
  for (uint64_t b : bad)
    if (b != UINT64_MAX)
      return fail(CDS_ERR_INVALID,
                  "%s: adjacency is not symmetric: %u lists %u but not back",
                  path, (uint32_t)(b >> 32), (uint32_t)b);
  uint64_t kept = 0;
  uint32_t maxd = 0;
  for (uint32_t v = 0; v < n; ++v) {
    kept += len[v];
    maxd = std::max(maxd, len[v]);
  }
  if (kept >= (1ull << 32))
    return fail(CDS_ERR_RANGE, "graph has %llu adjacency entries (>= 2^32)",
                (unsigned long long)kept);
  uint32_t *off = host_alloc<uint32_t>((size_t)n + 1);
  uint32_t *nb = host_alloc<uint32_t>(kept ? kept : 1);
  if (!off || !nb) {
    free(off);
    free(nb);
    return fail(CDS_ERR_OOM, "host allocation failed");
  }
  off[0] = 0;
  for (uint32_t v = 0; v < n; ++v) off[v + 1] = off[v] + len[v];
  run_threads(nt, [&](unsigned t) {
    for (uint32_t v = vcut[t]; v < vcut[t + 1]; ++v)
      memcpy(nb + off[v], raw.data() + start[v], (size_t)len[v] * sizeof(uint32_t));
  });
  out->n = n;
  out->m = kept / 2;
  out->m_header = 0;
  out->offsets = off;
  out->neighbors = nb;
  out->max_degree = maxd;
  return CDS_OK;
}

// Parses the adjacency-list format with a hand-written tokenizer (the
// reference uses istringstream per line, cudaCode/src/graph.cpp:24-35): the
// text is cut at line starts into one chunk per thread, chunks are parsed in
// parallel, and the lists are assembled into the CSR in parallel.
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
  const char *text = buf.data();
  const char *p = text;
  const char *end = p + buf.size() - 1;
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
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
  if (!read_uint(n64)) return fail(CDS_ERR_IO, "%s: missing header n", path);
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  if (!read_uint(m64)) m64 = 0;  // a header with n only is tolerated
  while (p < end && *p != '\n') ++p;
  const uint32_t n = (uint32_t)n64;
  // chunk boundaries at line starts
  const unsigned nt = io_threads((size_t)(end - p), 1 << 20);
  std::vector<const char *> cut(nt + 1, end);
  cut[0] = p;
  for (unsigned t = 1; t < nt; ++t) {
    const char *c = p + (size_t)(end - p) * t / nt;
    if (c < cut[t - 1]) c = cut[t - 1];
    while (c < end && c[-1] != '\n') ++c;
    cut[t] = c;
  }
  std::vector<ChunkLists> chunks(nt);
  run_threads(nt, [&](unsigned t) { parse_chunk(text, cut[t], cut[t + 1], n, &chunks[t]); });
  const ParseError *first = nullptr;
  for (const auto &c : chunks)
    if (c.err.st != CDS_OK && (!first || c.err.at < first->at)) first = &c.err;
  if (first) {
    const unsigned long long line =
        1 + (unsigned long long)std::count(text, text + first->at, '\n');
    switch (first->kind) {
      case 1:
        return fail(CDS_ERR_IO, "%s:%llu: expected a vertex id", path, line);
      case 2:
        return fail(CDS_ERR_INVALID, "%s:%llu: vertex %llu >= n=%u", path, line,
                    (unsigned long long)first->a, n);
      case 3:
        return fail(CDS_ERR_IO, "%s:%llu: bad neighbor token", path, line);
      default:
        return fail(CDS_ERR_INVALID, "%s:%llu: neighbor %llu >= n=%u", path, line,
                    (unsigned long long)first->a, n);
    }
  }
  cds_status st = assemble_csr(n, chunks, path, out);
  if (st == CDS_OK) out->m_header = m64;
  return st;
}

// ---------------------------------------------------------------------------
// Binary CSR file
// ---------------------------------------------------------------------------

static const char kCsrMagic[8] = {'C', 'D', 'S', 'C', 'S', 'R', 0, 1};

namespace {
struct CsrHeader {
  char magic[8];
  uint32_t n;
  uint32_t max_degree;
  uint64_t m;
  uint64_t m_header;
};
}  // namespace

extern "C" cds_status cds_graph_save_csr(const cds_graph *g, const char *path) {
  if (!g || !path || (g->n && !g->offsets)) return fail(CDS_ERR_INVALID, "null argument");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(CDS_ERR_IO, "cannot open %s for writing", path);
  CsrHeader h{};
  memcpy(h.magic, kCsrMagic, sizeof h.magic);
  h.n = g->n;
  h.max_degree = g->max_degree;
  h.m = g->m;
  h.m_header = g->m_header;
  static const uint32_t zero = 0;
  const uint32_t *off = g->n ? g->offsets : &zero;
  const uint64_t entries = g->n ? g->offsets[g->n] : 0;
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 &&
            fwrite(off, sizeof(uint32_t), (size_t)g->n + 1, f) == (size_t)g->n + 1 &&
            (entries == 0 ||
             fwrite(g->neighbors, sizeof(uint32_t), entries, f) == entries);
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(CDS_ERR_IO, "write failed: %s", path);
  return CDS_OK;
}

// Loads a file written by cds_graph_save_csr and checks every invariant the
// engine relies on (monotone offsets, ids < n, strictly ascending lists
// without self loops, symmetry) before handing the graph out.
extern "C" cds_status cds_graph_load_csr(const char *path, cds_graph *out) {
  if (!path || !out) return fail(CDS_ERR_INVALID, "null argument");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(CDS_ERR_IO, "Graph file Open Failed: %s", path);
  CsrHeader h{};
  fseek(f, 0, SEEK_END);
  const long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (sz < (long)sizeof h || fread(&h, sizeof h, 1, f) != 1 ||
      memcmp(h.magic, kCsrMagic, sizeof h.magic) != 0) {
    fclose(f);
    return fail(CDS_ERR_IO, "%s: not a binary CSR graph file", path);
  }
  const uint64_t want_min = sizeof h + 4ull * ((uint64_t)h.n + 1);
  if ((uint64_t)sz < want_min || 2 * h.m >= (1ull << 32)) {
    fclose(f);
    return fail(CDS_ERR_IO, "%s: truncated or inconsistent CSR header", path);
  }
  const uint64_t entries = 2 * h.m;
  if ((uint64_t)sz != want_min + 4 * entries) {
    fclose(f);
    return fail(CDS_ERR_IO, "%s: size %ld does not match n=%u m=%llu", path, sz,
                h.n, (unsigned long long)h.m);
  }
  uint32_t *off = host_alloc<uint32_t>((size_t)h.n + 1);
  uint32_t *nb = host_alloc<uint32_t>(entries ? entries : 1);
  if (!off || !nb) {
    fclose(f);
    free(off);
    free(nb);
    return fail(CDS_ERR_OOM, "host allocation failed");
  }
  bool ok = fread(off, sizeof(uint32_t), (size_t)h.n + 1, f) == (size_t)h.n + 1 &&
            (entries == 0 || fread(nb, sizeof(uint32_t), entries, f) == entries);
  fclose(f);
  auto reject = [&](const char *what, uint32_t v) {
    free(off);
    free(nb);
    return fail(CDS_ERR_IO, "%s: invalid CSR (%s at vertex %u)", path, what, v);
  };
  if (!ok) return reject("short read", 0);
  if (off[0] != 0 || off[h.n] != entries) return reject("offsets do not span the lists", 0);
  uint32_t maxd = 0;
  for (uint32_t v = 0; v < h.n; ++v) {
    if (off[v + 1] < off[v]) return reject("offsets not monotone", v);
    maxd = std::max(maxd, off[v + 1] - off[v]);
  }
  if (maxd != h.max_degree) return reject("max degree differs from the header", 0);
  const unsigned nt = io_threads(entries + h.n, 1 << 16);
  std::vector<uint64_t> bad(nt, UINT64_MAX);
  run_threads(nt, [&](unsigned t) {
    const uint32_t v0 = (uint32_t)((uint64_t)h.n * t / nt);
    const uint32_t v1 = (uint32_t)((uint64_t)h.n * (t + 1) / nt);
    for (uint32_t v = v0; v < v1; ++v)
      for (uint32_t j = off[v]; j < off[v + 1]; ++j) {
        const uint32_t w = nb[j];
        const bool fine = w < h.n && w != v && (j == off[v] || nb[j - 1] < w) &&
                          std::binary_search(nb + off[w], nb + off[w + 1], v);
        if (!fine) {
          bad[t] = v;
          return;
        }
      }
  });
  for (uint64_t b : bad)
    if (b != UINT64_MAX)
      return reject("list out of range, unsorted, self loop or asymmetric", (uint32_t)b);
  out->n = h.n;
  out->m = h.m;
  out->m_header = h.m_header;
  out->offsets = off;
  out->neighbors = nb;
  out->max_degree = maxd;
  return CDS_OK;
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
