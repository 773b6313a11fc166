// Autotune cache: launch shapes per (device architecture, CU count, k, edge
// class), one text line each:
//
//   cds-tune 1 <arch> <cus> <k> <class> <list_threads> <list_chunk>
//            <peel_blocks> <peel_coarse_shift> <replay_blocks> <flow_blocks>
//            <best_ms>
//
// (one line; '#' starts a comment line).  The class of a graph is
// floor(log2(m)) of its undirected edge count.  Host code only: the engine
// (engine.cpp) owns the sweep, this file the format, the lookup and the
// atomic merge into a file.  No counterpart in the reference, whose launch
// shapes are compile-time constants.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include "internal.h"
#include "tune.h"

namespace cds {

uint32_t edge_class(uint64_t m) {
  uint32_t c = 0;
  while (m > 1) {
    m >>= 1;
    ++c;
  }
  return c;
}

bool tuning_valid(const cds_tuning &t) {
  const bool shape0 = t.list_threads == 0 && t.list_chunk == 0;
  const bool shape =
      (t.list_threads == 256 && (t.list_chunk == 64 || t.list_chunk == 32)) ||
      (t.list_threads == 512 && (t.list_chunk == 64 || t.list_chunk == 32)) ||
      (t.list_threads == 128 && t.list_chunk == 32);
  return (shape0 || shape) && t.peel_coarse_shift >= -1 && t.peel_coarse_shift <= 8 &&
         t.peel_blocks <= 4096 && t.replay_blocks <= 4096 && t.flow_blocks <= 4096;
}

std::string tune_line(const TuneEntry &e) {
  char buf[512];
  snprintf(buf, sizeof buf, "cds-tune 1 %s %u %u %u %u %u %u %d %u %u %.4f",
           e.arch.c_str(), e.cus, e.k, e.cls, e.t.list_threads, e.t.list_chunk,
           e.t.peel_blocks, e.t.peel_coarse_shift, e.t.replay_blocks,
           e.t.flow_blocks, e.best_ms);
  return buf;
}

// Parses one non-comment line; false if malformed.
static bool parse_line(const std::string &line, TuneEntry *out) {
  std::istringstream is(line);
  std::string tag;
  int version = 0;
  TuneEntry e;
  long long v[9];
  if (!(is >> tag >> version >> e.arch) || tag != "cds-tune" || version != 1)
    return false;
  for (long long &x : v)
    if (!(is >> x)) return false;
  if (!(is >> e.best_ms)) return false;
  std::string rest;
  if (is >> rest) return false;
  for (int i = 0; i < 9; ++i)
    if ((i != 6 && v[i] < 0) || v[i] > 0xFFFFFFFFll || v[i] < -1) return false;
  e.cus = (uint32_t)v[0];
  e.k = (uint32_t)v[1];
  e.cls = (uint32_t)v[2];
  e.t.list_threads = (uint32_t)v[3];
  e.t.list_chunk = (uint32_t)v[4];
  e.t.peel_blocks = (uint32_t)v[5];
  e.t.peel_coarse_shift = (int32_t)v[6];
  e.t.replay_blocks = (uint32_t)v[7];
  e.t.flow_blocks = (uint32_t)v[8];
  if (e.k < 2 || e.k > CDS_MAX_K || e.cls > 63 || !tuning_valid(e.t) ||
      !(e.best_ms >= 0) || !std::isfinite(e.best_ms))
    return false;
  *out = e;
  return true;
}

cds_status read_tune_file(const char *path, std::vector<TuneEntry> *out,
                          bool missing_ok) {
  out->clear();
  std::ifstream f(path);
  if (!f) {
    if (missing_ok && errno == ENOENT) return CDS_OK;
    return fail(CDS_ERR_IO, "cannot open tune cache %s", path);
  }
  std::string line;
  size_t no = 0;
  while (std::getline(f, line)) {
    ++no;
    size_t a = line.find_first_not_of(" \t\r");
    if (a == std::string::npos || line[a] == '#') continue;
    TuneEntry e;
    if (!parse_line(line, &e))
      return fail(CDS_ERR_IO, "tune cache %s line %zu is malformed", path, no);
    out->push_back(e);
  }
  return CDS_OK;
}

cds_status merge_tune_file(const char *path, const TuneEntry &e) {
  std::vector<TuneEntry> all;
  CHK_STATUS(read_tune_file(path, &all, true));
  bool replaced = false;
  for (TuneEntry &x : all)
    if (x.arch == e.arch && x.cus == e.cus && x.k == e.k && x.cls == e.cls) {
      x = e;
      replaced = true;
    }
  if (!replaced) all.push_back(e);
  const std::string tmp = std::string(path) + ".tmp" + std::to_string(getpid());
  {
    std::ofstream f(tmp, std::ios::trunc);
    if (!f) return fail(CDS_ERR_IO, "cannot write %s", tmp.c_str());
    f << "# cds_amd autotune cache: arch cus k class list_threads list_chunk "
         "peel_blocks peel_coarse_shift replay_blocks flow_blocks best_ms\n";
    for (const TuneEntry &x : all) f << tune_line(x) << "\n";
    if (!f.flush()) return fail(CDS_ERR_IO, "cannot write %s", tmp.c_str());
  }
  if (rename(tmp.c_str(), path) != 0) {
    unlink(tmp.c_str());
    return fail(CDS_ERR_IO, "cannot rename %s to %s", tmp.c_str(), path);
  }
  return CDS_OK;
}

const TuneEntry *nearest_entry(const std::vector<TuneEntry> &all,
                               const std::string &arch, uint32_t cus, uint32_t k,
                               uint64_t m) {
  const uint32_t c = edge_class(m);
  const TuneEntry *best = nullptr;
  uint32_t bd = 3;
  for (const TuneEntry &x : all) {
    if (x.arch != arch || x.cus != cus || x.k != k) continue;
    const uint32_t d = x.cls > c ? x.cls - c : c - x.cls;
    // nearest class; ties go to the larger class (its shapes were sized for
    // more work, the safer side for the cooperative launches)
    if (d < bd || (d == bd && best && x.cls > best->cls)) {
      best = &x;
      bd = d;
    }
  }
  return best;
}

}  // namespace cds

using namespace cds;

extern "C" cds_status cds_tuning_lookup(const char *path, const char *arch,
                                        uint32_t cus, uint32_t k, uint64_t m,
                                        cds_tuning *out) {
  if (!path || !arch || !out) return fail(CDS_ERR_INVALID, "null argument");
  std::vector<TuneEntry> all;
  CHK_STATUS(read_tune_file(path, &all, false));
  const TuneEntry *e = nearest_entry(all, arch, cus, k, m);
  if (!e)
    return fail(CDS_ERR_INVALID, "no tuning for %s/%u k=%u near class %u", arch, cus, k,
                edge_class(m));
  *out = e->t;
  return CDS_OK;
}
