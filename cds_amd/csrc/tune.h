// Autotune cache entries and file access (tune.cpp); host code only.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cds.h"

#define CHK_STATUS(expr)                \
  do {                                  \
    cds_status s_ = (expr);             \
    if (s_ != CDS_OK) return s_;        \
  } while (0)

namespace cds {

struct TuneEntry {
  std::string arch;  // gcnArchName without feature suffixes, e.g. gfx950
  uint32_t cus = 0, k = 0, cls = 0;
  cds_tuning t{};
  double best_ms = 0;
};

// floor(log2(m)), 0 for m <= 1.
uint32_t edge_class(uint64_t m);
// Field ranges and the listing shapes the kernels are built for.
bool tuning_valid(const cds_tuning &t);
std::string tune_line(const TuneEntry &e);
// Every entry of a cache file; a missing file is empty when missing_ok.
cds_status read_tune_file(const char *path, std::vector<TuneEntry> *out,
                          bool missing_ok);
// Replaces or appends e's (arch, cus, k, class) entry, atomically (rename).
cds_status merge_tune_file(const char *path, const TuneEntry &e);
// The entry of (arch, cus, k) whose class is nearest to m's, within 2.
const TuneEntry *nearest_entry(const std::vector<TuneEntry> &all,
                               const std::string &arch, uint32_t cus,
                               uint32_t k, uint64_t m);

}  // namespace cds
