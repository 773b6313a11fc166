// Internal helpers shared by the host side of the engine.
#pragma once
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/cds.h"

namespace cds {

// Sets the thread-local message returned by cds_last_error() and returns st.
cds_status fail(cds_status st, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));

// Copies the default tuning and the loaded tune-cache entries (engine.cpp).
void copy_tuning(cds_engine *dst, cds_engine *src);
// Marks an engine as sharing its device with another one, or not
// (engine.cpp): shared engines keep cooperative kernels within half the CUs.
void share_device(cds_engine *e, bool shared);
// Whether a graph with T k-cliques is solved on its clique core instead of
// being materialised whole (engine.cpp).
bool solves_on_core(uint32_t k, uint64_t T);

// Allocates a host array with malloc (so the C ABI can free it with free()).
template <class T> T *host_alloc(size_t count) {
  if (count == 0) count = 1;
  return static_cast<T *>(malloc(count * sizeof(T)));
}

// splitmix64: seeds the xoshiro state and hashes seeds.
inline uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// xoshiro256** PRNG (Blackman & Vigna), deterministic across hosts.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &w : s) w = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t v, int k) {
    return (v << k) | (v >> (64 - k));
  }
  inline uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // Uniform double in [0, 1) with 53 random bits.
  inline double uniform() { return (next() >> 11) * 0x1.0p-53; }
  // Uniform integer in [0, bound).
  inline uint64_t below(uint64_t bound) {
    return (uint64_t)((unsigned __int128)next() * bound >> 64);
  }
};

}  // namespace cds
