// (a) k-clique listing over the degree-ordered DAG.
//
// Reference: listAllCliques (cudaCode/main.cu:59-250) with
// listIntialCliques / listMidCliques / flushParitions / countCliques /
// writeFinalCliques (cudaCode/src/helpers.cu:175-574); oracle
// listCliques (cpuCode/c++/src/helpers.cpp:373-511).  Every k-clique is
// found exactly once, from its lowest-ranked vertex u, as u < v < w (< x)
// along DAG arcs.
//
// Work decomposition: an item is (u, j0): up to `chunk` consecutive
// out-neighbours v of u (32 or 64), so a hub's out-list is spread over many
// workgroups.  One workgroup (256 or 512 lanes) per item stages the sorted
// N+(u) in LDS and flattens the item's work sum_v d+(v) across its lanes
// (coalesced reads of each N+(v)); membership of w in N+(u) is a binary
// search in LDS that also yields w's position, so per-vertex clique
// degrees are accumulated in LDS counters indexed by position and flushed
// with one global atomic per touched vertex.  k = 4 adds an LDS bitset of
// P = N+(u) & N+(v) over N+(u)'s positions and probes each N+(w), w in P.
//
// Two passes: count (per-item totals + clique degrees) and write (cliques
// stored AoS, k ids each, at the item's scanned base).
#include "dev.h"
#include "kernels.h"

namespace cds {
namespace {

__global__ void k_item_counts(uint32_t n, uint32_t chunk,
                              const uint32_t *__restrict__ doff,
                              uint32_t *__restrict__ ipv) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x) {
    const uint32_t d = doff[v + 1] - doff[v];
    ipv[v] = (d + chunk - 1) / chunk;
  }
}

__global__ void k_item_fill(uint32_t n, uint32_t chunk,
                            const uint32_t *__restrict__ doff,
                            const uint64_t *__restrict__ ioff,
                            uint2 *__restrict__ items) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x) {
    const uint32_t d = doff[v + 1] - doff[v];
    const uint32_t nb = (d + chunk - 1) / chunk;
    const uint64_t base = ioff[v];
    for (uint32_t t = 0; t < nb; ++t)
      items[base + t] = make_uint2(v, t * chunk);
  }
}

template <int CHUNK, int WAVES>
struct ChunkTable {
  static_assert(CHUNK <= kWave, "the chunk scan is one wave");
  uint32_t pre[CHUNK + 1];  // prefix sums of d+(v) over the chunk
  uint32_t vid[CHUNK];
  uint32_t vbase[CHUNK];
  uint32_t wave_found[WAVES];
  uint32_t cursor;
  unsigned long long bound;  // sum over the chunk of min(d+(v), d+(u))
  unsigned long long sbase;  // staging: the item's reservation
};

constexpr unsigned long long kNoStage = ~0ull;

// Value ranges of a staged N+(u) for the k = 3 probes (k_list3), and when
// an item uses them: a long list and many probes (building the table costs
// each lane about one search).
constexpr uint32_t kListBucketBits = 8;
constexpr uint32_t kListBuckets = 1u << kListBucketBits;
constexpr uint32_t kListBucketMinDegree = 64;
constexpr uint32_t kListBucketMinProbes = 2048;

// First wave: load the chunk's v ids and scan their out-degrees.
template <class Table>
__device__ __forceinline__ void load_chunk(Table &t, const uint32_t *dadj,
                                           const uint32_t *doff, uint32_t ub,
                                           uint32_t j0, uint32_t jn, uint32_t du) {
  if (threadIdx.x < kWave) {
    uint32_t dv = 0;
    if (threadIdx.x < jn) {
      const uint32_t v = dadj[ub + j0 + threadIdx.x];
      const uint32_t vb = doff[v];
      dv = doff[v + 1] - vb;
      t.vid[threadIdx.x] = v;
      t.vbase[threadIdx.x] = vb;
    }
    uint32_t incl = dv;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      uint32_t x = __shfl_up(incl, off, kWave);
      if ((int)lane() >= off) incl += x;
    }
    if (threadIdx.x < jn) t.pre[threadIdx.x + 1] = incl;
    const uint64_t b = wave_sum64(dv < du ? dv : du);
    if (threadIdx.x == 0) {
      t.pre[0] = 0;
      t.cursor = 0;
      t.bound = b;
    }
  }
}

// Finishes an item in count mode: item total, cd[u], flush LDS counters.
template <int WAVES, class Table>
__device__ __forceinline__ void finish_count(
    Table &t, uint32_t found, uint64_t it, uint32_t u, bool staged,
    const uint32_t *su, const uint32_t *scnt, uint32_t du,
    uint32_t *item_count, unsigned long long *cd) {
  found = wave_sum(found);
  if (lane() == 0) t.wave_found[threadIdx.x / kWave] = found;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int w = 0; w < WAVES; ++w) tot += t.wave_found[w];
    item_count[it] = tot;
    if (tot) atomicAdd(&cd[u], (unsigned long long)tot);
  }
  if (staged)
    for (uint32_t i = threadIdx.x; i < du; i += blockDim.x) {
      const uint32_t c = scnt[i];
      if (c) atomicAdd(&cd[su[i]], (unsigned long long)c);
    }
}

// A triangle's three ids stored with one 12-byte instruction.
struct __attribute__((packed, aligned(4))) Clique3 {
  uint32_t w[3];
};

template <int THREADS, int CHUNK>
__global__ __launch_bounds__(THREADS) void k_list3(
    uint64_t num_items, const uint2 *__restrict__ items,
    const uint32_t *__restrict__ doff, const uint32_t *__restrict__ dadj,
    uint32_t lds_cap, uint32_t *__restrict__ item_count,
    unsigned long long *__restrict__ cd, const uint64_t *__restrict__ item_base,
    uint32_t *__restrict__ out, ListStage st) {
  extern __shared__ uint32_t dyn[];
  uint32_t *su = dyn;             // N+(u), sorted
  uint32_t *scnt = dyn + lds_cap; // clique-degree counters by position
  constexpr int WAVES = THREADS / kWave;
  __shared__ ChunkTable<CHUNK, WAVES> t;
  __shared__ uint32_t bstart[kListBuckets + 1];
  const bool write = out != nullptr;
  for (uint64_t it = blockIdx.x; it < num_items; it += gridDim.x) {
    const uint2 item = items[it];
    const uint32_t u = item.x, j0 = item.y;
    const uint32_t ub = doff[u], du = doff[u + 1] - ub;
    const uint32_t jn = min(du - j0, (uint32_t)CHUNK);
    const bool staged = du <= lds_cap;
    if (staged)
      for (uint32_t i = threadIdx.x; i < du; i += blockDim.x) {
        su[i] = dadj[ub + i];
        scnt[i] = 0;
      }
    load_chunk(t, dadj, doff, ub, j0, jn, du);
    __syncthreads();
    const bool stage = st.buf && !write;
    if (stage) {
      if (threadIdx.x == 0) {  // the item's worst case, reserved at once
        const unsigned long long b = t.bound;
        const unsigned long long at = b ? atomicAdd(st.cursor, b) : 0ull;
        t.sbase = at + b <= st.cap ? at : kNoStage;
        if (t.sbase == kNoStage) *st.over = 1;
        st.item_tmp[it] = t.sbase;
      }
      __syncthreads();
    }
    const uint32_t total = t.pre[jn];
    const uint32_t *nu = staged ? su : dadj + ub;
    const uint64_t obase = write ? item_base[it] : 0;
    // Items with many probes against a long N+(u) first cut its values into
    // kListBuckets ranges of 2^bsh ids (bstart[b] = first position in range
    // b), so a probe searches one range instead of the whole list.
    const bool bucketed = staged && du >= kListBucketMinDegree && total >= kListBucketMinProbes;
    uint32_t blo = 0, bspan = 0, bsh = 0;
    if (bucketed) {
      blo = su[0];
      bspan = su[du - 1] - blo;
      const uint32_t bits = bspan ? 32u - (uint32_t)__clz(bspan) : 0u;
      bsh = bits > kListBucketBits ? bits - kListBucketBits : 0u;
      for (uint32_t b = threadIdx.x; b <= kListBuckets; b += blockDim.x) {
        const uint64_t thr = (uint64_t)b << bsh;
        int l = 0, h = (int)du;
        while (l < h) {
          const int mid = (l + h) >> 1;
          if ((uint64_t)(su[mid] - blo) < thr)
            l = mid + 1;
          else
            h = mid;
        }
        bstart[b] = (uint32_t)l;
      }
      __syncthreads();
    }
    uint32_t found = 0;
    // lo = the last vi with pre[vi] <= i; i only grows, so lo moves on from
    // where it was (pre[jn] = total > i bounds it) and is searched for only
    // when i has passed its out-list
    int lo = 0;
    for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
      if (t.pre[lo + 1] <= i) {
        int l = lo + 1, hi = (int)jn - 1;
        while (l < hi) {
          const int mid = (l + hi + 1) >> 1;
          if (t.pre[mid] <= i)
            l = mid;
          else
            hi = mid - 1;
        }
        lo = l;
      }
      const uint32_t w = dadj[t.vbase[lo] + (i - t.pre[lo])];
      int p;
      if (bucketed) {
        const uint32_t off = w - blo;  // wraps past bspan when w < blo
        p = -1;
        if (off <= bspan) {
          const uint32_t b = off >> bsh, l0 = bstart[b];
          const int r = find_sorted(su + l0, (int)(bstart[b + 1] - l0), w);
          if (r >= 0) p = (int)l0 + r;
        }
      } else {
        p = find_sorted(nu, (int)du, w);
      }
      if (p < 0) continue;
      ++found;
      if (write) {
        const uint32_t slot = atomicAdd(&t.cursor, 1u);
        *reinterpret_cast<Clique3 *>(out + 3ull * (obase + slot)) = Clique3{{u, t.vid[lo], w}};
      } else if (staged) {
        atomicAdd(&scnt[p], 1u);
        atomicAdd(&scnt[j0 + lo], 1u);
        if (stage && t.sbase != kNoStage) {
          const uint32_t slot = atomicAdd(&t.cursor, 1u);
          *reinterpret_cast<Clique3 *>(st.buf + 3ull * (t.sbase + slot)) =
              Clique3{{u, t.vid[lo], w}};
        }
      } else {
        atomicAdd(&cd[w], 1ull);
        atomicAdd(&cd[t.vid[lo]], 1ull);
        if (stage && t.sbase != kNoStage) {
          const uint32_t slot = atomicAdd(&t.cursor, 1u);
          *reinterpret_cast<Clique3 *>(st.buf + 3ull * (t.sbase + slot)) =
              Clique3{{u, t.vid[lo], w}};
        }
      }
    }
    if (!write) finish_count<WAVES>(t, found, it, u, staged, su, scnt, du, item_count, cd);
    __syncthreads();
  }
}

// k = 4.  Requires du <= lds_cap for every u (checked by the host).
template <int THREADS, int CHUNK>
__global__ __launch_bounds__(THREADS) void k_list4(
    uint64_t num_items, const uint2 *__restrict__ items,
    const uint32_t *__restrict__ doff, const uint32_t *__restrict__ dadj,
    uint32_t lds_cap, uint32_t *__restrict__ item_count,
    unsigned long long *__restrict__ cd, const uint64_t *__restrict__ item_base,
    uint32_t *__restrict__ out) {
  extern __shared__ uint32_t dyn[];
  uint32_t *su = dyn;
  uint32_t *scnt = dyn + lds_cap;
  uint32_t *sbits = dyn + 2 * lds_cap;               // lds_cap/32 + 1 words
  uint32_t *plist = sbits + (lds_cap / 32 + 1);      // lds_cap
  constexpr int WAVES = THREADS / kWave;
  __shared__ ChunkTable<CHUNK, WAVES> t;
  __shared__ uint32_t s_np;
  const bool write = out != nullptr;
  const uint32_t wid = threadIdx.x / kWave;
  for (uint64_t it = blockIdx.x; it < num_items; it += gridDim.x) {
    const uint2 item = items[it];
    const uint32_t u = item.x, j0 = item.y;
    const uint32_t ub = doff[u], du = doff[u + 1] - ub;
    const uint32_t jn = min(du - j0, (uint32_t)CHUNK);
    for (uint32_t i = threadIdx.x; i < du; i += blockDim.x) {
      su[i] = dadj[ub + i];
      scnt[i] = 0;
    }
    if (threadIdx.x == 0) t.cursor = 0;
    __syncthreads();
    const uint64_t obase = write ? item_base[it] : 0;
    const uint32_t nwords = (du + 31) / 32;
    uint32_t found = 0;
    for (uint32_t vi = 0; vi < jn; ++vi) {
      const uint32_t pv = j0 + vi;
      const uint32_t v = su[pv];
      const uint32_t vb = doff[v], dv = doff[v + 1] - vb;
      for (uint32_t i = threadIdx.x; i < nwords; i += blockDim.x) sbits[i] = 0;
      if (threadIdx.x == 0) s_np = 0;
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < dv; i += blockDim.x) {
        const int p = find_sorted(su, (int)du, dadj[vb + i]);
        if (p >= 0) {
          atomicOr(&sbits[p >> 5], 1u << (p & 31));
          plist[atomicAdd(&s_np, 1u)] = (uint32_t)p;
        }
      }
      __syncthreads();
      const uint32_t np = s_np;
      for (uint32_t q = wid; q < np; q += WAVES) {
        const uint32_t pw = plist[q];
        const uint32_t w = su[pw];
        const uint32_t wb = doff[w], dw = doff[w + 1] - wb;
        for (uint32_t r = lane(); r < dw; r += kWave) {
          const uint32_t x = dadj[wb + r];
          const int px = find_sorted(su, (int)du, x);
          if (px < 0 || !((sbits[px >> 5] >> (px & 31)) & 1u)) continue;
          ++found;
          if (write) {
            const uint32_t slot = atomicAdd(&t.cursor, 1u);
            uint32_t *o = out + 4ull * (obase + slot);
            o[0] = u;
            o[1] = v;
            o[2] = w;
            o[3] = x;
          } else {
            atomicAdd(&scnt[px], 1u);
            atomicAdd(&scnt[pw], 1u);
            atomicAdd(&scnt[pv], 1u);
          }
        }
      }
      __syncthreads();
    }
    if (!write) finish_count<WAVES>(t, found, it, u, true, su, scnt, du, item_count, cd);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_list2(
    uint32_t n, const uint32_t *__restrict__ doff,
    const uint32_t *__restrict__ dadj, unsigned long long *__restrict__ cd,
    uint32_t *__restrict__ out) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       u < n; u += waves) {
    const uint32_t b = doff[u], e = doff[u + 1];
    for (uint32_t j = b + lane(); j < e; j += kWave) {
      const uint32_t v = dadj[j];
      if (out) {
        out[2ull * j] = u;
        out[2ull * j + 1] = v;
      } else {
        atomicAdd(&cd[v], 1ull);
      }
    }
    if (!out && lane() == 0 && e > b) atomicAdd(&cd[u], (unsigned long long)(e - b));
  }
}

}  // namespace

hipError_t list_item_counts(uint32_t n, uint32_t chunk, const uint32_t *doff,
                            uint32_t *items_per_vertex, hipStream_t s) {
  if (chunk != 32 && chunk != 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_item_counts, dim3(grid_for(n, 256, 8192)), dim3(256), 0,
                     s, n, chunk, doff, items_per_vertex);
  return hipGetLastError();
}

hipError_t list_item_fill(uint32_t n, uint32_t chunk, const uint32_t *doff,
                          const uint64_t *item_off, uint2 *items,
                          hipStream_t s) {
  if (chunk != 32 && chunk != 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_item_fill, dim3(grid_for(n, 256, 8192)), dim3(256), 0,
                     s, n, chunk, doff, item_off, items);
  return hipGetLastError();
}

namespace {
template <int THREADS, int CHUNK>
hipError_t launch_list(uint32_t k, uint64_t num_items, const uint2 *items,
                       const uint32_t *doff, const uint32_t *dadj,
                       uint32_t max_out_degree, uint32_t *item_count,
                       unsigned long long *cd, const uint64_t *item_base,
                       uint32_t *out, hipStream_t s, ListStage st) {
  const unsigned grid = (unsigned)(num_items < 16384 ? num_items : 16384);
  if (k == 3) {
    const uint32_t cap = max_out_degree < kListLdsCap ? max_out_degree : kListLdsCap;
    const size_t lds = 2ull * (cap ? cap : 1) * sizeof(uint32_t);
    hipLaunchKernelGGL((k_list3<THREADS, CHUNK>), dim3(grid), dim3(THREADS), lds, s,
                       num_items, items, doff, dadj, cap ? cap : 1, item_count,
                       cd, item_base, out, st);
    return hipGetLastError();
  }
  if (st.buf) return hipErrorInvalidValue;  // staging is for triangles only
  if (k == 4) {
    const uint32_t cap = max_out_degree ? max_out_degree : 1;
    if (cap > kList4LdsCap) return hipErrorInvalidValue;
    const size_t lds = (3ull * cap + cap / 32 + 1) * sizeof(uint32_t);
    hipLaunchKernelGGL((k_list4<THREADS, CHUNK>), dim3(grid), dim3(THREADS), lds, s,
                       num_items, items, doff, dadj, cap, item_count, cd,
                       item_base, out);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}
}  // namespace

hipError_t list_cliques(uint32_t k, ListShape shape, uint64_t num_items,
                        const uint2 *items, const uint32_t *doff,
                        const uint32_t *dadj, uint32_t max_out_degree,
                        uint32_t *item_count, uint64_t *clique_degree,
                        const uint64_t *item_base, uint32_t *out, hipStream_t s,
                        ListStage st) {
  if (num_items == 0) return hipSuccess;
  auto *cd = reinterpret_cast<unsigned long long *>(clique_degree);
#define CDS_LIST_SHAPE(TH, CH)                                                      \
  if (shape.threads == TH && shape.chunk == CH)                                     \
    return launch_list<TH, CH>(k, num_items, items, doff, dadj, max_out_degree,     \
                               item_count, cd, item_base, out, s, st);
  CDS_LIST_SHAPE(256, 64)
  CDS_LIST_SHAPE(256, 32)
  CDS_LIST_SHAPE(512, 64)
  CDS_LIST_SHAPE(512, 32)
  CDS_LIST_SHAPE(128, 32)
#undef CDS_LIST_SHAPE
  return hipErrorInvalidValue;
}

namespace {
// One wave per item: its staged triangles, contiguous, to their place.
__global__ __launch_bounds__(256) void k_list_unstage(uint64_t num_items,
                                                      const uint64_t *__restrict__ item_tmp,
                                                      const uint32_t *__restrict__ item_count,
                                                      const uint64_t *__restrict__ item_base,
                                                      const uint32_t *__restrict__ stage,
                                                      uint32_t *__restrict__ out) {
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
  for (uint64_t it = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       it < num_items; it += waves) {
    const uint64_t words = 3ull * item_count[it];
    const uint32_t *src = stage + 3ull * item_tmp[it];
    uint32_t *dst = out + 3ull * item_base[it];
    for (uint64_t i = lane(); i < words; i += kWave) dst[i] = src[i];
  }
}
}  // namespace

hipError_t list_unstage(uint64_t num_items, const uint64_t *item_tmp,
                        const uint32_t *item_count, const uint64_t *item_base,
                        const uint32_t *stage, uint32_t *out, hipStream_t s) {
  if (num_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_list_unstage, dim3(grid_for(num_items * kWave, 256, 16384)), dim3(256),
                     0, s, num_items, item_tmp, item_count, item_base, stage, out);
  return hipGetLastError();
}

hipError_t list_edges(uint32_t n, const uint32_t *doff, const uint32_t *dadj,
                      uint64_t *clique_degree, uint32_t *out, hipStream_t s) {
  hipLaunchKernelGGL(k_list2, dim3(grid_for((size_t)n * kWave, 256, 65535)),
                     dim3(256), 0, s, n, doff, dadj,
                     reinterpret_cast<unsigned long long *>(clique_degree), out);
  return hipGetLastError();
}

}  // namespace cds
