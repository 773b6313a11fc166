// Degree-ordered DAG (orientation) on the device.
//
// The reference orients by a host-computed core (degeneracy) order
// (cudaCode/src/graph.cpp:41-113, generateDegreeDAG/generateNeighborDAG at
// cudaCode/src/helpers.cu:112-173; oracle cpuCode/c++/src/helpers.cpp:349-371).
// Orientation does not change any clique count, degree, core number or
// flow; it only bounds the out-degree.  Here u -> v iff
// (deg u, u) < (deg v, v), which bounds d+(u) by sqrt(2m) and is computed in
// one pass without a host peel.  One wave per vertex; out-lists keep the
// ascending id order of the input lists (order-preserving ballot
// compaction), so every N+(u) is sorted.
#include "dev.h"
#include "kernels.h"

namespace cds {
namespace {

__device__ __forceinline__ bool points_up(uint32_t v, uint32_t dv, uint32_t w,
                                          uint32_t dw) {
  return dw > dv || (dw == dv && w > v);
}

__global__ __launch_bounds__(256) void k_dag_count(
    uint32_t n, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ adj, uint32_t *__restrict__ dcnt) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t v = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       v < n; v += waves) {
    const uint32_t b = off[v], e = off[v + 1], dv = e - b;
    uint32_t c = 0;
    for (uint32_t j = b + lane(); j < e; j += kWave) {
      const uint32_t w = adj[j];
      c += points_up(v, dv, w, off[w + 1] - off[w]);
    }
    c = wave_sum(c);
    if (lane() == 0) dcnt[v] = c;
  }
}

__global__ __launch_bounds__(256) void k_dag_fill(
    uint32_t n, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ adj, const uint32_t *__restrict__ doff,
    uint32_t *__restrict__ dadj) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t v = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       v < n; v += waves) {
    const uint32_t b = off[v], e = off[v + 1], dv = e - b;
    uint32_t out = doff[v];
    for (uint32_t j = b; j < e; j += kWave) {
      const uint32_t jj = j + lane();
      bool up = false;
      uint32_t w = 0;
      if (jj < e) {
        w = adj[jj];
        up = points_up(v, dv, w, off[w + 1] - off[w]);
      }
      unsigned cnt;
      const unsigned slot = wave_slot(up, &cnt);
      if (up) dadj[out + slot] = w;
      out += cnt;
    }
  }
}

// out[0] = max d+(v) (as a 64-bit max), out[1] = sum_v (deg(v) - d+(v)) d+(v),
// the listing work W: every arc (u, v) contributes d+(v) candidates;
// out[2] = sum over arcs (u, v) of min(d+(u), d+(v)), a bound on the
// triangles (each arc closes at most that many).  One wave per vertex u.
__global__ __launch_bounds__(256) void k_dag_stats(uint32_t n, const uint32_t *__restrict__ off,
                                                   const uint32_t *__restrict__ dcnt,
                                                   const uint32_t *__restrict__ doff,
                                                   const uint32_t *__restrict__ dadj,
                                                   unsigned long long *out) {
  unsigned long long mx = 0, w = 0, tb = 0;
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n;
       u += waves) {
    const uint64_t d = dcnt[u], deg = off[u + 1] - off[u];
    if (lane() == 0) {
      mx = d > mx ? d : mx;
      w += (deg - d) * d;
    }
    for (uint32_t p = doff[u] + lane(); p < doff[u + 1]; p += kWave) {
      const uint64_t dv = dcnt[dadj[p]];
      tb += dv < d ? dv : d;
    }
  }
  tb = wave_sum64(tb);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long x = __shfl_xor(mx, o, kWave);
    mx = x > mx ? x : mx;
  }
  w = wave_sum64(w);
  if (lane() == 0) {
    if (mx) atomicMax(&out[0], mx);
    if (w) atomicAdd(&out[1], w);
    if (tb) atomicAdd(&out[2], tb);
  }
}

}  // namespace

hipError_t dag_stats(uint32_t n, const uint32_t *off, const uint32_t *dcnt,
                     const uint32_t *doff, const uint32_t *dadj, unsigned long long *out,
                     hipStream_t s) {
  hipError_t e = hipMemsetAsync(out, 0, 3 * sizeof(unsigned long long), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_dag_stats, dim3(grid_for((size_t)n * kWave, 256, 8192)), dim3(256), 0, s,
                     n, off, dcnt, doff, dadj, out);
  return hipGetLastError();
}

hipError_t dag_count(uint32_t n, const uint32_t *off, const uint32_t *adj,
                     uint32_t *dcnt, hipStream_t s) {
  hipLaunchKernelGGL(k_dag_count, dim3(grid_for((size_t)n * kWave, 256, 65535)),
                     dim3(256), 0, s, n, off, adj, dcnt);
  return hipGetLastError();
}

hipError_t dag_fill(uint32_t n, const uint32_t *off, const uint32_t *adj,
                    const uint32_t *doff, uint32_t *dadj, hipStream_t s) {
  hipLaunchKernelGGL(k_dag_fill, dim3(grid_for((size_t)n * kWave, 256, 65535)),
                     dim3(256), 0, s, n, off, adj, doff, dadj);
  return hipGetLastError();
}

}  // namespace cds
