// (c) Exact max flow of one component's vertex/clique network.
//
// Reference GPU path: createFlowNetwork / preFlow / pushRelabel /
// globalRelabel (cudaCode/src/helpers.cu:1173-1492), double capacities and
// a cooperative grid.  Oracle: createFlownetwork + edmondsKarp
// (cpuCode/c++/src/helpers.cpp:1302-1405), float capacities.
//
// Network (oracle orientation): s->v cap cd_v, v->t cap fl(alpha*k),
// v->clique cap 1, clique->v cap k-1.  Every capacity is multiplied by
// U = 2^sc so fl(alpha*k) is an integer; flows are int64 and exact.
//
// The kernel solves the REVERSED network (source t, sink s) because the
// oracle's answer set is the vertices reachable from s in the residual of
// a maximum flow, i.e. the minimal min-cut side, and in the reversed
// network that set is exactly "can reach the sink" after a maximum
// PREflow, which a backward BFS from the sink (the global relabel)
// computes.  The arcs t->v start saturated (excess fl(alpha*k)*U at every
// vertex) and are never used again.
//
// Representation is implicit: per (clique, member slot) one int64 x = net
// flow vertex->clique in the reversed network, so residual(v->c) =
// (k-1)U - x and residual(c->v) = U + x; per vertex rvs = residual(v->sink).
//
// Schedule: one workgroup; each round has a decision phase (every active
// node picks its lowest residual neighbour from a consistent snapshot of
// heights and residuals: push there if lower, else relabel) and an apply
// phase; the two pushes on an arc pair can never meet in one round.  A
// global relabel (exact BFS distances) runs every kGlobalRelabel rounds and
// whenever no node is active; the solve ends only when that BFS finds no
// node with excess that can reach the sink, which certifies the preflow
// maximum (sink excess = max flow) and yields the min-cut side.
#include "dev.h"
#include "kernels.h"

namespace cds {
namespace {

constexpr int kFlowThreads = 1024;
constexpr uint32_t kGlobalRelabel = 48;

struct Net {
  uint32_t k, V, T, N;
  int INF;
  uint32_t vbase;
  uint64_t sbase;
  long long U;
  const uint32_t *mem;
  const uint64_t *ioff;
  const uint32_t *inc;
  long long *x, *rvs, *ex, *damt;
  int *h;
  long long *dtgt;
  uint32_t *q[2];
};

// Exact BFS from the sink over residual arcs reversed; sets h = distance
// (INF if the sink is unreachable) and returns the number of nodes with
// excess that can reach the sink.
__device__ uint32_t global_relabel(const Net &g, uint32_t *s_qn,
                                   uint32_t *s_act) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t a = tid; a < g.N; a += kFlowThreads) g.h[a] = g.INF;
  if (tid == 0) {
    s_qn[0] = 0;
    s_qn[1] = 0;
    *s_act = 0;
  }
  __syncthreads();
  for (uint32_t v = tid; v < g.V; v += kFlowThreads)
    if (g.rvs[v] > 0) {
      g.h[v] = 1;
      g.q[0][atomicAdd(&s_qn[0], 1u)] = v;
    }
  int cur = 0, d = 1;
  for (;;) {
    __syncthreads();
    const uint32_t ncur = s_qn[cur];
    if (ncur == 0) break;
    uint32_t *qc = g.q[cur], *qn = g.q[1 - cur];
    for (uint32_t i = tid; i < ncur; i += kFlowThreads) {
      const uint32_t a = qc[i];
      if (a < g.V) {
        // arcs into vertex a: clique -> a, residual U + x
        for (uint64_t p = g.ioff[a]; p < g.ioff[a + 1]; ++p) {
          const uint64_t s = g.inc[p] - g.sbase;
          if (g.U + g.x[s] <= 0) continue;
          const uint32_t c = g.V + (uint32_t)(s / g.k);
          if (atomicCAS(&g.h[c], g.INF, d + 1) == g.INF)
            qn[atomicAdd(&s_qn[1 - cur], 1u)] = c;
        }
      } else {
        // arcs into clique a: member -> a, residual (k-1)U - x
        const uint64_t base = (uint64_t)(a - g.V) * g.k;
        for (uint32_t j = 0; j < g.k; ++j) {
          if ((long long)(g.k - 1) * g.U - g.x[base + j] <= 0) continue;
          const uint32_t v = g.mem[base + j] - g.vbase;
          if (atomicCAS(&g.h[v], g.INF, d + 1) == g.INF)
            qn[atomicAdd(&s_qn[1 - cur], 1u)] = v;
        }
      }
    }
    __syncthreads();
    if (tid == 0) s_qn[cur] = 0;
    cur = 1 - cur;
    ++d;
  }
  uint32_t act = 0;
  for (uint32_t a = tid; a < g.N; a += kFlowThreads)
    act += (g.ex[a] > 0 && g.h[a] < g.INF);
  act = wave_sum(act);
  if (lane() == 0 && act) atomicAdd(s_act, act);
  __syncthreads();
  return *s_act;
}

__global__ __launch_bounds__(kFlowThreads) void k_max_flow(
    FlowProblem P, FlowState S, long long cap_t, long long U,
    uint32_t max_rounds) {
  __shared__ uint32_t s_qn[2], s_act, s_any;
  Net g;
  g.k = P.k;
  g.V = P.V;
  g.T = P.T;
  g.N = P.V + P.T;
  g.INF = (int)g.N + 1;
  g.vbase = P.vbase;
  g.sbase = P.sbase;
  g.U = U;
  g.mem = P.mem;
  g.ioff = P.ioff;
  g.inc = P.inc;
  g.x = S.x;
  g.rvs = S.rvs;
  g.ex = S.excess;
  g.h = S.h;
  g.dtgt = S.dtgt;
  g.damt = S.damt;
  g.q[0] = S.queue0;
  g.q[1] = S.queue1;
  const uint32_t tid = threadIdx.x;
  const uint64_t slots = (uint64_t)g.T * g.k;

  for (uint64_t s = tid; s < slots; s += kFlowThreads) g.x[s] = 0;
  for (uint32_t v = tid; v < g.V; v += kFlowThreads) {
    g.rvs[v] = (long long)P.cd[v] * U;
    g.ex[v] = cap_t;
  }
  for (uint32_t c = tid; c < g.T; c += kFlowThreads) g.ex[g.V + c] = 0;
  if (tid == 0) {
    S.misc[0] = 0;  // flow into the sink
    S.misc[1] = 0;  // rounds
    S.misc[2] = 0;  // 1 = round limit hit
    S.misc[3] = 0;  // global relabels
  }
  __syncthreads();
  uint32_t active = global_relabel(g, s_qn, &s_act);
  uint32_t rounds = 0, relabels = 1;
  unsigned long long sink_flow = 0;
  bool limit = false;
  while (active) {
    if (rounds >= max_rounds) {
      limit = true;
      break;
    }
    // ---- decision phase (reads a consistent snapshot) ----
    if (tid == 0) s_any = 0;
    __syncthreads();
    uint32_t mine = 0;
    for (uint32_t a = tid; a < g.N; a += kFlowThreads) {
      long long dt = -1, amt = 0;
      const long long e = g.ex[a];
      const int ha = g.h[a];
      if (e > 0 && ha < g.INF) {
        ++mine;
        int hmin = g.INF;
        long long best = -1, bres = 0;
        if (a < g.V) {
          const long long r = g.rvs[a];
          if (r > 0) {
            hmin = 0;  // the sink
            best = 0;
            bres = r;
          } else {
            for (uint64_t p = g.ioff[a]; p < g.ioff[a + 1]; ++p) {
              const uint64_t s = g.inc[p] - g.sbase;
              const long long res = (long long)(g.k - 1) * U - g.x[s];
              if (res <= 0) continue;
              const int hc = g.h[g.V + (uint32_t)(s / g.k)];
              if (hc < hmin) {
                hmin = hc;
                best = (long long)s + 1;
                bres = res;
              }
            }
          }
        } else {
          const uint64_t base = (uint64_t)(a - g.V) * g.k;
          for (uint32_t j = 0; j < g.k; ++j) {
            const long long res = U + g.x[base + j];
            if (res <= 0) continue;
            const int hv = g.h[g.mem[base + j] - g.vbase];
            if (hv < hmin) {
              hmin = hv;
              best = (long long)(base + j) + 1;
              bres = res;
            }
          }
        }
        if (best >= 0 && ha > hmin) {
          dt = best;
          amt = e < bres ? e : bres;
        } else {
          dt = -2;
          amt = best >= 0 ? (hmin + 1 < g.INF ? hmin + 1 : g.INF) : g.INF;
        }
      }
      g.dtgt[a] = dt;
      g.damt[a] = amt;
    }
    mine = wave_sum(mine);
    if (lane() == 0 && mine) atomicAdd(&s_any, mine);
    __syncthreads();
    if (s_any == 0) {
      active = global_relabel(g, s_qn, &s_act);
      ++relabels;
      continue;
    }
    // ---- apply phase ----
    for (uint32_t a = tid; a < g.N; a += kFlowThreads) {
      const long long dt = g.dtgt[a];
      if (dt == -1) continue;
      const long long amt = g.damt[a];
      if (dt == -2) {
        g.h[a] = (int)amt;
        continue;
      }
      atomicAdd((unsigned long long *)&g.ex[a], (unsigned long long)(-amt));
      if (a < g.V) {
        if (dt == 0) {
          g.rvs[a] -= amt;
          sink_flow += (unsigned long long)amt;
        } else {
          const uint64_t s = (uint64_t)(dt - 1);
          g.x[s] += amt;
          atomicAdd((unsigned long long *)&g.ex[g.V + (uint32_t)(s / g.k)],
                    (unsigned long long)amt);
        }
      } else {
        const uint64_t s = (uint64_t)(dt - 1);
        g.x[s] -= amt;
        atomicAdd((unsigned long long *)&g.ex[g.mem[s] - g.vbase],
                  (unsigned long long)amt);
      }
    }
    __syncthreads();
    ++rounds;
    if (rounds % kGlobalRelabel == 0) {
      active = global_relabel(g, s_qn, &s_act);
      ++relabels;
    }
  }
  // A final exact BFS leaves h = distance to the sink; the min-cut side
  // (reachable from s in the oracle's residual) is h < INF.
  if (!limit) global_relabel(g, s_qn, &s_act);
  sink_flow = wave_sum64(sink_flow);
  if (lane() == 0 && sink_flow) atomicAdd(&S.misc[0], sink_flow);
  for (uint32_t v = tid; v < g.V; v += kFlowThreads)
    S.reach[v] = g.h[v] < g.INF ? 1 : 0;
  if (tid == 0) {
    S.misc[1] = rounds;
    S.misc[2] = limit ? 1 : 0;
    S.misc[3] = relabels;
  }
}

__global__ void k_count_in_reach(uint32_t T, uint32_t k,
                                 const uint32_t *__restrict__ mem,
                                 uint32_t vbase, const int *__restrict__ reach,
                                 unsigned long long *out) {
  unsigned long long c = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T;
       t += gridDim.x * blockDim.x) {
    bool all = true;
    for (uint32_t j = 0; j < k; ++j)
      all &= reach[mem[(uint64_t)t * k + j] - vbase] != 0;
    c += all;
  }
  c = wave_sum64(c);
  if (lane() == 0 && c) atomicAdd(out, c);
}

__global__ void k_count_in_mask(uint64_t T, uint32_t k,
                                const uint32_t *__restrict__ cl,
                                const uint8_t *__restrict__ mask,
                                unsigned long long *out) {
  unsigned long long c = 0;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T;
       t += (uint64_t)gridDim.x * blockDim.x) {
    bool all = true;
    for (uint32_t j = 0; j < k; ++j) all &= mask[cl[t * k + j]] != 0;
    c += all;
  }
  c = wave_sum64(c);
  if (lane() == 0 && c) atomicAdd(out, c);
}

}  // namespace

hipError_t count_cliques_in_reach(uint32_t T, uint32_t k, const uint32_t *mem,
                                  uint32_t vbase, const int *reach,
                                  unsigned long long *out, hipStream_t s) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(*out), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_count_in_reach, dim3(grid_for(T, 256, 8192)), dim3(256),
                     0, s, T, k, mem, vbase, reach, out);
  return hipGetLastError();
}

hipError_t count_cliques_in_mask(uint64_t T, uint32_t k, const uint32_t *cl,
                                 const uint8_t *mask, unsigned long long *out,
                                 hipStream_t s) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(*out), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_count_in_mask, dim3(grid_for(T, 256, 8192)), dim3(256),
                     0, s, T, k, cl, mask, out);
  return hipGetLastError();
}

hipError_t max_flow(const FlowProblem &P, const FlowState &S, long long cap_t,
                    long long unit, uint32_t max_rounds, hipStream_t s) {
  hipLaunchKernelGGL(k_max_flow, dim3(1), dim3(kFlowThreads), 0, s, P, S, cap_t,
                     unit, max_rounds);
  return hipGetLastError();
}

}  // namespace cds
