// ORACLE — test infrastructure only.  CPU restatement of the reference's
// serial k-clique densest-subgraph pipeline (wajidmanzoor/CDS, cpuCode/c++).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load or run this; the engine never links it.
//
// Every stage follows the reference function cited beside it, including its
// quirks (tie rules, float bound arithmetic, cumulative upper bound, float
// density).  Two implementations exist for the two expensive stages:
//
//   peel  PEEL_FAITHFUL  O(n^2) argmin + neighbourhood re-enumeration per
//                        removed vertex, as cpuCode/c++/src/helpers.cpp:16-228.
//         PEEL_FAST      the same removal sequence from a (degree, id)
//                        ordered set over a vertex->clique incidence.
//   flow  FLOW_FLOAT_EK  float Edmonds-Karp over the string-keyed clique map
//                        (helpers.cpp:790-956, :1256-1420): the reference's
//                        arithmetic when built with the same flags.
//         FLOW_EXACT     int64 Dinic on the same network, capacities scaled by
//                        2^s so fl(alpha*k) is an integer; the decision
//                        |F - target| < 1e-2f is evaluated exactly.
//
// Parity pins (what checks each function; see tests/test_oracle.py):
//   listing      brute force on small graphs, K_n closed forms, the
//                reference's logged 10-vertex example, the reference build.
//   peel         brute-force sequential min removal, FAITHFUL == FAST,
//                reference build.
//   core/prune/CC brute force on small graphs, reference build.
//   flow         FLOW_EXACT vs brute-force max_A k T(A) - c|A| on tiny
//                components; FLOAT_EK vs the reference build.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <queue>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

typedef unsigned int ui;

namespace {

enum { PEEL_FAITHFUL = 0, PEEL_FAST = 1 };
enum { FLOW_FLOAT_EK = 0, FLOW_EXACT = 1 };

// --------------------------------------------------------------------------
// Degeneracy order, cpuCode/c++/src/helpers.cpp:230-315 (getlistingOrder +
// coreDecompose).  Quirk kept: the reverse-order slot is written inside the
// neighbour loop, so isolated vertices never get a slot and the unwritten
// slots (value 0) overwrite order[0].  It only changes the orientation.
// --------------------------------------------------------------------------
static std::vector<ui> listing_order(const std::vector<std::vector<ui>> &adj) {
  const ui n = adj.size();
  std::vector<ui> deg(n), core(n), rev(n, 0);
  ui maxd = 0;
  for (ui i = 0; i < n; ++i) {
    deg[i] = adj[i].size();
    core[i] = deg[i];
    maxd = std::max(maxd, deg[i]);
  }
  std::vector<ui> bins(maxd + 1, 0), pos(n), sorted(n);
  for (ui i = 0; i < n; ++i) bins[deg[i]]++;
  ui start = 0;
  for (ui d = 0; d <= maxd; ++d) {
    ui c = bins[d];
    bins[d] = start;
    start += c;
  }
  for (ui i = 0; i < n; ++i) {
    pos[i] = bins[deg[i]];
    sorted[pos[i]] = i;
    bins[deg[i]]++;
  }
  for (ui d = maxd; d > 0; --d) bins[d] = bins[d - 1];
  bins[0] = 0;
  for (ui i = 0; i < n; ++i) {
    const ui v = sorted[i];
    for (ui u : adj[v]) {
      if (core[u] > core[v]) {
        const ui du = core[u], pu = pos[u], pw = bins[du], w = sorted[pw];
        if (u != w) {
          pos[u] = pw;
          pos[w] = pu;
          sorted[pu] = w;
          sorted[pw] = u;
        }
        bins[du]++;
        core[u]--;
      }
      rev[n - i - 1] = v;  // inside the loop, as the reference does
    }
  }
  std::vector<ui> order(n, 0);
  for (ui i = 0; i < n; ++i) order[rev[i]] = i + 1;
  return order;
}

// Arcs toward larger order, adjacency order kept (helpers.cpp:349-371).
static std::vector<std::vector<ui>> orient(const std::vector<std::vector<ui>> &adj,
                                          const std::vector<ui> &order) {
  std::vector<std::vector<ui>> dag(adj.size());
  for (ui i = 0; i < adj.size(); ++i)
    for (ui u : adj[i])
      if (order[u] > order[i]) dag[i].push_back(u);
  return dag;
}

// Recursive listing with a label array and in-place partitioning of the DAG
// lists: the shared skeleton of listCliques (helpers.cpp:373-511),
// listCliqueContainsVertex (:582-708) and listCliqueRecord (:832-956).
// emit(partial, u, v) is called once per k-clique partial + {u, v}, in the
// reference's enumeration order.
struct Lister {
  std::vector<std::vector<ui>> dag;
  std::vector<ui> label, vcount;
  std::vector<ui> partial;

  template <class Emit>
  void run(ui k, const std::vector<ui> &cand, Emit &emit) {
    if (k == 2) {
      for (ui u : cand)
        for (ui j = 0; j < vcount[u]; ++j) emit(partial, u, dag[u][j]);
      return;
    }
    for (ui u : cand) {
      std::vector<ui> next;
      for (ui w : dag[u])
        if (label[w] == k) {
          label[w] = k - 1;
          next.push_back(w);
        }
      for (ui w : next) {
        auto &L = dag[w];
        int idx = 0;
        for (int mpos = (int)L.size() - 1; mpos > idx; --mpos) {
          if (label[L[mpos]] == k - 1) {
            while (idx < mpos && label[L[idx]] == k - 1) idx++;
            if (label[L[idx]] != k - 1) std::swap(L[mpos], L[idx]);
          }
        }
        if (!L.empty() && label[L[idx]] == k - 1) idx++;
        vcount[w] = idx;
      }
      partial.push_back(u);
      run(k - 1, next, emit);
      partial.pop_back();
      for (ui w : next) label[w] = k;
    }
  }

  template <class Emit>
  void list_all(const std::vector<std::vector<ui>> &adj, ui k, Emit &&emit) {
    const ui n = adj.size();
    dag = orient(adj, listing_order(adj));
    label.assign(n, k);
    vcount.resize(n);
    for (ui i = 0; i < n; ++i) vcount[i] = dag[i].size();
    partial.clear();
    std::vector<ui> cand(n);
    std::iota(cand.begin(), cand.end(), 0u);
    run(k, cand, emit);
  }
};

// --------------------------------------------------------------------------
// Result container
// --------------------------------------------------------------------------
struct Component {
  int size = 0;
  std::vector<ui> reverse_map;  // local -> core-local id
  // FLOW_FLOAT_EK keeps the reference's string-keyed map (its iteration order
  // fixes the flow network's node order); FLOW_EXACT keeps a plain list.
  std::unordered_map<std::string, std::vector<int>> clique_map;
  std::vector<int> cliques;  // FLOW_EXACT: k members (local ids) per clique
  std::vector<long> clique_degree;
  long total_cliques = 0;
  double density = 0;
};

struct Result {
  ui k = 0, n = 0;
  unsigned long long m = 0;
  long total_cliques = 0;  // graph->totalCliques (long)
  std::vector<long> clique_degree;
  std::vector<long> removal_vertex, removal_value;  // rows 1..n
  std::vector<double> row_density;                  // rows 0..n
  std::vector<double> row_remaining;                // results[i][3]
  double max_density = 0;
  int lower_bound = 0;  // k' = ceil(max row density)
  long kmax = 0;
  int del_vertex_index = 0;
  int del_clique_count = 0;
  double core_density = 0;
  ui core_size = 0;
  long core_edges = 0;
  std::vector<ui> core_reverse_map;
  long core_cliques = 0;
  int pruned_edges = 0;
  std::vector<Component> comps;
  double paper_lb = 0;
  float ub1_val = 0, ub2_val = 0;
  double density = 0;  // densestSubgraph.density
  ui size = 0;
  std::vector<ui> vertices;
  long set_cliques = 0;
  std::vector<float> tr_lb, tr_ub, tr_alpha;
  std::vector<int> tr_comp, tr_equal;
  std::vector<double> tr_margin;
  int flow_networks = 0;
  int quirk_flags = 0;  // bit0 sentinel regime, bit1 float-inexact caps
  double t_list = 0, t_peel = 0, t_flow = 0, t_total = 0;
};

// --------------------------------------------------------------------------
// (a) full-graph listing: T and per-vertex clique degree
// (cliqueEnumerationFast, helpers.cpp:317-347 + listCliques :373-511).
// --------------------------------------------------------------------------
static void list_full(const std::vector<std::vector<ui>> &adj, ui k, Result &R,
                      std::vector<ui> *members) {
  const ui n = adj.size();
  R.clique_degree.assign(n, 0);
  R.total_cliques = 0;
  Lister L;
  L.list_all(adj, k, [&](const std::vector<ui> &p, ui a, ui b) {
    R.total_cliques++;
    for (ui x : p) R.clique_degree[x]++;
    R.clique_degree[a]++;
    R.clique_degree[b]++;
    if (members) {
      members->insert(members->end(), p.begin(), p.end());
      members->push_back(a);
      members->push_back(b);
    }
  });
}

// --------------------------------------------------------------------------
// (b) clique-core peel, helpers.cpp:16-150.  results rows = (removed vertex,
// its current clique degree, density after removal, cliques left).
// --------------------------------------------------------------------------

// Cliques containing v among unmarked vertices, counted per member, by
// listing the induced subgraph on v's unmarked closed neighbourhood with v
// as local vertex 0 (get2Dneighborhood, cliqueEnumerationSubgraph,
// listCliqueContainsVertex: helpers.cpp:152-228, :513-708).
static void neighbourhood_counts(const std::vector<std::vector<ui>> &adj, ui k,
                                 ui v, const std::vector<int> &mark,
                                 std::vector<int> &tag, std::vector<int> &local,
                                 std::vector<std::pair<ui, long>> &out) {
  out.clear();
  std::vector<ui> members{v};
  tag[v] = 1;
  for (ui u : adj[v])
    if (!mark[u] && !tag[u]) {
      tag[u] = 2;
      members.push_back(u);
    }
  const ui cnt = members.size();
  for (ui i = 0; i < cnt; ++i) local[members[i]] = i;
  std::vector<std::vector<ui>> sub(cnt);
  for (ui i = 0; i < cnt; ++i)
    for (ui u : adj[members[i]])
      if (tag[u] && u != members[i]) sub[i].push_back(local[u]);
  for (ui u : members) tag[u] = 0;
  std::vector<long> deg(cnt, 0);
  Lister L;
  L.list_all(sub, k, [&](const std::vector<ui> &p, ui a, ui b) {
    bool has = (a == 0 || b == 0);
    for (ui x : p) has |= (x == 0);
    if (!has) return;
    for (ui x : p) deg[x]++;
    deg[a]++;
    deg[b]++;
  });
  for (ui i = 0; i < cnt; ++i)
    if (deg[i] > 0) out.emplace_back(members[i], deg[i]);
}

static void peel(const std::vector<std::vector<ui>> &adj, ui k,
                 const std::vector<ui> &members, int mode,
                 Result &R) {
  const ui n = adj.size();
  std::vector<long> cur(R.clique_degree);
  // `int totalCliques` sums the long degrees, then divides by k through a
  // double (helpers.cpp:38-48).
  int total = 0;
  for (ui i = 0; i < n; ++i) total += R.clique_degree[i];
  total = total / static_cast<double>(k);
  R.row_density.assign(n + 1, 0.0);
  R.row_remaining.assign(n + 1, 0.0);
  R.removal_vertex.assign(n + 1, 0);
  R.removal_value.assign(n + 1, 0);
  R.row_density[0] = total / static_cast<double>(n);
  R.row_remaining[0] = total;
  std::vector<int> mark(n, 0);

  auto finish_row = [&](ui i) {
    R.row_remaining[i] = total;
    R.row_density[i] = (n - i > 0) ? total / static_cast<double>(n - i) : 0.0;
  };

  if (mode == PEEL_FAITHFUL) {
    std::vector<int> tag(n, 0), local(n, 0);
    std::vector<std::pair<ui, long>> hits;
    for (ui i = 1; i <= n; ++i) {
      ui index = 0;
      long best = 0xFFFFFF;
      for (ui j = 0; j < n; ++j)
        if (best > cur[j] && mark[j] == 0) {
          best = cur[j];
          index = j;
        }
      if (mark[index]) R.quirk_flags |= 1;
      R.removal_vertex[i] = index;
      R.removal_value[i] = cur[index];
      if (R.clique_degree[index] > 0) {
        neighbourhood_counts(adj, k, index, mark, tag, local, hits);
        long del = 0;
        for (auto &h : hits) {
          del += h.second;
          cur[h.first] -= h.second;
        }
        total -= del / (long)k;
      } else {
        cur[index] = 0;
      }
      finish_row(i);
      mark[index] = 1;
    }
    return;
  }

  // PEEL_FAST: the same sequence.  argmin over unmarked vertices of
  // (current clique degree, id) comes from a tournament tree; the cliques
  // killed by removing v are the alive cliques on v's incidence list (a CSR
  // whose row lengths are the initial clique degrees).  Members and
  // incidence are flat arrays: k + k 32-bit words per clique.
  const uint64_t T = members.size() / k;
  std::vector<uint64_t> ioff(n + 1, 0);
  for (ui v = 0; v < n; ++v) ioff[v + 1] = ioff[v] + (uint64_t)R.clique_degree[v];
  std::vector<ui> inc(ioff[n]);
  {
    std::vector<uint64_t> fill(ioff.begin(), ioff.end() - 1);
    for (uint64_t c = 0; c < T; ++c)
      for (ui j = 0; j < k; ++j) inc[fill[members[c * k + j]]++] = (ui)c;
  }
  std::vector<char> alive(T, 1);
  // Tournament tree over (degree, id) keys; a removed vertex holds the
  // largest key.
  typedef unsigned __int128 key_t;
  const key_t kGone = ~(key_t)0;
  size_t leaves = 1;
  while (leaves < n) leaves <<= 1;
  std::vector<key_t> tree(2 * leaves, kGone);
  auto key_of = [&](ui v) { return ((key_t)(uint64_t)cur[v] << 32) | v; };
  for (ui v = 0; v < n; ++v) tree[leaves + v] = key_of(v);
  for (size_t i = leaves - 1; i >= 1; --i) tree[i] = std::min(tree[2 * i], tree[2 * i + 1]);
  auto set_leaf = [&](ui v, key_t key) {
    size_t i = leaves + v;
    tree[i] = key;
    for (i >>= 1; i >= 1; i >>= 1) {
      const key_t m = std::min(tree[2 * i], tree[2 * i + 1]);
      if (tree[i] == m) break;
      tree[i] = m;
    }
  };
  for (ui i = 1; i <= n; ++i) {
    const key_t top = tree[1];
    const ui index = (ui)(uint64_t)top;
    // Above 0xFFFFFF the reference selects index 0 even if it is already
    // removed (helpers.cpp:93-100); that regime is flagged, not emulated.
    if ((long)(uint64_t)(top >> 32) > 0xFFFFFF) R.quirk_flags |= 1;
    set_leaf(index, kGone);
    R.removal_vertex[i] = index;
    R.removal_value[i] = cur[index];
    long del = 0;
    for (uint64_t e = ioff[index]; e < ioff[index + 1]; ++e) {
      const ui c = inc[e];
      if (!alive[c]) continue;
      alive[c] = 0;
      ++del;
      for (ui j = 0; j < k; ++j) {
        const ui x = members[(uint64_t)c * k + j];
        if (x == index || mark[x]) continue;
        cur[x]--;
        set_leaf(x, key_of(x));
      }
    }
    cur[index] = 0;
    total -= del;
    finish_row(i);
    mark[index] = 1;
  }
}

// --------------------------------------------------------------------------
// Densest core, helpers.cpp:710-788.
// --------------------------------------------------------------------------
static void locate_core(const std::vector<std::vector<ui>> &adj, Result &R,
                        std::vector<std::vector<ui>> &core_graph) {
  const ui n = adj.size();
  double mx = R.row_density[0];
  R.kmax = 0;
  for (ui i = 1; i < n; ++i) {
    if (mx < R.row_density[i]) mx = R.row_density[i];
    if (R.kmax < R.removal_value[i]) R.kmax = R.removal_value[i];
  }
  R.max_density = mx;
  const int lb = (int)std::ceil(mx);
  R.lower_bound = lb;
  int index = 1;
  std::vector<int> del(n, 0);
  for (; index < (int)n; ++index) {
    if (R.removal_value[index] >= lb) break;
    del[R.removal_vertex[index]] = -1;
  }
  int next = 0;
  for (ui i = 0; i < n; ++i)
    if (del[i] == 0) del[i] = next++;
  core_graph.assign(next, {});
  R.core_reverse_map.assign(next, 0);
  for (ui i = 0; i < n; ++i) {
    if (del[i] == -1) continue;
    R.core_reverse_map[del[i]] = i;
    for (ui u : adj[i])
      if (del[u] != -1) core_graph[del[i]].push_back(del[u]);
  }
  R.core_size = next;
  long e2 = 0;
  for (auto &l : core_graph) e2 += l.size();
  R.core_edges = e2 / 2;
  R.del_vertex_index = index - 1;
  R.del_clique_count = (int)(R.row_remaining[0] - R.row_remaining[index - 1]);
  R.core_density = R.row_density[index - 1];
}

// --------------------------------------------------------------------------
// Re-listing inside the core (helpers.cpp:790-956), pruning (:958-993) and
// connected components (:995-1116).
// --------------------------------------------------------------------------
static void components(const std::vector<std::vector<ui>> &core_graph, ui k,
                       int flow_mode, Result &R) {
  const ui cn = core_graph.size();
  // Re-list with string keys: the map's iteration order is what the
  // reference's flow network is built from.  FLOW_EXACT keeps the cliques
  // as a flat list of k members each (order-independent from here on).
  std::unordered_map<std::string, std::vector<int>> clique_data;
  std::vector<int> clique_list;
  {
    Lister L;
    L.list_all(core_graph, k, [&](const std::vector<ui> &p, ui a, ui b) {
      if (flow_mode == FLOW_FLOAT_EK) {
        std::vector<int> arr(k + 1);
        for (ui x = 0; x < p.size(); ++x) arr[x] = p[x];
        arr[k - 2] = a;
        arr[k - 1] = b;
        arr[k] = 1;
        std::string key;
        for (ui x : p) key += std::to_string(x) + " ";
        key += std::to_string(a) + " " + std::to_string(b);
        clique_data[key] = arr;
      } else {
        for (ui x : p) clique_list.push_back((int)x);
        clique_list.push_back((int)a);
        clique_list.push_back((int)b);
      }
    });
  }
  const size_t nflat = clique_list.size() / k;
  if (flow_mode == FLOW_FLOAT_EK) {
    R.core_cliques = clique_data.size();
  } else {
    R.core_cliques = nflat;
  }
  // Prune: keep core edges covered by at least one core clique.  Covered
  // pairs are marked in a bit matrix when the core is small enough, else in
  // sorted per-vertex lists; both hold the same set.
  const bool dense = cn <= 32768;
  std::vector<uint64_t> bits(dense ? ((size_t)cn * cn + 63) / 64 : 0);
  std::vector<std::vector<ui>> valid(dense ? 0 : cn);
  auto mark_pair = [&](int a, int b) {
    if (dense) {
      const size_t i = (size_t)a * cn + b;
      bits[i >> 6] |= 1ull << (i & 63);
    } else {
      valid[a].push_back(b);
    }
  };
  auto mark_clique = [&](const int *arr) {
    for (ui i = 0; i < k; ++i)
      for (ui j = 0; j < k; ++j)
        if (i != j) mark_pair(arr[i], arr[j]);
  };
  if (flow_mode == FLOW_FLOAT_EK)
    for (auto &e : clique_data) mark_clique(e.second.data());
  else
    for (size_t c = 0; c < nflat; ++c) mark_clique(&clique_list[c * k]);
  for (auto &v : valid) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  auto covered = [&](ui a, ui b) {
    if (dense) {
      const size_t i = (size_t)a * cn + b;
      return (bits[i >> 6] >> (i & 63)) & 1;
    }
    return (uint64_t)std::binary_search(valid[a].begin(), valid[a].end(), b);
  };
  std::vector<std::vector<ui>> pruned(cn);
  int total_edges = 0;
  for (ui i = 0; i < cn; ++i)
    for (ui u : core_graph[i])
      if (covered(i, u)) {
        pruned[i].push_back(u);
        total_edges++;
      }
  R.pruned_edges = total_edges / 2;
  // BFS components from the smallest unvisited id; local ids ascending.
  std::vector<ui> status(cn, 0);
  int ncomp = 0;
  for (ui i = 0; i < cn; ++i) {
    if (status[i]) continue;
    ++ncomp;
    std::queue<ui> q;
    q.push(i);
    status[i] = ncomp;
    while (!q.empty()) {
      ui x = q.front();
      q.pop();
      for (ui y : pruned[x])
        if (!status[y]) {
          status[y] = ncomp;
          q.push(y);
        }
    }
  }
  std::vector<ui> old_to_new(cn), sizes(ncomp + 1, 0);
  for (ui i = 0; i < cn; ++i) old_to_new[i] = sizes[status[i]]++;
  R.comps.assign(ncomp, Component());
  for (int c = 1; c <= ncomp; ++c) {
    R.comps[c - 1].size = sizes[c];
    R.comps[c - 1].reverse_map.resize(sizes[c]);
    R.comps[c - 1].clique_degree.assign(sizes[c], 0);
  }
  for (ui i = 0; i < cn; ++i)
    R.comps[status[i] - 1].reverse_map[old_to_new[i]] = i;
  if (flow_mode == FLOW_FLOAT_EK) {
    if (ncomp == 1) {
      // One component: conComp.cliqueData is a copy of the whole map
      // (helpers.cpp:1026-1044).
      R.comps[0].clique_map = clique_data;
    } else {
      for (auto &e : clique_data) {
        std::vector<int> arr(e.second);
        const int cid = status[arr[0]] - 1;
        for (ui i = 0; i < k; ++i) arr[i] = old_to_new[arr[i]];
        R.comps[cid].clique_map[e.first] = arr;
      }
    }
    for (auto &C : R.comps) {
      for (auto &e : C.clique_map) {
        for (ui i = 0; i < k; ++i) C.clique_degree[e.second[i]] += e.second[k];
        C.total_cliques += e.second[k];
      }
    }
  } else {
    for (size_t c = 0; c < nflat; ++c) {
      const int *arr = &clique_list[c * k];
      Component &C = R.comps[status[arr[0]] - 1];
      for (ui i = 0; i < k; ++i) {
        const int x = (int)old_to_new[arr[i]];
        C.cliques.push_back(x);
        C.clique_degree[x] += 1;
      }
      C.total_cliques += 1;
    }
  }
  for (auto &C : R.comps)
    C.density = (double)C.total_cliques / ((double)C.size);
}

// --------------------------------------------------------------------------
// (c) flow decisions.  Network (helpers.cpp:1302-1341): s->v cap cd_v,
// v->t cap fl(alpha*k), v->clique cap 1, clique->v cap k-1.
// --------------------------------------------------------------------------
struct FEdge {
  int to, rev;
  float cap, max;
};

// Float Edmonds-Karp exactly as helpers.cpp:1256-1300, :1343-1420.
static void exact_float(std::vector<int> &res, Component &C, ui k, float lb,
                        float ub, int cid, Result &R) {
  float alpha = (lb + ub) * 0.5f;
  float bias = 1.0f / (C.size * (C.size - 1));
  if (bias < 1e-2f) bias = 1e-2f;
  const int total_nodes = C.size + C.total_cliques + 2;
  std::vector<std::vector<FEdge>> G(total_nodes);
  std::vector<int> sink_idx(total_nodes, -1);
  const int source = C.size + C.total_cliques, sink = source + 1;
  auto add = [&](int u, int v, float cap) {
    FEdge a{v, (int)G[v].size(), cap, cap};
    FEdge b{u, (int)G[u].size(), 0.0f, 0.0f};
    G[u].push_back(a);
    G[v].push_back(b);
  };
  int idx = C.size;
  for (auto &e : C.clique_map) {
    auto &clq = e.second;
    float count = clq[k];
    float w = count * (k - 1);
    for (ui i = 0; i < k; ++i) {
      int v = clq[i];
      add(idx, v, w);
      add(v, idx, count);
    }
    idx++;
  }
  for (int v = 0; v < C.size; ++v) {
    if (C.clique_degree[v] >= (1l << 24)) R.quirk_flags |= 2;
    add(source, v, (float)C.clique_degree[v]);
    add(v, sink, alpha * k);
    sink_idx[v] = G[v].size() - 1;
  }
  std::vector<std::pair<int, int>> parent(total_nodes);
  std::vector<int> q;
  auto augment = [&]() -> float {
    std::fill(parent.begin(), parent.end(), std::make_pair(-1, -1));
    q.clear();
    q.push_back(source);
    parent[source] = {source, -1};
    for (int qi = 0; qi < (int)q.size(); qi++) {
      int u = q[qi];
      if (u == sink) break;
      for (int ei = 0; ei < (int)G[u].size(); ei++) {
        FEdge &e = G[u][ei];
        if (parent[e.to].first == -1 && e.cap > 0) {
          parent[e.to] = {u, ei};
          q.push_back(e.to);
        }
      }
    }
    if (parent[sink].first == -1) return -1;
    float f = FLT_MAX;
    for (int v = sink; v != source;) {
      auto [u, ei] = parent[v];
      f = std::min(f, G[u][ei].cap);
      v = u;
    }
    return f;
  };
  res.assign(C.size, 1);
  float target = C.total_cliques * k;
  while ((ub - lb) > bias) {
    float sum = 0.0f, pushed;
    while ((pushed = augment()) != -1) {
      for (int v = sink; v != source;) {
        auto [u, ei] = parent[v];
        FEdge &e = G[u][ei];
        FEdge &rv = G[v][e.rev];
        e.cap -= pushed;
        rv.cap += pushed;
        v = u;
      }
      sum += pushed;
    }
    R.tr_comp.push_back(cid);
    R.tr_lb.push_back(lb);
    R.tr_ub.push_back(ub);
    R.tr_alpha.push_back(alpha);
    R.tr_margin.push_back((double)target - (double)sum);
    const bool eq = std::fabs(sum - target) < 1e-2f;
    R.tr_equal.push_back(eq);
    if (eq) {
      ub = alpha;
    } else {
      lb = alpha;
      for (int i = 0; i < C.size; i++) res[i] = parent[i].first;
    }
    alpha = (lb + ub) * 0.5f;
    for (auto &u : G)
      for (auto &e : u) e.cap = e.max;
    for (int v = 0; v < C.size; v++) {
      int si = sink_idx[v];
      G[v][si].cap = G[v][si].max = alpha * k;
    }
  }
}

// Exact max flow (Dinic, int64) on the same network with every capacity
// scaled by 2^s.  maxflow returns F * 2^s and leaves `level` >= 0 exactly on
// the nodes reachable from the source in the final residual graph (the
// minimal min-cut side, the same for every maximum flow).
//
// Arcs live in a CSR built from an arc list; each arc stores its reverse.
// A pair of opposite arcs u->v (cap a) and v->u (cap b) is one arc pair
// with residuals a and b: the same maximum flow value and the same residual
// reachability as two separate arcs with zero-capacity reverses (u->v has
// residual capacity unless it carries a net flow of a, in both forms).
struct Dinic {
  struct Arc {
    int u, v;
    int64_t a, b;
  };
  int n = 0;
  std::vector<Arc> list;
  std::vector<uint64_t> head;  // n + 1
  std::vector<int> to;
  std::vector<uint64_t> rev;
  std::vector<int64_t> cap, cap0;
  std::vector<int> level;
  std::vector<uint64_t> it;
  std::vector<int> queue;
  explicit Dinic(int nodes) : n(nodes) {}
  void add(int u, int v, int64_t a, int64_t b = 0) { list.push_back({u, v, a, b}); }
  void build() {
    head.assign(n + 1, 0);
    for (auto &e : list) {
      head[e.u + 1]++;
      head[e.v + 1]++;
    }
    for (int i = 0; i < n; ++i) head[i + 1] += head[i];
    const uint64_t m = head[n];
    to.resize(m);
    rev.resize(m);
    cap0.resize(m);
    std::vector<uint64_t> pos(head.begin(), head.end() - 1);
    for (auto &e : list) {
      const uint64_t i = pos[e.u]++, j = pos[e.v]++;
      to[i] = e.v;
      to[j] = e.u;
      rev[i] = j;
      rev[j] = i;
      cap0[i] = e.a;
      cap0[j] = e.b;
    }
    list.clear();
    list.shrink_to_fit();
    cap = cap0;
    level.assign(n, -1);
    it.assign(n, 0);
  }
  void reset() { std::copy(cap0.begin(), cap0.end(), cap.begin()); }
  bool bfs(int s, int t) {
    std::fill(level.begin(), level.end(), -1);
    queue.clear();
    level[s] = 0;
    queue.push_back(s);
    for (size_t qi = 0; qi < queue.size(); ++qi) {
      const int u = queue[qi];
      for (uint64_t e = head[u]; e < head[u + 1]; ++e)
        if (cap[e] > 0 && level[to[e]] < 0) {
          level[to[e]] = level[u] + 1;
          queue.push_back(to[e]);
        }
    }
    return level[t] >= 0;
  }
  int64_t dfs(int u, int t, int64_t f) {
    if (u == t) return f;
    for (uint64_t &e = it[u]; e < head[u + 1]; ++e) {
      const int v = to[e];
      if (cap[e] > 0 && level[v] == level[u] + 1) {
        const int64_t d = dfs(v, t, std::min(f, cap[e]));
        if (d > 0) {
          cap[e] -= d;
          cap[rev[e]] += d;
          return d;
        }
      }
    }
    return 0;
  }
  int64_t maxflow(int s, int t) {
    int64_t flow = 0;
    while (bfs(s, t)) {
      for (int i = 0; i < n; ++i) it[i] = head[i];
      while (int64_t f = dfs(s, t, INT64_MAX)) flow += f;
    }
    return flow;  // the last bfs left the residual reachability in level
  }
};

// |F - target| < 1e-2f decided exactly.  F_scaled = F * 2^s; target is the
// float (float)(T*k); 1e-2f = 5368709 * 2^-29.
static bool float_equal_exact(int64_t f_scaled, float target, int s) {
  __int128 t = (__int128)(int64_t)target << s;  // target is an integer
  __int128 d = (__int128)f_scaled - t;
  if (d < 0) d = -d;
  // d / 2^s < 5368709 / 2^29  <=>  d * 2^29 < 5368709 * 2^s
  return d * ((__int128)1 << 29) < (__int128)5368709 << s;
}

static void exact_int(std::vector<int> &res, Component &C, ui k, float lb,
                      float ub, int cid, Result &R) {
  float alpha = (lb + ub) * 0.5f;
  float bias = 1.0f / (C.size * (C.size - 1));
  if (bias < 1e-2f) bias = 1e-2f;
  res.assign(C.size, 1);
  const float target = C.total_cliques * k;
  const int T = C.total_cliques, V = C.size;
  const int source = V + T, sink = source + 1;
  long maxcd = 0;
  for (long d : C.clique_degree) maxcd = std::max(maxcd, d);
  if (maxcd >= (1l << 24)) R.quirk_flags |= 2;
  if (!((ub - lb) > bias)) return;
  // The network's shape is the same for every alpha: build it once with
  // the clique and source capacities for unit U = 1, then scale them and
  // set the sink arcs per alpha.
  Dinic D(V + T + 2);
  for (int c = 0; c < T; ++c)
    for (ui i = 0; i < k; ++i) D.add(V + c, C.cliques[(size_t)c * k + i], (int64_t)k - 1, 1);
  for (int v = 0; v < V; ++v) D.add(source, v, C.clique_degree[v]);
  std::vector<uint64_t> sink_arc(V);
  for (int v = 0; v < V; ++v) D.add(v, sink, 0);
  D.build();
  for (int v = 0; v < V; ++v)
    for (uint64_t e = D.head[v]; e < D.head[v + 1]; ++e)
      if (D.to[e] == sink) sink_arc[v] = e;
  const std::vector<int64_t> unit_caps = D.cap0;
  while ((ub - lb) > bias) {
    const float cap_t = alpha * k;
    int e2;
    std::frexp(cap_t, &e2);
    int s = std::max(0, 24 - e2);
    const int64_t ct = (int64_t)std::ldexp((double)cap_t, s);
    for (size_t e = 0; e < unit_caps.size(); ++e) D.cap0[e] = unit_caps[e] << s;
    for (int v = 0; v < V; ++v) D.cap0[sink_arc[v]] = ct;
    D.reset();
    const int64_t F = D.maxflow(source, sink);
    R.tr_comp.push_back(cid);
    R.tr_lb.push_back(lb);
    R.tr_ub.push_back(ub);
    R.tr_alpha.push_back(alpha);
    R.tr_margin.push_back((double)C.total_cliques * k - std::ldexp((double)F, -s));
    const bool eq = float_equal_exact(F, target, s);
    R.tr_equal.push_back(eq);
    if (eq) {
      ub = alpha;
    } else {
      lb = alpha;
      for (int i = 0; i < V; i++) res[i] = D.level[i] >= 0 ? 1 : -1;
    }
    alpha = (lb + ub) * 0.5f;
  }
}

// UB1 / UB2 (helpers.cpp:1168-1198) in the form the reference build
// evaluates them under -ffast-math: num / exp(x) is turned into
// num * exp(-x), the division by dem into a multiply by 1/dem, and the
// square sum is FMA-contracted.  Written out and compiled without fast-math
// so this file's own compiler decisions cannot change the values.
#pragma GCC push_options
#pragma GCC optimize("no-fast-math")
static float ub1_value(ui k, long total) {
  const float kf = (float)k;
  float log_fact = 0.0f;
  for (ui i = 1; i <= k; ++i) log_fact = log_fact + logf((float)i);
  const float num = powf((float)total, (kf - 1.0f) / kf);
  return num * expf(-(log_fact / kf));
}
static float ub2_value(ui k, long total, const std::vector<long> &cd) {
  const float dem = (float)(k * total);
  const float inv = 1.0f / dem;
  float acc = 0.0f;
  for (size_t v = 0; v < cd.size(); v++) {
    const float pv = (float)cd[v] * inv;
    acc = fmaf(pv, pv, acc);
  }
  return acc * (float)total;
}
#pragma GCC pop_options

// dynamicExact, helpers.cpp:1120-1254 (bounds carried across components).
static void dynamic_exact(ui k, int flow_mode, bool ub1, bool ub2, Result &R) {
  Component *current = &R.comps[0];
  float lower = 0.0f;
  for (auto &C : R.comps)
    if (lower < C.density) {
      lower = C.density;
      current = &C;
    }
  if (std::ceil(lower) < std::ceil(R.core_density)) lower = R.core_density;
  float upper = R.kmax;
  R.vertices.resize(current->size);
  for (int i = 0; i < current->size; ++i)
    R.vertices[i] = R.core_reverse_map[current->reverse_map[i]];
  R.size = current->size;
  R.density = lower;
  std::vector<int> res;
  for (ui cid = 0; cid < R.comps.size(); ++cid) {
    Component &C = R.comps[cid];
    if (ub1) {
      R.ub1_val = ub1_value(k, C.total_cliques);
      if (R.ub1_val < upper) upper = R.ub1_val;
    }
    if (ub2) {
      R.ub2_val = ub2_value(k, C.total_cliques, C.clique_degree);
      if (R.ub2_val < upper) upper = R.ub2_val;
    }
    const size_t before = R.tr_alpha.size();
    if (flow_mode == FLOW_FLOAT_EK)
      exact_float(res, C, k, lower, upper, cid, R);
    else
      exact_int(res, C, k, lower, upper, cid, R);
    if (R.tr_alpha.size() > before) R.flow_networks++;
    long clique_count = 0;
    ui vertex_count = 0;
    auto count_clique = [&](const int *t, long weight) {
      ui i = 0;
      for (; i < k; ++i)
        if (res[t[i]] == -1) break;
      if (i == k) clique_count += weight;
    };
    if (flow_mode == FLOW_FLOAT_EK)
      for (auto &e : C.clique_map) count_clique(e.second.data(), e.second[k]);
    else
      for (size_t c = 0; c < C.cliques.size() / k; ++c) count_clique(&C.cliques[c * k], 1);
    for (int i = 0; i < C.size; i++)
      if (res[i] != -1) vertex_count++;
    if (vertex_count == 0) vertex_count = C.size;
    float density = (float)clique_count / (float)vertex_count;
    if (density > R.density) {
      R.density = density;
      lower = density;
      R.size = vertex_count;
      R.set_cliques = clique_count;
      R.vertices.clear();
      for (int i = 0; i < C.size; i++)
        if (res[i] != -1)
          R.vertices.push_back(R.core_reverse_map[C.reverse_map[i]]);
    }
  }
  double lbp = 0;
  for (auto &C : R.comps)
    if (lbp < C.density) lbp = C.density;
  R.paper_lb = lbp;
}

// Number of k-cliques inside a vertex set (exact fraction numerator).
static long cliques_in_set(const std::vector<std::vector<ui>> &adj, ui k,
                           const std::vector<ui> &set) {
  std::vector<int> in(adj.size(), -1);
  for (ui i = 0; i < set.size(); ++i) in[set[i]] = i;
  std::vector<std::vector<ui>> sub(set.size());
  for (ui i = 0; i < set.size(); ++i)
    for (ui u : adj[set[i]])
      if (in[u] >= 0) sub[i].push_back(in[u]);
  long t = 0;
  Lister L;
  L.list_all(sub, k, [&](const std::vector<ui> &, ui, ui) { ++t; });
  return t;
}

static void solve(const std::vector<std::vector<ui>> &adj, ui k, int peel_mode,
                  int flow_mode, bool ub1, bool ub2, Result &R) {
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  R.k = k;
  R.n = adj.size();
  unsigned long long e2 = 0;
  for (auto &l : adj) e2 += l.size();
  R.m = e2 / 2;
  std::vector<ui> members;
  list_full(adj, k, R, peel_mode == PEEL_FAST ? &members : nullptr);
  auto t1 = clk::now();
  peel(adj, k, members, peel_mode, R);
  members.clear();
  members.shrink_to_fit();
  auto t2 = clk::now();
  std::vector<std::vector<ui>> core_graph;
  locate_core(adj, R, core_graph);
  components(core_graph, k, flow_mode, R);
  auto t3 = clk::now();
  dynamic_exact(k, flow_mode, ub1, ub2, R);
  std::sort(R.vertices.begin(), R.vertices.end());
  R.set_cliques = cliques_in_set(adj, k, R.vertices);
  auto t4 = clk::now();
  R.t_list = std::chrono::duration<double, std::milli>(t1 - t0).count();
  R.t_peel = std::chrono::duration<double, std::milli>(t2 - t1).count();
  R.t_flow = std::chrono::duration<double, std::milli>(t4 - t3).count();
  R.t_total = std::chrono::duration<double, std::milli>(t4 - t0).count();
}

// ---- JSON output ----------------------------------------------------------
struct Json {
  std::string s;
  bool first = true;
  void key(const char *k) {
    if (!first) s += ",";
    first = false;
    s += "\"";
    s += k;
    s += "\":";
  }
  // Non-finite values are written as the tokens Python's json accepts; the
  // test is on the bit pattern since this file is built with -ffast-math.
  static void put(std::string &s, double v) {
    uint64_t b;
    memcpy(&b, &v, sizeof b);
    const bool nonfinite = ((b >> 52) & 0x7FF) == 0x7FF;
    if (nonfinite && (b & 0xFFFFFFFFFFFFFull)) {
      s += "NaN";
    } else if (nonfinite) {
      s += (b >> 63) ? "-Infinity" : "Infinity";
    } else {
      char buf[64];
      snprintf(buf, sizeof buf, "%.17g", v);
      s += buf;
    }
  }
  void num(const char *k, double v) {
    key(k);
    put(s, v);
  }
  void inum(const char *k, long long v) {
    key(k);
    s += std::to_string(v);
  }
  template <class T> void arr(const char *k, const std::vector<T> &v) {
    key(k);
    s += "[";
    for (size_t i = 0; i < v.size(); ++i) {
      if (i) s += ",";
      if constexpr (std::is_floating_point<T>::value) {
        put(s, (double)v[i]);
      } else {
        s += std::to_string(v[i]);
      }
    }
    s += "]";
  }
};

static std::string to_json(const Result &R, bool rows) {
  Json j;
  j.s = "{";
  j.inum("k", R.k);
  j.inum("n", R.n);
  j.inum("m", R.m);
  j.inum("total_cliques", R.total_cliques);
  j.arr("clique_degree", R.clique_degree);
  // core number of each vertex = running max of removal values
  std::vector<long> core(R.n, 0);
  long run = 0;
  for (ui i = 1; i <= R.n; ++i) {
    run = std::max(run, R.removal_value[i]);
    core[R.removal_vertex[i]] = run;
  }
  j.arr("core", core);
  if (rows) {
    j.arr("removal_vertex", R.removal_vertex);
    j.arr("removal_value", R.removal_value);
    j.arr("row_density", R.row_density);
  }
  j.num("max_density", R.max_density);
  j.inum("k_prime", R.lower_bound);
  j.inum("kmax", R.kmax);
  j.num("core_density", R.core_density);
  j.inum("core_size", R.core_size);
  j.inum("core_edges", R.core_edges);
  j.inum("core_cliques", R.core_cliques);
  j.inum("deleted_cliques", R.del_clique_count);
  j.inum("pruned_edges", R.pruned_edges);
  j.inum("num_components", R.comps.size());
  std::vector<long> csize, ccl;
  for (auto &C : R.comps) {
    csize.push_back(C.size);
    ccl.push_back(C.total_cliques);
  }
  j.arr("component_size", csize);
  j.arr("component_cliques", ccl);
  j.num("paper_lower_bound", R.paper_lb);
  j.num("ub1", R.ub1_val);
  j.num("ub2", R.ub2_val);
  j.num("density", R.density);
  j.inum("size", R.size);
  j.arr("vertices", R.vertices);
  j.inum("set_cliques", R.set_cliques);
  j.arr("trace_component", R.tr_comp);
  j.arr("trace_lb", R.tr_lb);
  j.arr("trace_ub", R.tr_ub);
  j.arr("trace_alpha", R.tr_alpha);
  j.arr("trace_equal", R.tr_equal);
  j.arr("trace_margin", R.tr_margin);
  j.inum("flow_networks", R.flow_networks);
  j.inum("quirk_flags", R.quirk_flags);
  j.num("t_list_ms", R.t_list);
  j.num("t_peel_ms", R.t_peel);
  j.num("t_flow_ms", R.t_flow);
  j.num("t_total_ms", R.t_total);
  j.s += "}";
  return j.s;
}

static std::vector<std::vector<ui>> adj_from_csr(ui n, const ui *off,
                                                 const ui *nbr) {
  std::vector<std::vector<ui>> adj(n);
  for (ui v = 0; v < n; ++v) adj[v].assign(nbr + off[v], nbr + off[v + 1]);
  return adj;
}

}  // namespace

extern "C" {

// Full pipeline; returns a malloc'd JSON document (free with oracle_free).
char *oracle_solve_json(ui n, const ui *offsets, const ui *neighbors, ui k,
                        int peel_mode, int flow_mode, int ub1, int ub2,
                        int with_rows) {
  auto adj = adj_from_csr(n, offsets, neighbors);
  Result R;
  solve(adj, k, peel_mode, flow_mode, ub1 != 0, ub2 != 0, R);
  std::string s = to_json(R, with_rows != 0);
  char *out = (char *)malloc(s.size() + 1);
  memcpy(out, s.c_str(), s.size() + 1);
  return out;
}

// Listing and the fast peel only: the removal rows (vertex, value) and the
// cliques left after each row, as JSON (free with oracle_free).
char *oracle_peel_rows_json(ui n, const ui *offsets, const ui *neighbors, ui k) {
  auto adj = adj_from_csr(n, offsets, neighbors);
  Result R;
  R.k = k;
  R.n = n;
  std::vector<ui> members;
  list_full(adj, k, R, &members);
  peel(adj, k, members, PEEL_FAST, R);
  Json j;
  j.s = "{";
  j.inum("total_cliques", R.total_cliques);
  j.arr("removal_vertex", R.removal_vertex);
  j.arr("removal_value", R.removal_value);
  j.arr("row_remaining", R.row_remaining);
  j.s += "}";
  char *out = (char *)malloc(j.s.size() + 1);
  memcpy(out, j.s.c_str(), j.s.size() + 1);
  return out;
}

// Listing only: T and per-vertex clique degrees.
long oracle_count_cliques(ui n, const ui *offsets, const ui *neighbors, ui k,
                          long *clique_degree) {
  auto adj = adj_from_csr(n, offsets, neighbors);
  Result R;
  list_full(adj, k, R, nullptr);
  if (clique_degree)
    for (ui v = 0; v < n; ++v) clique_degree[v] = R.clique_degree[v];
  return R.total_cliques;
}

// Exact max flow of one bipartite vertex/clique network at capacity c_t on
// the sink arcs (all capacities already integers): returns F and writes the
// source side reachability of the minimal min cut for the V vertices.
long long oracle_component_flow(int V, int T, ui k, const int *members,
                                const long *cd, long long cap_t,
                                long long unit, int *reach) {
  Dinic D(V + T + 2);
  const int source = V + T, sink = source + 1;
  for (int c = 0; c < T; ++c)
    for (ui i = 0; i < k; ++i)
      D.add(V + c, members[c * k + i], (long long)(k - 1) * unit, unit);
  for (int v = 0; v < V; ++v) {
    D.add(source, v, cd[v] * unit);
    D.add(v, sink, cap_t);
  }
  D.build();
  long long F = D.maxflow(source, sink);
  for (int v = 0; v < V; ++v) reach[v] = D.level[v] >= 0;
  return F;
}

void oracle_free(char *p) { free(p); }
}
