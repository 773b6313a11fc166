// ORACLE — test infrastructure only.  Driver for the reference's own CPU
// sources (cpuCode/c++/src/*.cpp, compiled unmodified from /root/reference by
// oracle/Makefile into oracle/_ref/).  It runs the stages of CDS::DSD
// (cpuCode/c++/src/helpers.cpp:1422-1589) through the public methods declared
// in cpuCode/c++/inc/helpers.h:52-75 and prints the intermediate values as
// one JSON document, so tests can pin the restatement in cds_oracle.cpp
// against the reference itself.
//
// usage: dsd_harness <graph> <k> <ub1> <ub2> [rows]
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "graph.h"
#include "helpers.h"

// JSON has no literal for non-finite values; they are printed as the
// tokens Python's json module accepts.  The test is on the bit pattern
// because this file is built with -ffast-math, under which isnan() folds.
static void put_val(double v) {
  uint64_t b;
  memcpy(&b, &v, sizeof b);
  const bool nonfinite = ((b >> 52) & 0x7FF) == 0x7FF;
  if (nonfinite && (b & 0xFFFFFFFFFFFFFull))
    printf("NaN");
  else if (nonfinite)
    printf((b >> 63) ? "-Infinity" : "Infinity");
  else
    printf("%.17g", v);
}

static void put_num(bool &first, const char *k, double v) {
  printf("%s\"%s\":", first ? "" : ",", k);
  put_val(v);
  first = false;
}

template <class T>
static void put_arr(bool &first, const char *k, const std::vector<T> &v) {
  printf("%s\"%s\":[", first ? "" : ",", k);
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) printf(",");
    put_val((double)v[i]);
  }
  printf("]");
  first = false;
}

int main(int argc, char **argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: %s <graph> <k> <ub1> <ub2> [rows]\n", argv[0]);
    return 2;
  }
  Graph graph{std::string(argv[1])};
  Motif motif;
  motif.size = (ui)atoi(argv[2]);
  motif.type = 0;
  motif.count = 0;
  const bool ub1 = atoi(argv[3]) != 0, ub2 = atoi(argv[4]) != 0;
  const bool rows = argc > 5 && atoi(argv[5]) != 0;
  CDS cds(&graph, &motif, ub1, ub2);

  std::vector<std::vector<double>> results;
  cds.cliqueCoreDecompose(results);
  DensestCoreData core;
  cds.locateDensestCore(results, core);
  unordered_map<string, vector<int>> clique_data;
  vector<long> cd_core;
  cds.cliqueEnumerationListRecord(core.graph, clique_data, cd_core, motif.size);
  vector<vector<ui>> pruned;
  int valid_edges = cds.pruneInvalidEdges(core.graph, pruned, clique_data);
  const size_t core_cliques = clique_data.size();
  vector<ConnectedComponentData> comps;
  cds.connectedComponentDecompose(pruned, clique_data, comps);
  finalResult best;
  float ub1_val = 0, ub2_val = 0;
  cds.dynamicExact(comps, core, best, ub1_val, ub2_val, ub1, ub2);

  bool first = true;
  printf("{");
  put_num(first, "k", motif.size);
  put_num(first, "n", graph.n);
  put_num(first, "total_cliques", graph.totalCliques);
  put_arr(first, "clique_degree", graph.cliqueDegree);
  std::vector<double> core_num(graph.n, 0), rv, rval, rd;
  double run = 0;
  for (ui i = 1; i <= graph.n; ++i) {
    run = std::max(run, results[i][1]);
    core_num[(ui)results[i][0]] = run;
  }
  put_arr(first, "core", core_num);
  if (rows) {
    for (ui i = 0; i <= graph.n; ++i) {
      rv.push_back(results[i][0]);
      rval.push_back(results[i][1]);
      rd.push_back(results[i][2]);
    }
    put_arr(first, "removal_vertex", rv);
    put_arr(first, "removal_value", rval);
    put_arr(first, "row_density", rd);
  }
  put_num(first, "k_prime", core.lowerBound);
  put_num(first, "kmax", core.maxCliqueCore);
  put_num(first, "core_density", core.density);
  put_num(first, "core_size", core.graph.size());
  long e2 = 0;
  for (auto &l : core.graph) e2 += l.size();
  put_num(first, "core_edges", e2 / 2);
  put_num(first, "core_cliques", core_cliques);
  put_num(first, "deleted_cliques", core.delCliqueCount);
  put_num(first, "pruned_edges", valid_edges);
  put_num(first, "num_components", comps.size());
  std::vector<double> csize, ccl;
  double lbp = 0;
  for (auto &c : comps) {
    csize.push_back(c.size);
    ccl.push_back(c.totalCliques);
    if (lbp < c.density) lbp = c.density;
  }
  put_arr(first, "component_size", csize);
  put_arr(first, "component_cliques", ccl);
  put_num(first, "paper_lower_bound", lbp);
  put_num(first, "ub1", ub1_val);
  put_num(first, "ub2", ub2_val);
  put_num(first, "density", best.density);
  put_num(first, "size", best.size);
  put_arr(first, "vertices", best.verticies);
  printf("}\n");
  return 0;
}
