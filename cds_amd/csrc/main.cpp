// cds_main: drop-in for the reference GPU binary.
//
//   cds_main <graph> <k> <pSize> <cpSize> <glBufferSize> <partitionSize> <earlyStop>
//
// Same argv (cudaCode/main.cu:865-887), same input format
// (cudaCode/src/graph.cpp:7-40) and the same report keys
// (cudaCode/src/graph.cpp:39, :59; cudaCode/main.cu:861, :994-995,
// :1069-1104), followed by the result block of the reference's CPU oracle
// (cpuCode/c++/src/helpers.cpp:1580-1588: K, Density, Size, densest
// subgraph) and one exact line `Density_exact=<cliques>/<size>`.
// pSize, cpSize, glBufferSize and partitionSize are capacity hints of the
// reference's fixed partitions; this engine sizes its buffers from the
// graph, so they are accepted and validated but change nothing.  earlyStop
// must be -1 (exact search); other values are rejected.
//
// Environment (none changes the printed report):
//   CDS_CSR_CACHE=<dir>  keep a binary CSR copy of each text graph in <dir>,
//                        keyed by file name, size and modification time, and
//                        load it instead of parsing the text again
//   CDS_JSON=<path>      also write the result as one JSON object to <path>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cds.h"

static int die(const char *what) {
  fprintf(stderr, "cds_main: %s: %s\n", what, cds_last_error());
  return 2;
}

static cds_status load_graph(const char *path, cds_graph *g) {
  const char *dir = getenv("CDS_CSR_CACHE");
  struct stat st;
  if (!dir || !*dir || stat(path, &st) != 0) return cds_graph_load_adjlist(path, g);
  const char *base = strrchr(path, '/');
  base = base ? base + 1 : path;
  const std::string cache = std::string(dir) + "/" + base + "." +
                            std::to_string((long long)st.st_size) + "." +
                            std::to_string((long long)st.st_mtime) + ".csr";
  if (cds_graph_load_csr(cache.c_str(), g) == CDS_OK) return CDS_OK;
  const cds_status s = cds_graph_load_adjlist(path, g);
  if (s == CDS_OK) {
    // written under a temporary name and renamed, so a reader never sees a
    // partial file; a cache that cannot be written is skipped silently
    const std::string tmp = cache + ".tmp" + std::to_string((long long)getpid());
    if (cds_graph_save_csr(g, tmp.c_str()) == CDS_OK)
      rename(tmp.c_str(), cache.c_str());
    else
      unlink(tmp.c_str());
  }
  return s;
}

static void write_json(const cds_result *r) {
  const char *path = getenv("CDS_JSON");
  if (!path || !*path) return;
  std::vector<char> buf(cds_result_json(r, nullptr, 0) + 1);
  cds_result_json(r, buf.data(), buf.size());
  FILE *f = fopen(path, "w");
  if (!f || fprintf(f, "%s\n", buf.data()) < 0 || fclose(f) != 0)
    fprintf(stderr, "cds_main: cannot write %s\n", path);
}

int main(int argc, const char *argv[]) {
  if (argc != 8) {
    printf("Server wrong input parameters!\n");
    return 1;
  }
  cds_params p;
  memset(&p, 0, sizeof p);
  p.k = (uint32_t)atoi(argv[2]);
  p.p_size = (uint32_t)atoi(argv[3]);
  p.cp_size = (uint32_t)atoi(argv[4]);
  p.gl_buffer_size = (uint32_t)atoi(argv[5]);
  p.partition_size = (uint32_t)atoi(argv[6]);
  p.early_stop = atoi(argv[7]);
  p.ub1 = 1;
  p.ub2 = 1;
  if (p.early_stop != -1) {
    fprintf(stderr, "cds_main: earlyStop=%d: only -1 (exact) is supported\n",
            p.early_stop);
    return 1;
  }
  cds_graph g;
  memset(&g, 0, sizeof g);
  if (load_graph(argv[1], &g) != CDS_OK) return die("load");
  printf("n=%u, m=%llu\n", g.n,
         (unsigned long long)(g.m_header ? g.m_header : g.m));
  printf("Max Degree=%u\n", g.max_degree);
  // launch shapes: CDS_TUNE_CACHE if set, else the tuned cache shipped
  // beside the binary (cds_amd/tune_cache.txt for cds_amd/bin/cds_main)
  cds_opts o;
  memset(&o, 0, sizeof o);
  std::string shipped;
  if (!getenv("CDS_TUNE_CACHE")) {
    char exe[4096];
    const ssize_t len = readlink("/proc/self/exe", exe, sizeof exe - 1);
    if (len > 0) {
      exe[len] = 0;
      std::string dir(exe);
      dir = dir.substr(0, dir.rfind('/'));
      shipped = dir + "/../tune_cache.txt";
      struct stat st;
      if (stat(shipped.c_str(), &st) == 0) o.tune_cache = shipped.c_str();
    }
  }
  cds_engine *e = nullptr;
  if (cds_engine_create(&o, &e) != CDS_OK) return die("engine");
  cds_dgraph *dg = nullptr;
  if (cds_graph_upload(e, &g, &dg) != CDS_OK) return die("upload");
  cds_result r;
  if (cds_solve(e, dg, &p, &r) != CDS_OK) return die("solve");
  if (r.total_cliques == 0) {
    printf("Max Density 0 Size 0\n");
    printf("Time taken: %g ms\n", r.times.total);
  } else {
    printf("Final Max Density %g\n", (double)r.density);
    printf("Total Cliques=%llu\n", (unsigned long long)r.total_cliques);
    printf("Densest Core Density=%g\n", r.peel_max_density);
    printf("Total Components=%u\n", r.num_components);
    printf("\n");
    printf("DAG Time =%g ms\n", r.times.dag);
    printf("Clique Listing Time =%g ms\n", r.times.listing);
    printf("Core Decompose Time =%g ms\n", r.times.core_decompose);
    printf("Densest Core Time =%g ms\n", r.times.densest_core);
    printf("Prune Time =%g ms\n", r.times.prune);
    printf("Connected Component Time =%g ms\n", r.times.components);
    printf("Exact Time =%g ms\n", r.times.exact);
    printf("Time taken=%g ms\n", r.times.total);
    printf("\n");
    printf("Mac Clique Core=%llu\n", (unsigned long long)r.kmax);
    printf("lb=");
    for (uint32_t i = 0; i < r.trace_len; ++i) printf("%g ", (double)r.trace_lb[i]);
    printf("\nub=");
    for (uint32_t i = 0; i < r.trace_len; ++i) printf("%g ", (double)r.trace_ub[i]);
    printf("\neub=");
    if (p.ub2) printf("%g ", (double)r.ub2_val);
    printf("\n");
  }
  printf("K: %u\n", r.k);
  printf("Density: %g\n", (double)r.density);
  printf("Size: %u\n", r.size);
  printf("densest subgraph: ");
  for (uint32_t i = 0; i < r.set_size; ++i) printf("%u ", r.vertices[i]);
  printf("\n");
  printf("Density_exact=%llu/%u\n", (unsigned long long)r.set_cliques,
         r.set_size);
  write_json(&r);
  cds_result_free(&r);
  cds_dgraph_free(e, dg);
  cds_engine_destroy(e);
  cds_graph_free(&g);
  return 0;
}
