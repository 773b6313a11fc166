// ORACLE — test infrastructure only.  Timing harness for the reference's
// OpenMP CPU variant (cpuCode/parallel), compiled with that variant's own
// sources where they lie (oracle/Makefile, target `ref`).  It does what
// cpuCode/parallel/main.cpp:4-19 does, except that the thread count comes
// from the command line instead of THREAD_COUNT (cpuCode/parallel/inc/
// common.h:38, fixed at 2), so the variant can be timed at its built-in two
// threads and at every core.  The variant sizes its per-thread state with
// omp_get_max_threads() (cpuCode/parallel/src/helpers.cpp:328, :996,
// :1445), so any count is valid.  Its answer depends on the thread
// schedule (it is not a parity oracle); only its run time is used.
#include "common.h"
#include "graph.h"
#include "helpers.h"

int main(int argc, const char *argv[]) {
  if (argc != 6) {
    fprintf(stderr, "usage: dsd_parallel <graph> <motif> <ub1> <ub2> <threads>\n");
    return 1;
  }
  const int threads = atoi(argv[5]);
  if (threads < 1) {
    fprintf(stderr, "dsd_parallel: threads must be >= 1\n");
    return 1;
  }
  omp_set_num_threads(threads);
  Graph graph{std::string(argv[1])};
  Motif motif{std::string(argv[2])};
  CDS cds(&graph, &motif, atoi(argv[3]) != 0, atoi(argv[4]) != 0);
  cds.DSD();
  return 0;
}
