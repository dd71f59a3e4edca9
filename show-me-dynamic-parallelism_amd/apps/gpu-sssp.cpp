// gpu-sssp: the reference's SSSP driver (apps/sssp/src/sssp.cpp) on the HIP
// library.  Same flags, same -s numbering, same -d output (outputCost).
#include <stdlib.h>

#include "cli_common.h"

static void usage() {
  // apps/sssp/src/sssp.cpp:19-38
  fprintf(stderr, "\n");
  fprintf(stderr, "Usage:  gpu-sssp [option]\n");
  fprintf(stderr, "\nOptions:\n");
  fprintf(stderr, "    --help,-h      print this message\n");
  fprintf(stderr, "    --verbose,-v   basic verbosity level\n");
  fprintf(stderr, "    --debug,-d     enhanced verbosity level\n");
  fprintf(stderr, "\nOther:\n");
  fprintf(stderr, "    --import,-i <graph_file>           import graph file\n");
  fprintf(stderr, "    --format,-f <number>               specify the input format\n");
  fprintf(stderr, "                 0 - DIMACS9\n");
  fprintf(stderr, "                 1 - DIMACS10\n");
  fprintf(stderr, "                 2 - SLNDC\n");
  fprintf(stderr, "    --solution,-s <number>             specify the solution\n");
  fprintf(stderr, "                 0 - no dynamic parallelism (unordered + thread queue)\n");
  fprintf(stderr, "                 1 - naive: heavy vertex expanded at once by its wave\n");
  fprintf(stderr, "                 2 - warp(wave)-level consolidation\n");
  fprintf(stderr, "                 3 - block(workgroup)-level consolidation\n");
  fprintf(stderr, "                 4 - grid-level consolidation\n");
  fprintf(stderr, "                 6 - tuned: edge-balanced LDS tiles + grid chunks\n");
  fprintf(stderr, "    --device,-e <number>               select the device\n");
  fprintf(stderr, "    --threshold <degree>               heavy-vertex degree (default 64)\n");
  fprintf(stderr, "    --tune <cache.json>                autotune the tuned solution\n");
}

int main(int argc, char** argv) {
  CliConf c;
  if (!cli_parse(argc, argv, &c, usage)) return 0;
  cli_print_conf(c);
  if (!c.graph_file) return 0;
  smdp_hgraph* g = nullptr;
  if (int rc = cli_load(c, &g)) return rc;
  smdp_ctx* ctx = nullptr;
  double t0 = cli_now();
  if (int rc = cli_context(c, &ctx)) return rc;
  double init_time = cli_now() - t0;
  smdp_csr csr = cli_csr(g);
  if (int rc = cli_apply_tuning(c, ctx, &csr, SMDP_APP_SSSP, g->source)) return rc;
  int32_t* cost = (int32_t*)malloc(sizeof(int32_t) * (g->n > 0 ? g->n : 1));
  smdp_stats st;
  int rc = smdp_sssp(ctx, &csr, g->source, c.solution, cost, &st);
  if (rc != SMDP_OK) return cli_fail("SSSP", rc);
  if (c.verbose) {
    fprintf(stdout, "HIP runtime initialization:\t\t%lf\n", init_time);
    fprintf(stdout, "HIP kernel execution:\t\t%lf\n", st.ms_kernel / 1e3);
    fprintf(stdout, "HIP total (malloc + H2D + kernels + D2H):\t\t%lf\n", st.ms_total / 1e3);
    fprintf(stdout, "Rounds: %llu  arcs scanned: %llu  frontier entries: %llu\n", (unsigned long long)st.rounds,
            (unsigned long long)st.edges_relaxed, (unsigned long long)st.vertices_expanded);
    fprintf(stdout, "GTEPS: %.3f\n", st.ms_kernel > 0 ? st.m_reached / (st.ms_kernel * 1e6) : 0.0);
    cli_print_counters(stdout, st);
  }
  if (c.debug) {
    fprintf(stderr, "Iteration: %llu\n", (unsigned long long)st.rounds);
    output_cost(stdout, g->source, g->n, cost);
  }
  free(cost);
  smdp_destroy(ctx);
  smdp_hgraph_free(g);
  return 0;
}
