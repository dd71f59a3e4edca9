// gpu-graph-color: the reference's coloring driver
// (apps/graph-color/src/graph_color.cpp) on the HIP library.  Vertex-ID
// priority by default, as in the reference; --prio-seed S uses seeded
// uniform 32-bit priorities with ties broken by vertex id.
#include <stdlib.h>

#include "cli_common.h"

static void usage() {
  fprintf(stderr, "\n");
  fprintf(stderr, "Usage:  gpu-graph-color [option]\n");
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
  fprintf(stderr, "                 0 - no dynamic parallelism (lane per vertex)\n");
  fprintf(stderr, "                 1 - naive: heavy vertex scanned at once by its wave\n");
  fprintf(stderr, "                 2 - warp(wave)-level consolidation\n");
  fprintf(stderr, "                 3 - block(workgroup)-level consolidation\n");
  fprintf(stderr, "                 4 - grid-level consolidation\n");
  fprintf(stderr, "                 6 - tuned\n");
  fprintf(stderr, "    --device,-e <number>               select the device\n");
  fprintf(stderr, "    --prio-seed <seed>                 random vertex priorities\n");
}

int main(int argc, char** argv) {
  CliConf c;
  if (!cli_parse(argc, argv, &c, usage)) return 0;
  cli_print_conf(c);
  if (!c.graph_file) return 0;
  smdp_hgraph* g = nullptr;
  if (int rc = cli_load(c, &g)) return rc;
  smdp_ctx* ctx = nullptr;
  if (int rc = cli_context(c, &ctx)) return rc;
  smdp_csr csr = cli_csr(g);
  if (int rc = cli_apply_tuning(c, ctx, &csr, SMDP_APP_COLOR, 0)) return rc;
  uint32_t* prio = nullptr;
  if (c.prio_seed >= 0) {
    prio = (uint32_t*)malloc(sizeof(uint32_t) * (g->n > 0 ? g->n : 1));
    smdp_gen_prio(g->n, (uint64_t)c.prio_seed, prio);
  }
  int32_t* color = (int32_t*)malloc(sizeof(int32_t) * (g->n > 0 ? g->n : 1));
  smdp_stats st;
  int rc = smdp_color(ctx, &csr, prio, c.solution, color, &st);
  if (rc != SMDP_OK) return cli_fail("coloring", rc);
  if (c.verbose) {
    fprintf(stdout, "HIP kernel execution:\t\t%lf\n", st.ms_kernel / 1e3);
    fprintf(stdout, "HIP total (malloc + H2D + kernels + D2H):\t\t%lf\n", st.ms_total / 1e3);
    cli_print_counters(stdout, st);
  }
  if (c.debug) {
    fprintf(stderr, "Graph Coloring ends in %llu iterations.\n", (unsigned long long)st.rounds);
    output_color(stdout, g->n, color);
  }
  free(color);
  free(prio);
  smdp_destroy(ctx);
  smdp_hgraph_free(g);
  return 0;
}
