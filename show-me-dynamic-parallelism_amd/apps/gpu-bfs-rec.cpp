// gpu-bfs-rec: the reference's recursive-BFS driver (apps/bfs-rec/src/bfs_rec.cpp)
// on the HIP library.  bfs-rec numbers its solutions differently from the
// other apps (bfs_rec.cpp:30-36); they are mapped to the library's solutions.
#include <stdlib.h>

#include "cli_common.h"

static void usage() {
  fprintf(stderr, "\n");
  fprintf(stderr, "Usage:  gpu-bfs-rec [option]\n");
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
  fprintf(stderr, "                 0 - flat parallelism\n");
  fprintf(stderr, "                 1 - naive: improved vertex expanded by its wave\n");
  fprintf(stderr, "                 2 - hierarchical: every vertex expanded by a wave\n");
  fprintf(stderr, "                 3 - warp(wave)-level consolidation\n");
  fprintf(stderr, "                 4 - block(workgroup)-level consolidation\n");
  fprintf(stderr, "                 5 - grid-level consolidation\n");
  fprintf(stderr, "                 6 - tuned: edge-balanced LDS tiles + grid chunks\n");
  fprintf(stderr, "    --device,-e <number>               select the device\n");
}

int main(int argc, char** argv) {
  CliConf c;
  if (!cli_parse(argc, argv, &c, usage)) return 0;
  cli_print_conf(c);
  if (!c.graph_file) return 0;
  static const int map[] = {SMDP_SOL_FLAT, SMDP_SOL_NAIVE, SMDP_SOL_HIER, SMDP_SOL_WAVE,
                            SMDP_SOL_BLOCK, SMDP_SOL_GRID, SMDP_SOL_TUNED};
  if (c.solution < 0 || c.solution > 6) {
    printf("Unsopported solutions\n");
    return 1;
  }
  smdp_hgraph* g = nullptr;
  if (int rc = cli_load(c, &g)) return rc;
  printf("\n===MAIN=== :: [num_nodes,num_edges] = %u, %u\n", (unsigned)g->n, (unsigned)g->m);
  smdp_ctx* ctx = nullptr;
  if (int rc = cli_context(c, &ctx)) return rc;
  smdp_csr csr = cli_csr(g);
  CliConf tc = c;
  tc.solution = map[c.solution];
  if (int rc = cli_apply_tuning(tc, ctx, &csr, SMDP_APP_BFS, g->source)) return rc;
  int32_t* level = (int32_t*)malloc(sizeof(int32_t) * (g->n > 0 ? g->n : 1));
  int32_t* level_cpu = (int32_t*)malloc(sizeof(int32_t) * (g->n > 0 ? g->n : 1));
  smdp_stats st;
  int rc = smdp_bfs(ctx, &csr, g->source, map[c.solution], level, &st);
  if (rc != SMDP_OK) return cli_fail("BFS", rc);
  bfs_cpu(g, g->source, level_cpu);
  validate_arrays(g->n, level, level_cpu, "GPU bfs rec");
  if (c.verbose) {
    fprintf(stdout, "===MAIN=== :: HIP kernel execution:\t\t%.2lf ms.\n", st.ms_kernel);
    fprintf(stdout, "===MAIN=== :: HIP total (malloc + H2D + kernels + D2H):\t\t%.2lf ms.\n", st.ms_total);
    fprintf(stdout, "===MAIN=== :: levels: %llu  GTEPS: %.3f\n", (unsigned long long)st.rounds,
            st.ms_kernel > 0 ? st.m_reached / (st.ms_kernel * 1e6) : 0.0);
    cli_print_counters(stdout, st);
  }
  if (c.debug) output_level(stdout, g->n, level);
  free(level);
  free(level_cpu);
  smdp_destroy(ctx);
  smdp_hgraph_free(g);
  return 0;
}
