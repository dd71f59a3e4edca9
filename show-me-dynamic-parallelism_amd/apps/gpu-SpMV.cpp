// gpu-SpMV: the reference's SpMV driver (apps/SpMV/src/SpMV.cpp) on the HIP
// library.  The matrix is the graph's CSR; data and x are the reference's
// deterministic inputs (SpMV_wrapper.cu:105-110), here in FP32 with
// compensated accumulation.  The FP64 host product is computed for the error
// report printed under -v.
#include <math.h>
#include <stdlib.h>

#include "cli_common.h"

static void usage() {
  fprintf(stderr, "\n");
  fprintf(stderr, "Usage:  gpu-SpMV [option]\n");
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
  fprintf(stderr, "                 0 - no dynamic parallelism (lane per row)\n");
  fprintf(stderr, "                 1 - naive: heavy row reduced at once by its wave\n");
  fprintf(stderr, "                 2 - warp(wave)-level consolidation\n");
  fprintf(stderr, "                 3 - block(workgroup)-level consolidation\n");
  fprintf(stderr, "                 4 - grid-level consolidation\n");
  fprintf(stderr, "                 6 - tuned: LDS-staged row tiles + grid chunks\n");
  fprintf(stderr, "    --device,-e <number>               select the device\n");
  fprintf(stderr, "    --threshold <nnz>                  heavy-row length (default 64)\n");
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
  if (int rc = cli_apply_tuning(c, ctx, &csr, SMDP_APP_SPMV, 0)) return rc;
  size_t n = g->n > 0 ? g->n : 1, m = g->m > 0 ? g->m : 1;
  float* data = (float*)malloc(sizeof(float) * m);
  float* x = (float*)malloc(sizeof(float) * n);
  float* y = (float*)malloc(sizeof(float) * n);
  smdp_gen_spmv_reference_inputs(g->n, g->m, data, x);
  smdp_stats st;
  int rc = smdp_spmv_f32(ctx, &csr, data, x, c.solution, y, &st);
  if (rc != SMDP_OK) return cli_fail("SpMV", rc);
  if (c.verbose) {
    double linf = 0, ymax = 0;
    for (int32_t r = 0; r < g->n; ++r) {
      long double acc = 0;
      for (int32_t e = g->row[r]; e < g->row[r + 1]; ++e) acc += (long double)data[e] * x[g->col[e]];
      linf = fmax(linf, fabs((double)acc - (double)y[r]));
      ymax = fmax(ymax, fabs((double)acc));
    }
    fprintf(stdout, "HIP kernel execution:\t\t%lf\n", st.ms_kernel / 1e3);
    fprintf(stdout, "HIP total (malloc + H2D + kernels + D2H):\t\t%lf\n", st.ms_total / 1e3);
    fprintf(stdout, "L-inf error vs FP64: %.3e (max |y| %.3e)\n", linf, ymax);
    cli_print_counters(stdout, st);
  }
  if (c.debug) {
    for (int32_t i = 0; i < g->n; ++i) {
      fprintf(stderr, "x[%d] = %f\n", i, x[i]);
      fprintf(stderr, "y[%d] = %f\n", i, y[i]);
    }
  }
  free(data);
  free(x);
  free(y);
  smdp_destroy(ctx);
  smdp_hgraph_free(g);
  return 0;
}
