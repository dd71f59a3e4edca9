// gpu-pg: the reference's PageRank driver (apps/pagerank/src/pagerank.cpp) on
// the HIP library.  Pull PageRank over the inverted graph (invertGraph) with
// the out-degrees of the original graph.  Defaults keep the reference's
// damping 0.5 and total rank 1000 (pagerank_wrapper.cu:7, pagerank.h:15) and
// its initial ranks T/N (pagerank.cpp:103-106); the stop rule is an FP64 L1
// residual (--tol, relative to the total rank) or --max-iter iterations.
// Under -v the same number of iterations is repeated on the host in FP64 and
// the L1 and L-inf distances of the device ranks from it are printed.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cli_common.h"

static void usage() {
  fprintf(stderr, "\n");
  fprintf(stderr, "Usage:  gpu-pg [option]\n");
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
  fprintf(stderr, "                 0 - no dynamic parallelism (pull, lane per vertex)\n");
  fprintf(stderr, "                 1 - naive: heavy vertex reduced at once by its wave\n");
  fprintf(stderr, "                 2 - warp(wave)-level consolidation\n");
  fprintf(stderr, "                 3 - block(workgroup)-level consolidation\n");
  fprintf(stderr, "                 4 - grid-level consolidation\n");
  fprintf(stderr, "                 6 - tuned: LDS-staged row tiles + grid chunks\n");
  fprintf(stderr, "    --device,-e <number>               select the device\n");
  fprintf(stderr, "    --damping <d>  --total-rank <T>  --tol <l1>  --max-iter <k>  --init-zero\n");
}

// `iters` FP64 pull iterations with the device's update rule: the mass of
// out-degree-0 vertices is spread evenly, r' = d*(sum_in r/outdeg + dang/N)
// + (1-d)*T/N.
static void host_pagerank_f64(const smdp_hgraph* ig, const int32_t* outdeg, double d, double T, int init_zero,
                              uint64_t iters, double* r) {
  const int32_t n = ig->n;
  double* nr = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  double* contrib = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  for (int32_t v = 0; v < n; ++v) r[v] = init_zero ? 0.0 : T / n;
  for (uint64_t it = 0; it < iters; ++it) {
    double dang = 0;
    for (int32_t u = 0; u < n; ++u) {
      if (outdeg[u] == 0) dang += r[u];
      contrib[u] = outdeg[u] ? r[u] / outdeg[u] : 0.0;
    }
    const double base = (1.0 - d) * T / n + d * dang / n;
    for (int32_t v = 0; v < n; ++v) {
      double s = 0;
      for (int32_t e = ig->row[v]; e < ig->row[v + 1]; ++e) s += contrib[ig->col[e]];
      nr[v] = d * s + base;
    }
    memcpy(r, nr, sizeof(double) * n);
  }
  free(nr);
  free(contrib);
}

int main(int argc, char** argv) {
  CliConf c;
  if (!cli_parse(argc, argv, &c, usage)) return 0;
  cli_print_conf(c);
  if (!c.graph_file) return 0;
  smdp_hgraph* g = nullptr;
  if (int rc = cli_load(c, &g)) return rc;
  smdp_hgraph* ig = nullptr;
  double t0 = cli_now();
  int rc = smdp_invert(g, &ig);
  if (rc != SMDP_OK) return cli_fail("inverting the graph", rc);
  if (c.verbose) fprintf(stderr, "Generate inverted graph:\t\t%lf\n", cli_now() - t0);
  int32_t* outdeg = (int32_t*)malloc(sizeof(int32_t) * (g->n > 0 ? g->n : 1));
  smdp_outdegree(g, outdeg);
  smdp_ctx* ctx = nullptr;
  if (int rc2 = cli_context(c, &ctx)) return rc2;
  smdp_csr icsr = cli_csr(ig);
  icsr.w = nullptr;
  if (int rc2 = cli_apply_tuning(c, ctx, &icsr, SMDP_APP_PAGERANK, 0)) return rc2;
  smdp_pr_opts o;
  o.damping = c.damping >= 0 ? (float)c.damping : 0.5f;
  o.total_rank = c.total_rank > 0 ? (float)c.total_rank : 1000.0f;
  o.tol_l1 = c.tol > 0 ? c.tol : 1e-6;
  o.max_iter = c.max_iter > 0 ? c.max_iter : 100;
  o.init_zero = c.init_zero;
  float* rank = (float*)malloc(sizeof(float) * (g->n > 0 ? g->n : 1));
  smdp_stats st;
  rc = smdp_pagerank_f32(ctx, &icsr, outdeg, &o, c.solution, rank, &st);
  if (rc != SMDP_OK) return cli_fail("PageRank", rc);
  if (c.verbose) {
    fprintf(stdout, "HIP kernel execution:\t\t%lf\n", st.ms_kernel / 1e3);
    fprintf(stdout, "HIP total (malloc + H2D + kernels + D2H):\t\t%lf\n", st.ms_total / 1e3);
    fprintf(stdout, "L1 residual: %.3e  dangling mass: %.6e\n", st.residual, st.dangling);
    double* ref = (double*)malloc(sizeof(double) * (g->n > 0 ? g->n : 1));
    host_pagerank_f64(ig, outdeg, o.damping, o.total_rank, o.init_zero, st.rounds, ref);
    double l1 = 0, linf = 0;
    for (int32_t v = 0; v < g->n; ++v) {
      const double e = fabs((double)rank[v] - ref[v]);
      l1 += e;
      linf = fmax(linf, e);
    }
    fprintf(stdout, "vs FP64 host (%llu iterations): L1 %.3e  L-inf %.3e  (total rank %g)\n",
            (unsigned long long)st.rounds, l1, linf, (double)o.total_rank);
    free(ref);
    cli_print_counters(stdout, st);
  }
  if (c.debug) {
    fprintf(stderr, "PageRank ends in %llu iterations.\n", (unsigned long long)st.rounds);
    output_rank(stdout, g->n, rank);
  }
  free(rank);
  free(outdeg);
  smdp_destroy(ctx);
  smdp_hgraph_free(ig);
  smdp_hgraph_free(g);
  return 0;
}
