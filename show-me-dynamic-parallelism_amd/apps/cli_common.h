// Shared pieces of the command-line drivers (gpu-sssp, gpu-SpMV, gpu-pg,
// gpu-graph-color, gpu-bfs-rec).  They keep the reference's flags
// (apps/sssp/src/sssp.cpp:59-98, identical in every app), input formats and
// text outputs (common/global.cpp:361-399).
#pragma once
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "smdp.h"

struct CliConf {
  bool verbose = false;
  bool debug = false;
  int data_set_format = 0;
  int solution = 0;
  int device_num = 0;
  const char* graph_file = nullptr;
  // additions (not in the reference)
  const char* tune_cache = nullptr;  // --tune <cache.json>: run the autotuner for the tuned solution
  double damping = -1, total_rank = -1, tol = -1;
  int max_iter = -1;
  int init_zero = 0;
  long long prio_seed = -1;  // --prio-seed S: random 32-bit priorities for coloring
  int threshold = 0;         // --threshold T: heavy degree (reference THRESHOLD=64)
};

// Parses the reference flags plus the additions.  Returns 0 when the program
// should exit 0 (usage printed, or no arguments), 1 to continue.
int cli_parse(int argc, char** argv, CliConf* c, void (*usage)());
void cli_print_conf(const CliConf& c);

// Reads the graph named by -i/-f.  Prints the reference's messages and returns
// a process exit code (0 on success).
int cli_load(const CliConf& c, smdp_hgraph** g);

// Creates the device context; applies --threshold and --tune.
int cli_context(const CliConf& c, smdp_ctx** ctx);
int cli_apply_tuning(const CliConf& c, smdp_ctx* ctx, const smdp_csr* csr, int app, int32_t src);

int cli_fail(const char* what, int rc);
double cli_now();
// Prints the work counters of a call (-v): the heavy vertices (or rows) handed
// to each queue level, which stand where the reference's GPU_PROFILE build
// counts nested kernel calls (apps/sssp/src/sssp_kernel.cu:14-25), plus the
// kernels launched and, for SSSP/BFS, the atomics issued and won.
void cli_print_counters(FILE* f, const smdp_stats& st);

// Reference text outputs.
void output_cost(FILE* f, int32_t source, int32_t n, const int32_t* cost);  // global.cpp:369-375
void output_rank(FILE* f, int32_t n, const float* rank);                      // global.cpp:385-391
void output_color(FILE* f, int32_t n, const int32_t* color);                  // global.cpp:393-399
void output_level(FILE* f, int32_t n, const int32_t* level);                  // global.cpp:361-367

// Host-side checks the drivers print (the reference's own bfs_cpu +
// validateArrays, apps/bfs-rec/src/bfs_rec.cpp:102-141, and FP64 error
// reports for the FP32 kernels).
void bfs_cpu(const smdp_hgraph* g, int32_t source, int32_t* level);
int validate_arrays(int32_t n, const int32_t* a, const int32_t* b, const char* message);

static inline smdp_csr cli_csr(const smdp_hgraph* g) {
  smdp_csr c;
  c.n = g->n;
  c.m = g->m;
  c.row = g->row;
  c.col = g->col;
  c.w = g->w;
  return c;
}
