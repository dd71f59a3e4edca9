#include "cli_common.h"

#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <deque>

double cli_now() {
  struct timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + t.tv_usec * 1e-6;
}

int cli_parse(int argc, char** argv, CliConf* c, void (*usage)()) {
  // apps/sssp/src/sssp.cpp:59-98: unknown flags are ignored.
  if (argc < 2) {
    usage();
    return 0;
  }
  auto value = [&](int& i) -> const char* {
    ++i;
    if (i >= argc) {
      fprintf(stderr, "Missing value for %s.\n", argv[i - 1]);
      return nullptr;
    }
    return argv[i];
  };
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    const char* v = nullptr;
    if (!strcmp(a, "-h") || !strcmp(a, "--help")) {
      usage();
      return 0;
    } else if (!strcmp(a, "-v") || !strcmp(a, "--verbose")) {
      c->verbose = true;
    } else if (!strcmp(a, "-d") || !strcmp(a, "--debug")) {
      c->debug = true;
    } else if (!strcmp(a, "-f") || !strcmp(a, "--format")) {
      if ((v = value(i))) c->data_set_format = atoi(v);
    } else if (!strcmp(a, "-s") || !strcmp(a, "--solution")) {
      if ((v = value(i))) c->solution = atoi(v);
    } else if (!strcmp(a, "-e") || !strcmp(a, "--device")) {
      if ((v = value(i))) c->device_num = atoi(v);
    } else if (!strcmp(a, "-i") || !strcmp(a, "--import")) {
      if ((v = value(i))) c->graph_file = v;
      else fprintf(stderr, "Graph file name missing.\n");
    } else if (!strcmp(a, "--tune")) {
      if ((v = value(i))) c->tune_cache = v;
    } else if (!strcmp(a, "--threshold")) {
      if ((v = value(i))) c->threshold = atoi(v);
    } else if (!strcmp(a, "--damping")) {
      if ((v = value(i))) c->damping = atof(v);
    } else if (!strcmp(a, "--total-rank")) {
      if ((v = value(i))) c->total_rank = atof(v);
    } else if (!strcmp(a, "--tol")) {
      if ((v = value(i))) c->tol = atof(v);
    } else if (!strcmp(a, "--max-iter")) {
      if ((v = value(i))) c->max_iter = atoi(v);
    } else if (!strcmp(a, "--init-zero")) {
      c->init_zero = 1;
    } else if (!strcmp(a, "--prio-seed")) {
      if ((v = value(i))) c->prio_seed = atoll(v);
    }
  }
  return 1;
}

void cli_print_conf(const CliConf& c) {
  // apps/sssp/src/sssp.cpp:40-49
  fprintf(stderr, "\nCONFIGURATION:\n");
  if (c.graph_file) {
    fprintf(stderr, "- Graph file: %s\n", c.graph_file);
    fprintf(stderr, "- Graph format: %d\n", c.data_set_format);
  }
  fprintf(stderr, "- Solution: %d\n", c.solution);
  if (c.verbose && c.debug) fprintf(stderr, "- verbose mode\n");
  if (c.debug) fprintf(stderr, "- debug mode\n");
}

int cli_fail(const char* what, int rc) {
  fprintf(stderr, "Error: %s: %s", what, smdp_strerror(rc));
  const char* h = smdp_last_hip_error();
  if (rc == SMDP_ERR_DEVICE && h && *h) fprintf(stderr, " (%s)", h);
  fprintf(stderr, "\n");
  return 1;
}

int cli_load(const CliConf& c, smdp_hgraph** g) {
  *g = nullptr;
  FILE* f = fopen(c.graph_file, "r");
  if (!f) {
    fprintf(stderr, "Error: NULL file pointer.\n");
    return 1;
  }
  fclose(f);
  double t0 = cli_now();
  int rc = smdp_read_graph(c.graph_file, c.data_set_format, g);
  if (rc == SMDP_ERR_FORMAT) {
    fprintf(stderr, "Wrong code for dataset\n");
    return 1;
  }
  if (rc != SMDP_OK) return cli_fail("reading the graph", rc);
  if (c.verbose) {
    fprintf(stderr, "Source node is: %d\n", (*g)->source + 1);
    fprintf(stderr, "Read data and AdjList to CSR:\t\t%lf\n", cli_now() - t0);
    fprintf(stderr, "Node no: %d\n", (*g)->n);
    fprintf(stderr, "Edge no: %d\n", (*g)->m);
  }
  return 0;
}

int cli_context(const CliConf& c, smdp_ctx** ctx) {
  int rc = smdp_create(c.device_num, ctx);
  if (rc != SMDP_OK) return cli_fail("selecting the device", rc);
  if (c.threshold > 0) {
    for (int app = 0; app < SMDP_APP_COUNT; ++app) {
      smdp_params p;
      smdp_get_params(*ctx, app, &p);
      p.heavy_threshold = c.threshold;
      smdp_set_params(*ctx, app, &p);
    }
  }
  if (c.debug) {
    char name[256], arch[64];
    int32_t cus = 0;
    smdp_device_info(*ctx, name, sizeof name, arch, sizeof arch, &cus);
    fprintf(stderr, "Choose HIP device: %d (%s, %s, %d CUs)\n", c.device_num, name, arch, cus);
  }
  return 0;
}

int cli_apply_tuning(const CliConf& c, smdp_ctx* ctx, const smdp_csr* csr, int app, int32_t src) {
  if (!c.tune_cache || c.solution != SMDP_SOL_TUNED) return 0;
  smdp_graph* g = nullptr;
  int rc = smdp_graph_create(ctx, csr, &g);
  if (rc != SMDP_OK) return cli_fail("uploading the graph for tuning", rc);
  smdp_params chosen;
  rc = smdp_autotune(g, app, src, c.tune_cache, &chosen);
  smdp_graph_destroy(g);
  if (rc != SMDP_OK) return cli_fail("autotuning", rc);
  if (c.verbose)
    fprintf(stderr, "Tuned: split=%d chunk=%d wg_per_cu=%d heavy=%d\n", chosen.split_threshold, chosen.chunk,
            chosen.wg_per_cu, chosen.heavy_threshold);
  return 0;
}

void cli_print_counters(FILE* f, const smdp_stats& st) {
  fprintf(f, "Heavy work queued: wave %llu  workgroup %llu  grid %llu  (next-frontier pushes %llu)\n",
          (unsigned long long)st.q_push[1], (unsigned long long)st.q_push[2], (unsigned long long)st.q_push[3],
          (unsigned long long)st.q_push[0]);
  fprintf(f, "Kernel launches: %llu  atomics tried: %llu  won: %llu\n", (unsigned long long)st.kernel_launches,
          (unsigned long long)st.atomics_tried, (unsigned long long)st.atomics_won);
}

void output_cost(FILE* f, int32_t source, int32_t n, const int32_t* cost) {
  for (int32_t i = 0; i < n; ++i) fprintf(f, "From Node %d -> %d is %d\n", source + 1, i + 1, cost[i]);
}

void output_rank(FILE* f, int32_t n, const float* rank) {
  fprintf(stderr, "------- Rank value ------\n");
  for (int32_t i = 0; i < n; ++i) fprintf(f, "Node %d: %0.6lf\n", i + 1, (double)rank[i]);
}

void output_color(FILE* f, int32_t n, const int32_t* color) {
  fprintf(stderr, "------- Color value ------\n");
  for (int32_t i = 0; i < n; ++i) fprintf(f, "Node %d: %d\n", i + 1, color[i]);
}

void output_level(FILE* f, int32_t n, const int32_t* level) {
  for (int32_t i = 0; i < n; ++i) fprintf(f, "Node %d is in level %d\n", i + 1, level[i]);
}

void bfs_cpu(const smdp_hgraph* g, int32_t source, int32_t* level) {
  // FIFO traversal re-inserting a vertex whenever its level improves.
  for (int32_t i = 0; i < g->n; ++i) level[i] = SMDP_BFS_UNDEF;
  level[source] = 0;
  std::deque<int32_t> work;
  work.push_back(source);
  while (!work.empty()) {
    int32_t node = work.front();
    work.pop_front();
    int32_t next = level[node] + 1;
    for (int32_t e = g->row[node]; e < g->row[node + 1]; ++e) {
      int32_t nb = g->col[e];
      if (level[nb] == SMDP_BFS_UNDEF || level[nb] > next) {
        level[nb] = next;
        work.push_back(nb);
      }
    }
  }
}

int validate_arrays(int32_t n, const int32_t* a, const int32_t* b, const char* message) {
  for (int32_t i = 0; i < n; ++i) {
    if (a[i] != b[i]) {
      printf("ERROR: validation error at %d: %s !\n", i, message);
      printf("Node %d : %d v.s. %d\n", i, a[i], b[i]);
      return 0;
    }
  }
  printf("PASS: %s !\n", message);
  return 1;
}
