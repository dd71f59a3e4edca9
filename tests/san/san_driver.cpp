// Sanitizer driver for the host-side code: the graph readers and writer
// (sequential and threaded DIMACS9 paths, DIMACS10, SLNDC, malformed files),
// the seeded generators, inversion, the pinned-upload thread team run against
// a simulated copy engine, and (with -DSAN_ORACLE) the oracle's sequential
// and OpenMP functions.  Built with AddressSanitizer + UndefinedBehavior-
// Sanitizer and, without the oracle, with ThreadSanitizer (Makefile beside
// this file); exits 0 and prints "san: ok" when every check held.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "smdp.h"
#include "stage_team.h"

#ifdef SAN_ORACLE
extern "C" {
int orc_dijkstra(int n, const int* row, const int* col, const int* w, int src, int* dist);
int orc_bellman_ford(int n, const int* row, const int* col, const int* w, int src, int* dist);
long long orc_sssp_worklist(int n, const int* row, const int* col, const int* w, int src, int* cost,
                            long long* edges);
long long orc_sssp_worklist_par(int n, const int* row, const int* col, const int* w, int src, int* cost,
                                long long* edges, int* threads);
int orc_bfs_fifo(int n, const int* row, const int* col, int src, int* level);
int orc_bfs_par(int n, const int* row, const int* col, int src, int* level);
int orc_tree(int n, const int* row, const int* col, int root, int* height, int* desc);
int orc_color_rounds(int n, const int* row, const int* col, const unsigned* prio, int* color);
int orc_color_closed(int n, const int* row, const int* col, const unsigned* prio, int* color);
int orc_color_layers_par(int n, const int* row, const int* col, const unsigned* prio, int* color, int* ncolors);
int orc_spmv_f64(int n, const int* row, const int* col, const float* val, const float* x, double* y);
int orc_spmv_f64_par(int n, const int* row, const int* col, const float* val, const float* x, double* y);
int orc_pagerank_f64(int n, const int* in_row, const int* in_col, const int* outdeg, double d, double T,
                     int init_zero, double tol, int max_iter, int fixed_iters, double* rank, int* iters,
                     double* residual);
int orc_pagerank_f64_par(int n, const int* in_row, const int* in_col, const int* outdeg, double d, double T,
                         int init_zero, int fixed_iters, double* rank, double* residual);
}
#endif

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      fprintf(stderr, "san: check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                  \
    }                                                              \
  } while (0)

static std::string tmpdir;

static std::string write_text(const char* name, const std::string& text) {
  std::string p = tmpdir + "/" + name;
  FILE* f = fopen(p.c_str(), "w");
  fwrite(text.data(), 1, text.size(), f);
  fclose(f);
  return p;
}

static bool same(const smdp_hgraph* a, const smdp_hgraph* b) {
  if (a->n != b->n || a->m != b->m || a->source != b->source) return false;
  if (memcmp(a->row, b->row, sizeof(int32_t) * (a->n + 1))) return false;
  if (a->m && memcmp(a->col, b->col, sizeof(int32_t) * a->m)) return false;
  if ((a->w == nullptr) != (b->w == nullptr)) return false;
  return !a->w || !a->m || !memcmp(a->w, b->w, sizeof(int32_t) * a->m);
}

// Every row of a valid CSR is in range and the offsets never fall.
static bool valid_csr(const smdp_hgraph* g) {
  if (g->row[0] != 0 || g->row[g->n] != g->m) return false;
  for (int32_t v = 0; v < g->n; ++v)
    if (g->row[v + 1] < g->row[v]) return false;
  for (int32_t e = 0; e < g->m; ++e)
    if (g->col[e] < 0 || g->col[e] >= g->n) return false;
  return true;
}

// write -> read back: a file above a few MB goes through the threaded reader
static void roundtrip(const smdp_hgraph* g, const char* name) {
  std::string p = tmpdir + "/" + name;
  CHECK(smdp_write_dimacs9(p.c_str(), g) == SMDP_OK);
  smdp_hgraph* r = nullptr;
  CHECK(smdp_read_graph(p.c_str(), SMDP_FMT_DIMACS9, &r) == SMDP_OK);
  // the reader builds rows in reverse file order: write + read twice is the identity
  if (r) {
    CHECK(valid_csr(r));
    std::string p2 = p + "2";
    CHECK(smdp_write_dimacs9(p2.c_str(), r) == SMDP_OK);
    smdp_hgraph* r2 = nullptr;
    CHECK(smdp_read_graph(p2.c_str(), SMDP_FMT_DIMACS9, &r2) == SMDP_OK);
    if (r2) CHECK(same(g, r2));
    smdp_hgraph_free(r2);
    unlink(p2.c_str());
  }
  smdp_hgraph_free(r);
  unlink(p.c_str());
}

static void readers() {
  // malformed files return an error and leave nothing behind
  const char* bad9[] = {
      "",                                        // empty
      "p sp 3 1\na 1 9 5\n",                     // head out of range
      "p sp 3 2\na 1 2 5\n",                     // fewer arcs than declared
      "p sp 3 1\na 1\n",                         // truncated arc line
      "p sp 3 1\ns 7\na 1 2 5\n",                // source out of range
      "p\n",                                     // truncated problem line
  };
  for (const char* t : bad9) {
    smdp_hgraph* g = nullptr;
    CHECK(smdp_read_graph(write_text("bad.gr", t).c_str(), SMDP_FMT_DIMACS9, &g) != SMDP_OK);
    CHECK(g == nullptr);
  }
  smdp_hgraph* g = nullptr;
  CHECK(smdp_read_graph((tmpdir + "/missing.gr").c_str(), SMDP_FMT_DIMACS9, &g) == SMDP_ERR_IO);
  // well-formed small files of all three formats
  CHECK(smdp_read_graph(write_text("ok.gr", "c x\np sp 4 3\ns 2\na 1 2 7\na 2 3 1\na 2 4 9\n").c_str(),
                        SMDP_FMT_DIMACS9, &g) == SMDP_OK);
  if (g) CHECK(g->n == 4 && g->m == 3 && g->source == 1 && valid_csr(g));
  smdp_hgraph_free(g);
  g = nullptr;
  CHECK(smdp_read_graph(write_text("ok.graph", "% c\n3 2\n2 3\n1\n1\n").c_str(), SMDP_FMT_DIMACS10, &g) ==
        SMDP_OK);
  if (g) CHECK(g->n == 3 && g->m == 4 && valid_csr(g));
  smdp_hgraph_free(g);
  g = nullptr;
  CHECK(smdp_read_graph(write_text("bad.graph", "3 2\n2 9\n1\n1\n").c_str(), SMDP_FMT_DIMACS10, &g) != SMDP_OK);
  CHECK(smdp_read_graph(
            write_text("ok.txt", "# a\n# b\n# Nodes: 5 Edges: 4\n# FromNodeId ToNodeId\n0 3\n3 7\n7 0\n3 0\n").c_str(),
            SMDP_FMT_SLNDC, &g) == SMDP_OK);
  if (g) CHECK(g->n == 3 && g->m == 4 && valid_csr(g));
  smdp_hgraph_free(g);
  g = nullptr;
  CHECK(smdp_read_graph(write_text("bad.txt", "# a\n# b\n# Nodes: 2 Edges: 1\n# x\n0 9\n").c_str(), SMDP_FMT_SLNDC,
                        &g) != SMDP_OK);
  CHECK(smdp_read_graph(write_text("bad2.txt", "# a\n").c_str(), SMDP_FMT_SLNDC, &g) != SMDP_OK);
}

static void generators_and_roundtrips() {
  smdp_hgraph* g = nullptr;
  // 65536 x 8 arcs: an 8 MB file, read by the threaded DIMACS9 reader
  CHECK(smdp_gen_uniform(65536, 8, 1, 100, 1, &g) == SMDP_OK);
  if (g) {
    CHECK(valid_csr(g));
    roundtrip(g, "uniform.gr");
    smdp_hgraph* inv = nullptr;
    CHECK(smdp_invert(g, &inv) == SMDP_OK);
    if (inv) {
      CHECK(valid_csr(inv) && inv->m == g->m);
      smdp_hgraph* back = nullptr;
      CHECK(smdp_invert(inv, &back) == SMDP_OK);
      if (back) CHECK(valid_csr(back) && back->m == g->m);
      smdp_hgraph_free(back);
    }
    smdp_hgraph_free(inv);
    std::vector<int32_t> od(g->n), srcs(64);
    CHECK(smdp_outdegree(g, od.data()) == SMDP_OK);
    CHECK(smdp_gen_sources(g, 64, 9, srcs.data()) == SMDP_OK);
    for (int32_t s : srcs) CHECK(s >= 0 && s < g->n && od[s] > 0);
  }
  smdp_hgraph_free(g);
  g = nullptr;
  for (int sym = 0; sym < 2; ++sym) {
    CHECK(smdp_gen_rmat(14, 16, 0.57, 0.19, 0.19, 1, 255, sym, 1, 3, &g) == SMDP_OK);
    if (g) {
      CHECK(valid_csr(g));
      roundtrip(g, "rmat.gr");
    }
    smdp_hgraph_free(g);
    g = nullptr;
  }
  CHECK(smdp_gen_rmat(12, 8, 0.57, 0.19, 0.19, 0, 0, 0, 0, 2, &g) == SMDP_OK);
  if (g) CHECK(valid_csr(g) && g->w == nullptr);
  smdp_hgraph_free(g);
  g = nullptr;
  CHECK(smdp_gen_zipf_tree(1 << 16, 1.5, 4096, 5, &g) == SMDP_OK);
  if (g) CHECK(valid_csr(g) && g->m == g->n - 1);
  smdp_hgraph_free(g);
  g = nullptr;
  for (int i = 0; i < 8; ++i) {
    CHECK(smdp_gen_small_graph(i, 1024, 65536, 4.0, 32.0, 1, 100, 4, &g) == SMDP_OK);
    if (g) CHECK(valid_csr(g) && g->n >= 1024 && g->n <= 65536);
    smdp_hgraph_free(g);
    g = nullptr;
  }
  std::vector<float> f(100003);
  CHECK(smdp_gen_uniform_f32((int64_t)f.size(), -1.0f, 1.0f, 2, 6, f.data()) == SMDP_OK);
  for (float x : f) CHECK(x >= -1.0f && x < 1.0f);
  std::vector<uint32_t> pr(70001);
  CHECK(smdp_gen_prio((int32_t)pr.size(), 5, pr.data()) == SMDP_OK);
  std::vector<float> data(1000), x(200);
  CHECK(smdp_gen_spmv_reference_inputs(200, 1000, data.data(), x.data()) == SMDP_OK);
  CHECK(smdp_gen_uniform(0, 8, 1, 100, 1, &g) != SMDP_OK || (g && g->n == 0));
  smdp_hgraph_free(g);
}

// The upload team against a simulated copy engine: every issued copy runs on
// a thread of its own that reads the staging buffer while the team fills the
// other one; wait() joins it.  A team that refilled a buffer still being read
// would race with that thread.
static void staging_team() {
  const int64_t chunk = 4099;
  std::vector<int32_t> a(chunk), b(chunk);
  int32_t* stage[2] = {a.data(), b.data()};
  for (int nt : {1, 3, 8}) {
    for (int64_t count : {int64_t(0), int64_t(1), chunk - 1, chunk, 2 * chunk, 5 * chunk + 3}) {
      std::vector<int32_t> src((size_t)count), dst((size_t)count, -1);
      for (int64_t i = 0; i < count; ++i) src[(size_t)i] = (int32_t)((i * 2654435761u) % 1000);
      std::thread engine[2];
      std::atomic<bool> bad{false};
      bool ok = smdp::stage_team(
          nt, src.data(), count, chunk, stage, 1000, &bad,
          [&](int buf) {
            engine[buf].join();
            return true;
          },
          [&](int64_t off, int buf, int64_t len) {
            if (engine[buf].joinable()) engine[buf].join();
            engine[buf] = std::thread([&, off, buf, len] {
              memcpy(dst.data() + off, stage[buf], sizeof(int32_t) * (size_t)len);
            });
            return true;
          });
      for (auto& e : engine)
        if (e.joinable()) e.join();
      CHECK(ok && !bad.load());
      CHECK(src == dst);
      // a value at the bound is reported
      if (count) {
        src[(size_t)(count / 2)] = 1000;
        std::atomic<bool> bad2{false};
        smdp::stage_team(
            nt, src.data(), count, chunk, stage, 1000, &bad2, [](int) { return true; },
            [](int64_t, int, int64_t) { return true; });
        CHECK(bad2.load());
      }
    }
  }
  // a failing copy is reported and the team still ends
  std::vector<int32_t> src(3 * chunk, 1);
  std::atomic<bool> bad{false};
  CHECK(!smdp::stage_team(
      4, src.data(), (int64_t)src.size(), chunk, stage, -1, &bad, [](int) { return true; },
      [](int64_t off, int, int64_t) { return off == 0; }));
}

#ifdef SAN_ORACLE
static void oracle() {
  smdp_hgraph* g = nullptr;
  CHECK(smdp_gen_rmat(13, 16, 0.57, 0.19, 0.19, 1, 255, 0, 1, 3, &g) == SMDP_OK);
  if (!g) return;
  const int n = g->n, src = g->source;
  std::vector<int> d1(n), d2(n), d3(n), d4(n), l1(n), l2(n);
  long long e1 = 0, e2 = 0;
  int threads = 0;
  CHECK(orc_dijkstra(n, g->row, g->col, g->w, src, d1.data()) == 0);
  CHECK(orc_bellman_ford(n, g->row, g->col, g->w, src, d2.data()) >= 0);
  CHECK(orc_sssp_worklist(n, g->row, g->col, g->w, src, d3.data(), &e1) >= 0);
  CHECK(orc_sssp_worklist_par(n, g->row, g->col, g->w, src, d4.data(), &e2, &threads) >= 0);
  // the threaded rounds may scan other arcs (update order), the distances agree
  CHECK(d1 == d2 && d1 == d3 && d1 == d4 && e1 > 0 && e2 > 0);
  CHECK(orc_bfs_fifo(n, g->row, g->col, src, l1.data()) >= 0);
  CHECK(orc_bfs_par(n, g->row, g->col, src, l2.data()) > 0);
  CHECK(l1 == l2);
  std::vector<float> val(g->m), x(n);
  smdp_gen_uniform_f32(g->m, -1.0f, 1.0f, 2, 6, val.data());
  smdp_gen_uniform_f32(n, 0.0f, 1.0f, 2, 9, x.data());
  std::vector<double> y1(n), y2(n);
  CHECK(orc_spmv_f64(n, g->row, g->col, val.data(), x.data(), y1.data()) == 0);
  CHECK(orc_spmv_f64_par(n, g->row, g->col, val.data(), x.data(), y2.data()) > 0);
  CHECK(y1 == y2);
  smdp_hgraph* ig = nullptr;
  CHECK(smdp_invert(g, &ig) == SMDP_OK);
  std::vector<int32_t> od(n);
  smdp_outdegree(g, od.data());
  std::vector<double> r1(n), r2(n);
  int it = 0;
  double res1 = 0, res2 = 0;
  CHECK(orc_pagerank_f64(n, ig->row, ig->col, od.data(), 0.85, 1.0, 0, 0.0, 0, 5, r1.data(), &it, &res1) == 0);
  CHECK(orc_pagerank_f64_par(n, ig->row, ig->col, od.data(), 0.85, 1.0, 0, 5, r2.data(), &res2) > 0);
  double diff = 0;
  for (int v = 0; v < n; ++v) diff += fabs(r1[v] - r2[v]);
  CHECK(it == 5 && diff < 1e-12);
  smdp_hgraph_free(ig);
  smdp_hgraph_free(g);
  g = nullptr;
  CHECK(smdp_gen_rmat(12, 8, 0.57, 0.19, 0.19, 0, 0, 1, 1, 5, &g) == SMDP_OK);
  if (g) {
    std::vector<uint32_t> pr(g->n);
    smdp_gen_prio(g->n, 5, pr.data());
    std::vector<int> c1(g->n), c2(g->n), c3(g->n);
    int k = 0;
    CHECK(orc_color_rounds(g->n, g->row, g->col, pr.data(), c1.data()) >= 0);
    CHECK(orc_color_closed(g->n, g->row, g->col, pr.data(), c2.data()) >= 0);
    CHECK(orc_color_layers_par(g->n, g->row, g->col, pr.data(), c3.data(), &k) > 0);
    CHECK(c1 == c2 && c1 == c3);
  }
  smdp_hgraph_free(g);
  g = nullptr;
  CHECK(smdp_gen_zipf_tree(1 << 14, 1.5, 4096, 5, &g) == SMDP_OK);
  if (g) {
    std::vector<int> h(g->n), dsc(g->n);
    CHECK(orc_tree(g->n, g->row, g->col, 0, h.data(), dsc.data()) >= 0);
    CHECK(dsc[0] == g->n - 1);  // every other node descends from the root
  }
  smdp_hgraph_free(g);
}
#endif

int main() {
  char tmpl[] = "/tmp/smdp_san_XXXXXX";
  if (!mkdtemp(tmpl)) return 2;
  tmpdir = tmpl;
  readers();
  generators_and_roundtrips();
  staging_team();
#ifdef SAN_ORACLE
  oracle();
#endif
  for (const char* f : {"bad.gr", "ok.gr", "ok.graph", "bad.graph", "ok.txt", "bad.txt", "bad2.txt"})
    unlink((tmpdir + "/" + f).c_str());
  rmdir(tmpl);
  if (failures) {
    fprintf(stderr, "san: %d checks failed\n", failures);
    return 1;
  }
  printf("san: ok\n");
  return 0;
}
