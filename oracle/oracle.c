/*
 * oracle.c -- CPU restatement of the reference's algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libsmdp.so, the gpu-*
 * command-line drivers) links, loads or calls this file.  It is used by
 * tests/ as the checker, by __graft_entry__.smoke() as the checker, and by
 * bench.py's cpu_baseline leg as the timed CPU path.
 *
 * Each function names the reference code it restates.  How each one is
 * pinned (known answers, library cross-checks, invariants, brute force) is
 * listed in the tests that exercise it: tests/test_oracle.py.
 *
 * Sentinels: SSSP unreachable 1073741824 (common/global.h:19), BFS
 * unreachable 0x7FFFFFFF (apps/bfs-rec/src/bfs_rec.h:13).
 */
#include <math.h>
#include <stdint.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF 1073741824
#define ORC_UNDEF 0x7FFFFFFF

/* ---------------------------------------------------------------- SSSP ---*/

/* Textbook Dijkstra with a binary heap (lazy deletion).  Shortest-path
 * distances are unique, so this is an exact oracle for every relaxation order
 * the reference's worklist Bellman-Ford (sssp_kernel.cu:40-61) may take. */
int orc_dijkstra(int n, const int* row, const int* col, const int* w, int src, int* dist) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  for (int i = 0; i < n; ++i) dist[i] = ORC_INF;
  int64_t cap = (int64_t)row[n] + 1;
  int64_t* hk = (int64_t*)malloc(sizeof(int64_t) * cap); /* key: dist<<32 | vertex */
  if (!hk) return -2;
  int64_t hn = 0;
#define HPUSH(key)                                               \
  do {                                                           \
    int64_t _i = hn++;                                           \
    hk[_i] = (key);                                              \
    while (_i > 0 && hk[(_i - 1) / 2] > hk[_i]) {                \
      int64_t _p = (_i - 1) / 2, _t = hk[_p];                    \
      hk[_p] = hk[_i];                                           \
      hk[_i] = _t;                                               \
      _i = _p;                                                   \
    }                                                            \
  } while (0)
  dist[src] = 0;
  HPUSH((int64_t)src);
  while (hn > 0) {
    int64_t top = hk[0];
    hk[0] = hk[--hn];
    for (int64_t i = 0;;) {
      int64_t l = 2 * i + 1, r = l + 1, s = i;
      if (l < hn && hk[l] < hk[s]) s = l;
      if (r < hn && hk[r] < hk[s]) s = r;
      if (s == i) break;
      int64_t t = hk[s];
      hk[s] = hk[i];
      hk[i] = t;
      i = s;
    }
    int u = (int)(top & 0xFFFFFFFF);
    int du = (int)(top >> 32);
    if (du != dist[u]) continue;
    for (int e = row[u]; e < row[u + 1]; ++e) {
      int v = col[e];
      int alt = du + w[e];
      if (alt < dist[v]) {
        dist[v] = alt;
        HPUSH(((int64_t)alt << 32) | (int64_t)v);
      }
    }
  }
#undef HPUSH
  free(hk);
  return 0;
}

/* Brute-force Bellman-Ford: full passes over every arc until nothing changes.
 * For tiny graphs only (pins Dijkstra by a different algorithm). */
int orc_bellman_ford(int n, const int* row, const int* col, const int* w, int src, int* dist) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  for (int i = 0; i < n; ++i) dist[i] = ORC_INF;
  dist[src] = 0;
  for (int pass = 0; pass < n; ++pass) {
    int changed = 0;
    for (int u = 0; u < n; ++u) {
      if (dist[u] == ORC_INF) continue;
      for (int e = row[u]; e < row[u + 1]; ++e)
        if (dist[u] + w[e] < dist[col[e]]) {
          dist[col[e]] = dist[u] + w[e];
          changed = 1;
        }
    }
    if (!changed) break;
  }
  return 0;
}

/* The reference's SSSP, restated sequentially: rounds of
 * "queue := {v : update[v]}, clear update" (unorder_generateQueue_kernel,
 * sssp_kernel.cu:496-506) followed by "for each queued u and arc (u,v,w):
 * alt = cost[u]+w; if cost[v] > alt then cost[v] = alt, update[v] = 1"
 * (unorder_threadQueue_kernel, :40-61), starting from cost = INF,
 * cost[src] = 0, update[src] = 1 (sssp_wrapper.cu:88-92).  Returns the round
 * count; *edges receives the arcs scanned.  This is the timed CPU baseline. */
long long orc_sssp_worklist(int n, const int* row, const int* col, const int* w, int src, int* cost,
                            long long* edges) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  char* update = (char*)calloc((size_t)n, 1);
  int* queue = (int*)malloc(sizeof(int) * (size_t)n);
  if (!update || !queue) {
    free(update);
    free(queue);
    return -2;
  }
  for (int i = 0; i < n; ++i) cost[i] = ORC_INF;
  cost[src] = 0;
  update[src] = 1;
  long long rounds = 0, scanned = 0;
  for (;;) {
    int qn = 0;
    for (int v = 0; v < n; ++v)
      if (update[v]) {
        update[v] = 0;
        queue[qn++] = v;
      }
    if (qn == 0) break;
    ++rounds;
    for (int i = 0; i < qn; ++i) {
      int u = queue[i];
      int cu = cost[u];
      for (int e = row[u]; e < row[u + 1]; ++e) {
        int v = col[e];
        int alt = cu + w[e];
        if (cost[v] > alt) {
          cost[v] = alt;
          update[v] = 1;
        }
      }
      scanned += row[u + 1] - row[u];
    }
  }
  free(update);
  free(queue);
  if (edges) *edges = scanned;
  return rounds;
}

/* The same rounds on all host threads (OpenMP): the flag scan and the arc
 * loop are split over threads, and "if cost[v] > alt then cost[v] = alt" is
 * an atomic minimum, so the distances are those of the sequential version.
 * Timed as the multi-core CPU baseline.  *threads receives the thread count. */
long long orc_sssp_worklist_par(int n, const int* row, const int* col, const int* w, int src, int* cost,
                                long long* edges, int* threads) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  char* update = (char*)calloc((size_t)n, 1);
  int* queue = (int*)malloc(sizeof(int) * (size_t)n);
  if (!update || !queue) {
    free(update);
    free(queue);
    return -2;
  }
  for (int i = 0; i < n; ++i) cost[i] = ORC_INF;
  cost[src] = 0;
  update[src] = 1;
  long long rounds = 0, scanned = 0;
  int nt = 1;
#pragma omp parallel
  {
#pragma omp single
    nt = omp_get_num_threads();
  }
  for (;;) {
    int qn = 0;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < n; ++v)
      if (update[v]) {
        update[v] = 0;
        queue[__atomic_fetch_add(&qn, 1, __ATOMIC_RELAXED)] = v;
      }
    if (qn == 0) break;
    ++rounds;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : scanned)
    for (int i = 0; i < qn; ++i) {
      int u = queue[i];
      int cu = __atomic_load_n(&cost[u], __ATOMIC_RELAXED);
      for (int e = row[u]; e < row[u + 1]; ++e) {
        int v = col[e];
        int alt = cu + w[e];
        int old = __atomic_load_n(&cost[v], __ATOMIC_RELAXED);
        while (old > alt && !__atomic_compare_exchange_n(&cost[v], &old, alt, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        if (old > alt) __atomic_store_n(&update[v], 1, __ATOMIC_RELAXED);
      }
      scanned += row[u + 1] - row[u];
    }
  }
  free(update);
  free(queue);
  if (edges) *edges = scanned;
  if (threads) *threads = nt;
  return rounds;
}

/* ----------------------------------------------------------------- BFS ---*/

/* bfs_cpu (apps/bfs-rec/src/bfs_rec.cpp:102-121): FIFO traversal that
 * re-inserts a vertex whenever its level improves. */
int orc_bfs_fifo(int n, const int* row, const int* col, int src, int* level) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  for (int i = 0; i < n; ++i) level[i] = ORC_UNDEF;
  level[src] = 0;
  /* A vertex can be re-queued only when its level improves; with FIFO order
   * on unit weights each vertex is queued once, so n slots suffice. */
  int64_t cap = (int64_t)n + (int64_t)row[n] + 1;
  int* q = (int*)malloc(sizeof(int) * (size_t)cap);
  if (!q) return -2;
  int64_t head = 0, tail = 0;
  q[tail++] = src;
  while (head < tail) {
    int node = q[head++];
    int next = level[node] + 1;
    for (int e = row[node]; e < row[node + 1]; ++e) {
      int nb = col[e];
      if (level[nb] == ORC_UNDEF || level[nb] > next) {
        level[nb] = next;
        if (tail < cap) q[tail++] = nb;
      }
    }
  }
  free(q);
  return 0;
}

/* ---------------------------------------------------------------- tree ---*/

/* Tree Height and Tree Descendant, the two recursive benchmarks that the
 * reference's README.md:12-13 names but does not ship.  The input is a tree
 * given as child lists (row/col) and its root; it is rejected (-3) when some
 * vertex has two or more in-arcs (counting every arc of the CSR).  Arcs into
 * the root are ignored (a cycle through the root).  For every vertex reached
 * from the root: height = 0 for a leaf, else 1 + the largest child height;
 * desc = the number of its descendants.  Unreached vertices get -1 in both.
 * Sequential: the vertices in BFS order from the root, then that order
 * backwards, each summing over its children.  Returns height[root]. */
int orc_tree(int n, const int* row, const int* col, int root, int* height, int* desc) {
  if (n <= 0 || root < 0 || root >= n) return -1;
  int* indeg = (int*)calloc((size_t)n, sizeof(int));
  int* order = (int*)malloc(sizeof(int) * (size_t)n);
  if (!indeg || !order) {
    free(indeg);
    free(order);
    return -2;
  }
  int bad = 0;
  for (int e = 0; e < row[n] && !bad; ++e)
    if (++indeg[col[e]] > 1) bad = 1;
  if (bad) {
    free(indeg);
    free(order);
    return -3;
  }
  for (int i = 0; i < n; ++i) height[i] = desc[i] = -1;
  int64_t head = 0, tail = 0;
  order[tail++] = root;
  height[root] = 0;  /* marks "reached" until the backward pass */
  while (head < tail) {
    const int u = order[head++];
    for (int e = row[u]; e < row[u + 1]; ++e)
      if (col[e] != root) {  /* in-degree <= 1: every other child is new */
        height[col[e]] = 0;
        order[tail++] = col[e];
      }
  }
  for (int64_t i = tail - 1; i >= 0; --i) {
    const int u = order[i];
    int h = 0, d = 0;
    for (int e = row[u]; e < row[u + 1]; ++e) {
      const int c = col[e];
      if (c == root) continue;
      if (1 + height[c] > h) h = 1 + height[c];
      d += 1 + desc[c];
    }
    height[u] = h;
    desc[u] = d;
  }
  const int hr = height[root];
  free(indeg);
  free(order);
  return hr;
}

/* ------------------------------------------------------------ coloring ---*/

/* u precedes v: (prio[u], u) < (prio[v], v); prio == NULL is the reference's
 * vertex-ID priority (nid < tid, graph_color_kernel.cu:40). */
static int precedes(const unsigned* prio, int u, int v) {
  if (!prio) return u < v;
  if (prio[u] != prio[v]) return prio[u] < prio[v];
  return u < v;
}

/* gclr_bitmap_kernel (graph_color_kernel.cu:28-49) driven by
 * gclr_nopruning_gpu (graph_color_wrapper.cu:158-174): in round c = 1, 2, ...
 * an uncoloured v takes colour c unless some u in row(v) that precedes v has
 * colour 0 or c.  Sequential sweep: a u coloured earlier in the same sweep
 * reads as c, one coloured later still reads 0, so both block -- exactly the
 * two values the concurrent kernel can observe.  Returns the number of
 * rounds that coloured something. */
int orc_color_rounds(int n, const int* row, const int* col, const unsigned* prio, int* color) {
  for (int i = 0; i < n; ++i) color[i] = 0;
  int c = 1, left = n;
  while (left > 0) {
    int colored = 0;
    for (int v = 0; v < n; ++v) {
      if (color[v] != 0) continue;
      int ok = 1;
      for (int e = row[v]; e < row[v + 1] && ok; ++e) {
        int u = col[e];
        if (precedes(prio, u, v) && (color[u] == 0 || color[u] == c)) ok = 0;
      }
      if (ok) {
        color[v] = c;
        ++colored;
      }
    }
    if (colored == 0) return -1; /* cannot happen: the least uncoloured vertex always colours */
    left -= colored;
    ++c;
  }
  return c - 1;
}

static const unsigned* g_sort_prio;
static int cmp_prec(const void* a, const void* b) {
  int u = *(const int*)a, v = *(const int*)b;
  if (u == v) return 0;
  return precedes(g_sort_prio, u, v) ? -1 : 1;
}

/* Closed form of the same rule: color(v) = 1 + max{ color(u) : u in row(v),
 * u precedes v } (empty max = 0), evaluated in one pass in precedence order.
 * Independent of the round structure; pinned against orc_color_rounds. */
int orc_color_closed(int n, const int* row, const int* col, const unsigned* prio, int* color) {
  int* order = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!order) return -2;
  for (int i = 0; i < n; ++i) order[i] = i;
  g_sort_prio = prio;
  qsort(order, (size_t)n, sizeof(int), cmp_prec);
  int maxc = 0;
  for (int k = 0; k < n; ++k) {
    int v = order[k], mx = 0;
    for (int e = row[v]; e < row[v + 1]; ++e) {
      int u = col[e];
      if (precedes(prio, u, v) && color[u] > mx) mx = color[u];
    }
    color[v] = mx + 1;
    if (color[v] > maxc) maxc = color[v];
  }
  free(order);
  return maxc;
}

/* ---------------------------------------------------------------- SpMV ---*/

/* csr_spmv_kernel (SpMV_kernel.cu:81-96): y[r] = sum_i data[i]*x[col[i]] in
 * row order.  FP32 inputs are widened exactly; products of two floats are
 * exact in double and the sum is accumulated in long double. */
int orc_spmv_f64(int n, const int* row, const int* col, const float* val, const float* x, double* y) {
  for (int r = 0; r < n; ++r) {
    long double acc = 0.0L;
    for (int e = row[r]; e < row[r + 1]; ++e) acc += (long double)((double)val[e] * (double)x[col[e]]);
    y[r] = (double)acc;
  }
  return 0;
}

/* ------------------------------------------------------------ PageRank ---*/

/* pagerank_thread_pull_kernel (pagerank_kernel.cu:95-120) with the dangling
 * mass of update_danglingrankarray_kernel + the host sum
 * (pagerank_wrapper.cu:205-222), in FP64:
 *   D = sum_{outdeg[u]==0} r[u]
 *   r'[v] = sum_{u in in(v)} d*r[u]/outdeg[u] + (1-d)*T/N + d*D/N
 * Start from 0 (init_zero, pagerank_wrapper.cu:115) or T/N
 * (pagerank.cpp:103-106).  With fixed_iters > 0 runs exactly that many
 * iterations (iteration-matched oracle); otherwise stops after the first
 * iteration whose sum|r'-r| < tol*T, or at max_iter.  *iters and *residual
 * receive the iteration count and the last residual. */
int orc_pagerank_f64(int n, const int* in_row, const int* in_col, const int* outdeg, double d, double T,
                     int init_zero, double tol, int max_iter, int fixed_iters, double* rank, int* iters,
                     double* residual) {
  if (n <= 0) return -1;
  double* r = rank;
  double* nr = (double*)malloc(sizeof(double) * (size_t)n);
  double* contrib = (double*)malloc(sizeof(double) * (size_t)n);
  if (!nr || !contrib) {
    free(nr);
    free(contrib);
    return -2;
  }
  for (int i = 0; i < n; ++i) r[i] = init_zero ? 0.0 : T / n;
  int it = 0;
  double res = 0.0;
  int limit = fixed_iters > 0 ? fixed_iters : max_iter;
  while (it < limit) {
    double dang = 0.0;
    for (int u = 0; u < n; ++u) {
      if (outdeg[u] == 0) dang += r[u];
      contrib[u] = outdeg[u] ? r[u] / outdeg[u] : 0.0;
    }
    double base = (1.0 - d) * T / n + d * dang / n;
    res = 0.0;
    for (int v = 0; v < n; ++v) {
      double s = 0.0;
      for (int e = in_row[v]; e < in_row[v + 1]; ++e) s += contrib[in_col[e]];
      nr[v] = d * s + base;
      res += fabs(nr[v] - r[v]);
    }
    memcpy(r, nr, sizeof(double) * (size_t)n);
    ++it;
    if (fixed_iters <= 0 && res < tol * T) break;
  }
  free(nr);
  free(contrib);
  if (iters) *iters = it;
  if (residual) *residual = res;
  return 0;
}

/* ------------------------------------------- OpenMP CPU baselines ---------*/
/* Parallel twins of the functions above, used only to time the CPU baseline
 * on all host cores.  Each returns the number of threads it used (< 0 on
 * error) and is checked against its sequential twin in tests/test_oracle.py. */

/* orc_spmv_f64 with rows split over threads: every row is summed exactly as
 * in the sequential version, so y is bit-identical. */
int orc_spmv_f64_par(int n, const int* row, const int* col, const float* val, const float* x, double* y) {
  int nt = 1;
#pragma omp parallel
  {
#pragma omp single
    nt = omp_get_num_threads();
#pragma omp for schedule(dynamic, 4096)
    for (int r = 0; r < n; ++r) {
      long double acc = 0.0L;
      for (int e = row[r]; e < row[r + 1]; ++e) acc += (long double)((double)val[e] * (double)x[col[e]]);
      y[r] = (double)acc;
    }
  }
  return nt;
}

/* orc_pagerank_f64 with fixed_iters iterations, rows split over threads.  The
 * dangling mass and the residual are OpenMP reductions, so they may differ
 * from the sequential sums in the last bits. */
int orc_pagerank_f64_par(int n, const int* in_row, const int* in_col, const int* outdeg, double d, double T,
                         int init_zero, int fixed_iters, double* rank, double* residual) {
  if (n <= 0 || fixed_iters <= 0) return -1;
  double* r = rank;
  double* nr = (double*)malloc(sizeof(double) * (size_t)n);
  double* contrib = (double*)malloc(sizeof(double) * (size_t)n);
  if (!nr || !contrib) {
    free(nr);
    free(contrib);
    return -2;
  }
  int nt = 1;
  double res = 0.0;
#pragma omp parallel
  {
#pragma omp single
    nt = omp_get_num_threads();
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) r[i] = init_zero ? 0.0 : T / n;
  for (int it = 0; it < fixed_iters; ++it) {
    double dang = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : dang)
    for (int u = 0; u < n; ++u) {
      if (outdeg[u] == 0) dang += r[u];
      contrib[u] = outdeg[u] ? r[u] / outdeg[u] : 0.0;
    }
    double base = (1.0 - d) * T / n + d * dang / n;
    res = 0.0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : res)
    for (int v = 0; v < n; ++v) {
      double s = 0.0;
      for (int e = in_row[v]; e < in_row[v + 1]; ++e) s += contrib[in_col[e]];
      nr[v] = d * s + base;
      res += fabs(nr[v] - r[v]);
    }
    double* t = r;
    r = nr;
    nr = t;
  }
  if (r != rank) {
    memcpy(rank, r, sizeof(double) * (size_t)n);
    nr = r;
  }
  free(nr);
  free(contrib);
  if (residual) *residual = res;
  return nt;
}

/* Appends to a shared array through a per-thread buffer: one atomic reserve
 * per ORC_APPEND entries instead of one per entry. */
#define ORC_APPEND 1024
typedef struct {
  int buf[ORC_APPEND];
  int k;
} orc_appender;
static void orc_append_flush(orc_appender* a, int* out, int64_t* len) {
  if (a->k == 0) return;
  const int64_t at = __atomic_fetch_add(len, (int64_t)a->k, __ATOMIC_RELAXED);
  memcpy(out + at, a->buf, sizeof(int) * (size_t)a->k);
  a->k = 0;
}
static void orc_append(orc_appender* a, int v, int* out, int64_t* len) {
  a->buf[a->k++] = v;
  if (a->k == ORC_APPEND) orc_append_flush(a, out, len);
}

/* Level-synchronous BFS: the frontier of level k is expanded by all threads,
 * a vertex joins level k+1 when a compare-and-swap claims it.  Levels equal
 * orc_bfs_fifo's (shortest hop counts). */
int orc_bfs_par(int n, const int* row, const int* col, int src, int* level) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  int* cur = (int*)malloc(sizeof(int) * (size_t)n);
  int* nxt = (int*)malloc(sizeof(int) * (size_t)n);
  if (!cur || !nxt) {
    free(cur);
    free(nxt);
    return -2;
  }
  int nt = 1;
#pragma omp parallel
  {
#pragma omp single
    nt = omp_get_num_threads();
#pragma omp for schedule(static)
    for (int i = 0; i < n; ++i) level[i] = ORC_UNDEF;
  }
  level[src] = 0;
  cur[0] = src;
  int64_t nc = 1;
  for (int lv = 0; nc > 0; ++lv) {
    int64_t nn = 0;
#pragma omp parallel
    {
      orc_appender ap;
      ap.k = 0;
#pragma omp for schedule(dynamic, 64)
      for (int64_t i = 0; i < nc; ++i) {
        const int u = cur[i];
        for (int e = row[u]; e < row[u + 1]; ++e) {
          const int v = col[e];
          int exp = ORC_UNDEF;
          if (__atomic_load_n(&level[v], __ATOMIC_RELAXED) == ORC_UNDEF &&
              __atomic_compare_exchange_n(&level[v], &exp, lv + 1, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
            orc_append(&ap, v, nxt, &nn);
        }
      }
      orc_append_flush(&ap, nxt, &nn);
    }
    int* t = cur;
    cur = nxt;
    nxt = t;
    nc = nn;
  }
  free(cur);
  free(nxt);
  return nt;
}

/* orc_color_closed by layers (Kahn's algorithm): pend[v] counts the
 * preceding row neighbours of v; the vertices with none form layer 1, and
 * the vertices whose count reaches zero while layer c is expanded form layer
 * c+1.  Expanding v walks the transposed CSR (built here, in parallel where
 * it can be).  The layer index is 1 + the longest chain of preceding
 * neighbours, which is the closed form's colour.  Returns the thread count;
 * *ncolors receives the number of colours. */
int orc_color_layers_par(int n, const int* row, const int* col, const unsigned* prio, int* color, int* ncolors) {
  if (n <= 0) return -1;
  const int64_t m = row[n];
  int* tdeg = (int*)calloc((size_t)n + 1, sizeof(int));
  int* tcol = (int*)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
  int* pend = (int*)malloc(sizeof(int) * (size_t)n);
  int* cur = (int*)malloc(sizeof(int) * (size_t)n);
  int* nxt = (int*)malloc(sizeof(int) * (size_t)n);
  if (!tdeg || !tcol || !pend || !cur || !nxt) {
    free(tdeg);
    free(tcol);
    free(pend);
    free(cur);
    free(nxt);
    return -2;
  }
  int nt = 1;
#pragma omp parallel
  {
#pragma omp single
    nt = omp_get_num_threads();
  }
  /* transposed CSR: tcol lists, for each u, the rows v that contain u */
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; ++e) __atomic_fetch_add(&tdeg[col[e] + 1], 1, __ATOMIC_RELAXED);
  for (int v = 0; v < n; ++v) tdeg[v + 1] += tdeg[v];
  int* fill = (int*)malloc(sizeof(int) * (size_t)n);
  if (!fill) {
    free(tdeg);
    free(tcol);
    free(pend);
    free(cur);
    free(nxt);
    return -2;
  }
  memcpy(fill, tdeg, sizeof(int) * (size_t)n);
  /* order inside a transposed row does not matter to the layers */
#pragma omp parallel for schedule(dynamic, 4096)
  for (int v = 0; v < n; ++v)
    for (int e = row[v]; e < row[v + 1]; ++e) tcol[__atomic_fetch_add(&fill[col[e]], 1, __ATOMIC_RELAXED)] = v;
  free(fill);
  int64_t nc = 0;
#pragma omp parallel
  {
    orc_appender ap;
    ap.k = 0;
#pragma omp for schedule(dynamic, 4096)
    for (int v = 0; v < n; ++v) {
      int c = 0;
      for (int e = row[v]; e < row[v + 1]; ++e) c += col[e] != v && precedes(prio, col[e], v);
      pend[v] = c;
      color[v] = 0;
      if (c == 0) orc_append(&ap, v, cur, &nc);
    }
    orc_append_flush(&ap, cur, &nc);
  }
  int c = 0;
  while (nc > 0) {
    ++c;
    int64_t nn = 0;
#pragma omp parallel
    {
      orc_appender ap;
      ap.k = 0;
#pragma omp for schedule(dynamic, 64)
      for (int64_t i = 0; i < nc; ++i) {
        const int u = cur[i];
        color[u] = c;
        for (int e = tdeg[u]; e < tdeg[u + 1]; ++e) {
          const int v = tcol[e];
          if (v != u && precedes(prio, u, v) && __atomic_sub_fetch(&pend[v], 1, __ATOMIC_RELAXED) == 0)
            orc_append(&ap, v, nxt, &nn);
        }
      }
      orc_append_flush(&ap, nxt, &nn);
    }
    int* t = cur;
    cur = nxt;
    nxt = t;
    nc = nn;
  }
  free(tdeg);
  free(tcol);
  free(pend);
  free(cur);
  free(nxt);
  if (ncolors) *ncolors = c;
  return nt;
}
