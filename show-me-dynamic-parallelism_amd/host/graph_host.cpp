// Host-side graph library: the reference's input formats and CSR semantics,
// and the seeded synthetic generators.  No device code.
//
// Reference behaviour kept here (common/global.cpp):
//   readInputDIMACS9  :29-64   'p sp N M', 's src', 'a u v w' lines, 1-based.
//   readInputDIMACS10 :66-105  METIS adjacency lines; M doubled; weights
//                               rand()%1000+1 after srand(7), in file order.
//   readInputSLNDC    :107-190 SNAP edge list; IDs remapped to 0..n-1 in
//                               ascending order; weights drawn per remapped
//                               vertex in ascending order, file order inside.
//   convertCSR        :192-235 each row holds its arcs in REVERSE file order
//                               (it pops from the back of each list).
//   invertGraph       :237-277 in-rows in descending source order.
// Deviation: where the reference would index out of bounds or loop forever on
// malformed input (arc count != header count, vertex id out of range,
// unparsable pair), these functions return SMDP_ERR_IO instead.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/smdp.h"
#include "rng.h"

namespace {

using smdp::GlibcRand;
using smdp::rng_at;

// Host worker threads: the hardware count, at most 16, or SMDP_HOST_THREADS.
unsigned worker_count() {
  unsigned t = std::thread::hardware_concurrency();
  if (const char* e = getenv("SMDP_HOST_THREADS")) t = (unsigned)std::max(1, atoi(e));
  if (t == 0) t = 1;
  if (t > 16) t = 16;
  return t;
}

// Run fn(begin, end) over [0, n) split into contiguous slices, one per thread.
void parallel_slices(int64_t n, const std::function<void(int64_t, int64_t)>& fn) {
  unsigned t = worker_count();
  if (n < (int64_t)1 << 16 || t == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (unsigned k = 0; k < t; ++k) {
    int64_t b = n * k / t, e = n * (k + 1) / t;
    pool.emplace_back([=, &fn] { fn(b, e); });
  }
  for (auto& th : pool) th.join();
}

smdp_hgraph* hg_alloc(int32_t n, int32_t m, bool weights) {
  smdp_hgraph* g = (smdp_hgraph*)calloc(1, sizeof(smdp_hgraph));
  if (!g) return nullptr;
  g->n = n;
  g->m = m;
  g->row = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  g->col = (int32_t*)malloc(sizeof(int32_t) * ((size_t)m > 0 ? (size_t)m : 1));
  g->w = weights ? (int32_t*)malloc(sizeof(int32_t) * ((size_t)m > 0 ? (size_t)m : 1)) : nullptr;
  if (!g->row || !g->col || (weights && !g->w)) {
    smdp_hgraph_free(g);
    return nullptr;
  }
  return g;
}

// Arcs in file order -> CSR whose rows hold the arcs in reverse file order
// (convertCSR pops the back of each adjacency list, common/global.cpp:215-221).
// Counting sort of the arc list [0, m) by key(e) into rows of `row` (n + 1
// offsets, filled here), on T threads that each own a contiguous slice of
// the arcs.  place(e, p) stores arc e at position p.  Within a row the arcs
// keep their list order (reverse = false) or the reverse of it (reverse =
// true): thread t's arcs of row v take the block after (or before) those of
// threads t' < t, so the result equals the sequential fill exactly.
template <typename Key, typename Place>
void parallel_row_fill(int32_t n, int64_t m, int32_t* row, bool reverse, Key key, Place place) {
  // per-thread counts: T x n ints, T kept so that they stay under 1 GiB
  int64_t T = std::min<int64_t>(worker_count(), std::max<int64_t>(1, ((int64_t)1 << 28) / std::max(1, n)));
  if (m < ((int64_t)1 << 16)) T = 1;
  std::vector<int32_t> cnt((size_t)(T * n), 0);
  auto slice = [&](int64_t t, int64_t* b, int64_t* e) {
    *b = m * t / T;
    *e = m * (t + 1) / T;
  };
  auto run = [&](const std::function<void(int64_t)>& fn) {
    std::vector<std::thread> pool;
    for (int64_t t = 1; t < T; ++t) pool.emplace_back([=, &fn] { fn(t); });
    fn(0);
    for (auto& th : pool) th.join();
  };
  run([&](int64_t t) {
    int64_t b, e;
    slice(t, &b, &e);
    int32_t* c = cnt.data() + t * n;
    for (int64_t k = b; k < e; ++k) ++c[key(k)];
  });
  // rows (totals in parallel, prefix in order), then each thread's start
  // inside each row
  parallel_slices(n, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) {
      int32_t tot = 0;
      for (int64_t t = 0; t < T; ++t) tot += cnt[(size_t)(t * n + v)];
      row[v + 1] = tot;
    }
  });
  row[0] = 0;
  for (int32_t v = 0; v < n; ++v) row[v + 1] += row[v];
  parallel_slices(n, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) {
      int32_t at = row[v];
      for (int64_t i = 0; i < T; ++i) {
        const int64_t t = reverse ? T - 1 - i : i;
        const int32_t c = cnt[(size_t)(t * n + v)];
        cnt[(size_t)(t * n + v)] = reverse ? at + c : at;  // reverse fills its block from the end
        at += c;
      }
    }
  });
  run([&](int64_t t) {
    int64_t b, e;
    slice(t, &b, &e);
    int32_t* c = cnt.data() + t * n;
    if (reverse) {
      for (int64_t k = b; k < e; ++k) place(k, --c[key(k)]);
    } else {
      for (int64_t k = b; k < e; ++k) place(k, c[key(k)]++);
    }
  });
}

int build_csr_reversed(int32_t n, const std::vector<int32_t>& tail, const std::vector<int32_t>& head,
                       const std::vector<int32_t>* wt, smdp_hgraph** out) {
  int64_t m = (int64_t)tail.size();
  if (m > INT32_MAX) return SMDP_ERR_IO;
  smdp_hgraph* g = hg_alloc(n, (int32_t)m, wt != nullptr);
  if (!g) return SMDP_ERR_NOMEM;
  // Each row holds its arcs in reverse read order: the first arc read lands last.
  parallel_row_fill(n, m, g->row, true, [&](int64_t e) { return tail[(size_t)e]; },
                    [&](int64_t e, int32_t p) {
                      g->col[p] = head[(size_t)e];
                      if (wt) g->w[p] = (*wt)[(size_t)e];
                    });
  *out = g;
  return SMDP_OK;
}

// Whole-file reader: one read of the file's size when it is known.
bool slurp(const char* path, std::string* s) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  if (fseek(f, 0, SEEK_END) == 0) {
    const long size = ftell(f);
    if (size > 0 && fseek(f, 0, SEEK_SET) == 0) {
      s->resize((size_t)size);
      const size_t k = fread(&(*s)[0], 1, (size_t)size, f);
      s->resize(k);
    } else {
      rewind(f);
    }
  } else {
    rewind(f);
  }
  char buf[1 << 16];
  size_t k;
  while ((k = fread(buf, 1, sizeof buf, f)) > 0) s->append(buf, k);
  fclose(f);
  return true;
}

// atoi() semantics on [p, end): optional spaces/tabs, sign, digits.
int32_t atoi_span(const char* p, const char* end) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
  bool neg = false;
  if (p < end && (*p == '-' || *p == '+')) neg = (*p++ == '-');
  int64_t v = 0;
  while (p < end && *p >= '0' && *p <= '9') v = v * 10 + (*p++ - '0');
  return (int32_t)(neg ? -v : v);
}

// strtok(s, " ") over one line: tokens split on ' ' only.
void split_spaces(const char* b, const char* e, std::vector<std::pair<const char*, const char*>>* toks) {
  toks->clear();
  const char* p = b;
  while (p < e) {
    while (p < e && *p == ' ') ++p;
    if (p >= e) break;
    const char* q = p;
    while (q < e && *q != ' ') ++q;
    toks->emplace_back(p, q);
    p = q;
  }
}

// The DIMACS9 rules of read_dimacs9 on all host threads.  The text is cut at
// line starts; each thread parses its slice ('p', 's' and 'a' lines) into
// arrays of its own, which are joined in file order.  A file that is not one
// 'p' line followed by the arcs (an 'a' before the 'p' line, or several 'p'
// lines, which the sequential rules handle line by line) returns 1 so that
// the caller uses the sequential reader.
int read_dimacs9_parallel(const std::string& s, smdp_hgraph** out) {
  const int64_t len = (int64_t)s.size();
  const int64_t T = std::min<int64_t>(worker_count(), std::max<int64_t>(1, len >> 20));
  if (T < 2) return 1;
  const char* base = s.data();
  std::vector<int64_t> cut((size_t)T + 1);
  cut[0] = 0;
  cut[(size_t)T] = len;
  for (int64_t t = 1; t < T; ++t) {
    int64_t c = std::max(cut[(size_t)t - 1], len * t / T);
    while (c < len && c > 0 && base[c - 1] != '\n') ++c;
    cut[(size_t)t] = c;
  }
  struct Part {
    std::vector<int32_t> tail, head, wt;
    int64_t first_a = -1;         // offset of the first 'a' line
    std::vector<int64_t> p_at;    // offsets of 'p' lines
    int32_t n = -1, mdecl = 0;    // from the last 'p' line of the slice
    int64_t s_at = -1;            // offset of the last 's' line
    int32_t src = 0;
    bool bad = false;
  };
  std::vector<Part> parts((size_t)T);
  std::vector<std::thread> pool;
  auto work = [&](int64_t t) {
    Part& P = parts[(size_t)t];
    std::vector<std::pair<const char*, const char*>> toks;
    const char* p = base + cut[(size_t)t];
    const char* end = base + cut[(size_t)t + 1];
    const char* fend = base + len;
    P.tail.reserve((size_t)((end - p) / 12));
    P.head.reserve((size_t)((end - p) / 12));
    P.wt.reserve((size_t)((end - p) / 12));
    while (p < end) {
      const char* eol = (const char*)memchr(p, '\n', fend - p);
      const char* le = eol ? eol + 1 : fend;
      const char c = *p;
      if (c == 'p') {
        if (le - p < 4) { P.bad = true; break; }
        split_spaces(p + 4, le, &toks);
        if (toks.size() < 2) { P.bad = true; break; }
        P.p_at.push_back(p - base);
        P.n = atoi_span(toks[0].first, toks[0].second);
        P.mdecl = atoi_span(toks[1].first, toks[1].second);
      } else if (c == 's') {
        split_spaces(p + 1, le, &toks);
        if (toks.empty()) { P.bad = true; break; }
        P.s_at = p - base;
        P.src = atoi_span(toks[0].first, toks[0].second) - 1;
      } else if (c == 'a') {
        split_spaces(p + 1, le, &toks);
        if (toks.size() < 3) { P.bad = true; break; }
        if (P.first_a < 0) P.first_a = p - base;
        P.tail.push_back(atoi_span(toks[0].first, toks[0].second) - 1);
        P.head.push_back(atoi_span(toks[1].first, toks[1].second) - 1);
        P.wt.push_back(atoi_span(toks[2].first, toks[2].second));
      }
      p = le;
    }
  };
  for (int64_t t = 1; t < T; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  // one 'p' line, before every 'a' line; otherwise the sequential rules decide
  int64_t np = 0, p_at = -1, first_a = -1, s_at = -1;
  int32_t n = -1, mdecl = 0, src = 0;
  bool bad = false;
  for (const Part& P : parts) {
    np += (int64_t)P.p_at.size();
    if (!P.p_at.empty()) {
      p_at = P.p_at.back();
      n = P.n;
      mdecl = P.mdecl;
    }
    if (first_a < 0 && P.first_a >= 0) first_a = P.first_a;
    if (P.s_at >= 0) {
      s_at = P.s_at;
      src = P.src;
    }
    bad = bad || P.bad;
  }
  if (np != 1 || (first_a >= 0 && first_a < p_at)) return 1;
  (void)s_at;
  if (bad || n < 0 || mdecl < 0) return SMDP_ERR_IO;
  std::vector<int64_t> off((size_t)T + 1, 0);
  for (int64_t t = 0; t < T; ++t) off[(size_t)t + 1] = off[(size_t)t] + (int64_t)parts[(size_t)t].tail.size();
  const int64_t m = off[(size_t)T];
  if (m != mdecl) return SMDP_ERR_IO;
  if (n > 0 && (src < 0 || src >= n)) return SMDP_ERR_IO;
  std::vector<int32_t> tail((size_t)m), head((size_t)m), wt((size_t)m);
  std::vector<char> range_ok((size_t)T, 1);
  pool.clear();
  auto join = [&](int64_t t) {
    const Part& P = parts[(size_t)t];
    const size_t o = (size_t)off[(size_t)t];
    for (size_t k = 0; k < P.tail.size(); ++k) {
      const int32_t u = P.tail[k], v = P.head[k];
      if (u < 0 || u >= n || v < 0 || v >= n) range_ok[(size_t)t] = 0;
      tail[o + k] = u;
      head[o + k] = v;
      wt[o + k] = P.wt[k];
    }
  };
  for (int64_t t = 1; t < T; ++t) pool.emplace_back(join, t);
  join(0);
  for (auto& th : pool) th.join();
  for (char ok : range_ok)
    if (!ok) return SMDP_ERR_IO;
  parts.clear();
  int rc = build_csr_reversed(n, tail, head, &wt, out);
  if (rc == SMDP_OK) (*out)->source = src;
  return rc;
}

int read_dimacs9(const std::string& s, smdp_hgraph** out) {
  // common/global.cpp:29-64
  const int prc = read_dimacs9_parallel(s, out);
  if (prc != 1) return prc;
  int32_t n = -1, mdecl = 0, src = 0;
  std::vector<int32_t> tail, head, wt;
  std::vector<std::pair<const char*, const char*>> toks;
  const char* p = s.data();
  const char* end = p + s.size();
  while (p < end) {
    const char* eol = (const char*)memchr(p, '\n', end - p);
    const char* le = eol ? eol + 1 : end;  // fgets keeps the newline
    char c = *p;
    if (c == 'p') {
      if (le - p < 4) return SMDP_ERR_IO;
      split_spaces(p + 4, le, &toks);
      if (toks.size() < 2) return SMDP_ERR_IO;
      n = atoi_span(toks[0].first, toks[0].second);
      mdecl = atoi_span(toks[1].first, toks[1].second);
      if (n < 0 || mdecl < 0) return SMDP_ERR_IO;
      tail.reserve(mdecl);
      head.reserve(mdecl);
      wt.reserve(mdecl);
    } else if (c == 's') {
      split_spaces(p + 1, le, &toks);
      if (toks.empty()) return SMDP_ERR_IO;
      src = atoi_span(toks[0].first, toks[0].second) - 1;
    } else if (c == 'a') {
      if (n < 0) return SMDP_ERR_IO;
      split_spaces(p + 1, le, &toks);
      if (toks.size() < 3) return SMDP_ERR_IO;
      int32_t u = atoi_span(toks[0].first, toks[0].second) - 1;
      int32_t v = atoi_span(toks[1].first, toks[1].second) - 1;
      int32_t w = atoi_span(toks[2].first, toks[2].second);
      if (u < 0 || u >= n || v < 0 || v >= n) return SMDP_ERR_IO;
      tail.push_back(u);
      head.push_back(v);
      wt.push_back(w);
    }
    p = le;
  }
  if (n < 0) return SMDP_ERR_IO;
  if ((int64_t)tail.size() != mdecl) return SMDP_ERR_IO;
  if (n > 0 && (src < 0 || src >= n)) return SMDP_ERR_IO;
  int rc = build_csr_reversed(n, tail, head, &wt, out);
  if (rc == SMDP_OK) (*out)->source = src;
  return rc;
}

int read_dimacs10(const std::string& s, smdp_hgraph** out) {
  // common/global.cpp:66-105
  const char* p = s.data();
  const char* end = p + s.size();
  auto next_line = [&](const char** b, const char** e) -> bool {
    if (p >= end) return false;
    const char* eol = (const char*)memchr(p, '\n', end - p);
    *b = p;
    *e = eol ? eol + 1 : end;
    p = *e;
    return true;
  };
  const char *b, *e;
  do {
    if (!next_line(&b, &e)) return SMDP_ERR_IO;
  } while (*b == '%');
  std::vector<std::pair<const char*, const char*>> toks;
  split_spaces(b, e, &toks);
  if (toks.size() < 2) return SMDP_ERR_IO;
  int32_t n = atoi_span(toks[0].first, toks[0].second);
  int64_t mdecl = 2 * (int64_t)atoi_span(toks[1].first, toks[1].second);
  if (n < 0 || mdecl < 0 || mdecl > INT32_MAX) return SMDP_ERR_IO;
  std::vector<int32_t> tail, head, wt;
  tail.reserve(mdecl);
  head.reserve(mdecl);
  wt.reserve(mdecl);
  GlibcRand rnd(7);
  for (int32_t i = 0; i < n; ++i) {
    if (!next_line(&b, &e)) return SMDP_ERR_IO;  // "Parsing adjacencylist fails!"
    split_spaces(b, e, &toks);
    for (auto& t : toks) {
      if (*t.first == '\n') break;  // the trailing newline token ends the list
      int32_t v = atoi_span(t.first, t.second) - 1;
      if (v < 0 || v >= n) return SMDP_ERR_IO;
      tail.push_back(i);
      head.push_back(v);
      wt.push_back(rnd.next() % 1000 + 1);
    }
  }
  if ((int64_t)tail.size() != mdecl) return SMDP_ERR_IO;
  int rc = build_csr_reversed(n, tail, head, &wt, out);
  if (rc == SMDP_OK) (*out)->source = 0;
  return rc;
}

int read_slndc(const std::string& s, smdp_hgraph** out) {
  // common/global.cpp:107-190
  const char* p = s.data();
  const char* end = p + s.size();
  const char* line[4];
  const char* lend[4];
  for (int k = 0; k < 4; ++k) {
    if (p >= end) return SMDP_ERR_IO;
    const char* eol = (const char*)memchr(p, '\n', end - p);
    line[k] = p;
    lend[k] = eol ? eol + 1 : end;
    p = lend[k];
  }
  if (lend[2] - line[2] < 9) return SMDP_ERR_IO;
  std::vector<std::pair<const char*, const char*>> toks;
  split_spaces(line[2] + 9, lend[2], &toks);  // "# Nodes: N Edges: M"
  if (toks.size() < 3) return SMDP_ERR_IO;
  int64_t ndecl = atoi_span(toks[0].first, toks[0].second);
  if (ndecl <= 0 || 2 * ndecl > INT32_MAX) return SMDP_ERR_IO;
  int64_t cap = 2 * ndecl;  // ids must be below 2N
  // Pairs "tail head" until EOF.
  std::vector<int32_t> rt, rh;
  std::vector<char> used((size_t)cap, 0);
  int32_t maxid = 0;
  while (true) {
    while (p < end && isspace((unsigned char)*p)) ++p;
    if (p >= end) break;
    char* q;
    errno = 0;
    long a = strtol(p, &q, 10);
    if (q == p) return SMDP_ERR_IO;
    p = q;
    while (p < end && isspace((unsigned char)*p)) ++p;
    if (p >= end) return SMDP_ERR_IO;
    long bb = strtol(p, &q, 10);
    if (q == p) return SMDP_ERR_IO;
    p = q;
    if (a < 0 || bb < 0 || a >= cap || bb >= cap) return SMDP_ERR_IO;
    used[a] = used[bb] = 1;
    maxid = std::max<int32_t>(maxid, (int32_t)std::max(a, bb));
    rt.push_back((int32_t)a);
    rh.push_back((int32_t)bb);
  }
  std::vector<int32_t> map((size_t)maxid + 1, -1);
  int32_t n = 0;
  for (int32_t i = 0; i <= maxid; ++i)
    if (used[i]) map[i] = n++;
  // Per old id, its heads in file order (tempNodeList[tail].push_back).
  int64_t m = (int64_t)rt.size();
  std::vector<int32_t> cnt((size_t)maxid + 2, 0);
  for (int64_t e = 0; e < m; ++e) cnt[rt[e] + 1]++;
  for (int32_t i = 0; i <= maxid; ++i) cnt[i + 1] += cnt[i];
  std::vector<int32_t> order((size_t)m);
  {
    std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < m; ++e) order[pos[rt[e]]++] = (int32_t)e;
  }
  // Weights drawn walking old ids ascending, file order inside each id.
  std::vector<int32_t> tail, head, wt;
  tail.reserve(m);
  head.reserve(m);
  wt.reserve(m);
  GlibcRand rnd(7);
  for (int32_t i = 0; i <= maxid; ++i) {
    if (map[i] < 0) continue;
    for (int32_t k = cnt[i]; k < cnt[i + 1]; ++k) {
      int32_t e = order[k];
      tail.push_back(map[i]);
      head.push_back(map[rh[e]]);
      wt.push_back(rnd.next() % 1000 + 1);
    }
  }
  int rc = build_csr_reversed(n, tail, head, &wt, out);
  if (rc == SMDP_OK) (*out)->source = 0;
  return rc;
}

// Stable counting sort of generated arcs by tail into a CSR (rows keep
// generation order).
int csr_from_arcs(int32_t n, const std::vector<int32_t>& tail, const std::vector<int32_t>& head,
                  const std::vector<int32_t>* wt, smdp_hgraph** out) {
  int64_t m = (int64_t)tail.size();
  if (m > INT32_MAX) return SMDP_ERR_ARG;
  smdp_hgraph* g = hg_alloc(n, (int32_t)m, wt != nullptr);
  if (!g) return SMDP_ERR_NOMEM;
  parallel_row_fill(n, m, g->row, false, [&](int64_t e) { return tail[(size_t)e]; },
                    [&](int64_t e, int32_t p) {
                      g->col[p] = head[(size_t)e];
                      if (wt) g->w[p] = (*wt)[(size_t)e];
                    });
  *out = g;
  return SMDP_OK;
}

}  // namespace

extern "C" {

void smdp_hgraph_free(smdp_hgraph* g) {
  if (!g) return;
  free(g->row);
  free(g->col);
  free(g->w);
  free(g);
}

int smdp_read_graph(const char* path, int format, smdp_hgraph** out) {
  if (!path || !out) return SMDP_ERR_ARG;
  *out = nullptr;
  if (format < SMDP_FMT_DIMACS9 || format > SMDP_FMT_SLNDC) return SMDP_ERR_FORMAT;
  std::string s;
  if (!slurp(path, &s)) return SMDP_ERR_IO;
  switch (format) {
    case SMDP_FMT_DIMACS9: return read_dimacs9(s, out);
    case SMDP_FMT_DIMACS10: return read_dimacs10(s, out);
    default: return read_slndc(s, out);
  }
}

int smdp_invert(const smdp_hgraph* g, smdp_hgraph** out) {
  // invertGraph, common/global.cpp:237-277: in-rows in descending source order.
  if (!g || !out) return SMDP_ERR_ARG;
  smdp_hgraph* r = hg_alloc(g->n, g->m, false);
  if (!r) return SMDP_ERR_NOMEM;
  // the source of arc e (arcs are in row order, so reversing the list order
  // gives descending sources in every in-row)
  std::vector<int32_t> src((size_t)g->m);
  parallel_slices(g->n, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u)
      for (int32_t k = g->row[u]; k < g->row[u + 1]; ++k) src[(size_t)k] = (int32_t)u;
  });
  parallel_row_fill(g->n, g->m, r->row, true, [&](int64_t e) { return g->col[e]; },
                    [&](int64_t e, int32_t p) { r->col[p] = src[(size_t)e]; });
  r->source = g->source;
  *out = r;
  return SMDP_OK;
}

int smdp_write_dimacs9(const char* path, const smdp_hgraph* g) {
  if (!path || !g) return SMDP_ERR_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return SMDP_ERR_IO;
  std::vector<char> buf(1 << 20);
  setvbuf(f, buf.data(), _IOFBF, buf.size());
  fprintf(f, "c generated by smdp\np sp %d %d\ns %d\n", g->n, g->m, g->source + 1);
  // Write each row in reverse so that reading it back (which reverses rows)
  // reproduces this exact CSR.
  for (int32_t u = 0; u < g->n; ++u)
    for (int32_t e = g->row[u + 1] - 1; e >= g->row[u]; --e)
      fprintf(f, "a %d %d %d\n", u + 1, g->col[e] + 1, g->w ? g->w[e] : 1);
  int bad = ferror(f);
  fclose(f);
  return bad ? SMDP_ERR_IO : SMDP_OK;
}

int smdp_gen_uniform(int32_t n, int32_t deg, int32_t wlo, int32_t whi, uint64_t seed,
                     smdp_hgraph** out) {
  if (!out || n <= 0 || deg < 0 || (int64_t)n * deg > INT32_MAX || wlo > whi) return SMDP_ERR_ARG;
  int32_t m = n * deg;
  smdp_hgraph* g = hg_alloc(n, m, true);
  if (!g) return SMDP_ERR_NOMEM;
  for (int32_t u = 0; u <= n; ++u) g->row[u] = u * deg;
  parallel_slices(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      g->col[i] = smdp::uniform_int(rng_at(seed, smdp::STREAM_UNIFORM_DST, i), 0, n - 1);
      g->w[i] = smdp::uniform_int(rng_at(seed, smdp::STREAM_WEIGHT, i), wlo, whi);
    }
  });
  g->source = 0;
  *out = g;
  return SMDP_OK;
}

int smdp_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, int32_t wlo, int32_t whi,
                  int symmetrize, int permute, uint64_t seed, smdp_hgraph** out) {
  if (!out || scale < 1 || scale > 30 || ef < 1) return SMDP_ERR_ARG;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return SMDP_ERR_ARG;
  int64_t n = (int64_t)1 << scale;
  int64_t mgen = (int64_t)ef << scale;
  int64_t m = symmetrize ? 2 * mgen : mgen;
  if (m > INT32_MAX || n > INT32_MAX) return SMDP_ERR_ARG;
  bool weighted = whi > 0;
  if (weighted && wlo > whi) return SMDP_ERR_ARG;
  // Quadrant thresholds on 16-bit uniforms.
  const uint32_t ta = (uint32_t)llround(a * 65536.0);
  const uint32_t tab = (uint32_t)llround((a + b) * 65536.0);
  const uint32_t tabc = (uint32_t)llround((a + b + c) * 65536.0);
  const int draws = (scale + 3) / 4;  // four 16-bit fields per 64-bit draw
  std::vector<int32_t> perm;
  if (permute) {
    // Fisher-Yates with counter draws.
    perm.resize(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    for (int64_t i = n - 1; i > 0; --i) {
      int64_t j = smdp::uniform_int(rng_at(seed, smdp::STREAM_PERM, i), 0, (int32_t)i);
      std::swap(perm[i], perm[j]);
    }
  }
  std::vector<int32_t> tail(m), head(m), wt;
  if (weighted) wt.resize(m);
  parallel_slices(mgen, [&](int64_t b0, int64_t e0) {
    for (int64_t e = b0; e < e0; ++e) {
      uint32_t src = 0, dst = 0;
      int level = 0;
      for (int d = 0; d < draws; ++d) {
        uint64_t r = rng_at(seed, smdp::STREAM_RMAT, (uint64_t)e * draws + d);
        for (int f = 0; f < 4 && level < scale; ++f, ++level) {
          uint32_t u = (uint32_t)(r >> (16 * f)) & 0xFFFFu;
          uint32_t sb = u >= tab ? 1u : 0u;
          uint32_t db = (u >= ta && u < tab) || u >= tabc ? 1u : 0u;
          src = (src << 1) | sb;
          dst = (dst << 1) | db;
        }
      }
      int32_t s = permute ? perm[src] : (int32_t)src;
      int32_t t = permute ? perm[dst] : (int32_t)dst;
      tail[e] = s;
      head[e] = t;
      int32_t w = weighted ? smdp::uniform_int(rng_at(seed, smdp::STREAM_WEIGHT, e), wlo, whi) : 0;
      if (weighted) wt[e] = w;
      if (symmetrize) {
        tail[mgen + e] = t;
        head[mgen + e] = s;
        if (weighted) wt[mgen + e] = w;
      }
    }
  });
  int rc = csr_from_arcs((int32_t)n, tail, head, weighted ? &wt : nullptr, out);
  if (rc != SMDP_OK) return rc;
  int32_t src;
  smdp_gen_sources(*out, 1, seed, &src);
  (*out)->source = src;
  return SMDP_OK;
}

int smdp_gen_zipf_tree(int32_t n, double s, int32_t cap, uint64_t seed, smdp_hgraph** out) {
  if (!out || n <= 0 || cap < 1 || s <= 0) return SMDP_ERR_ARG;
  std::vector<double> cdf((size_t)cap);
  double acc = 0;
  for (int32_t k = 1; k <= cap; ++k) {
    acc += std::pow((double)k, -s);
    cdf[k - 1] = acc;
  }
  for (auto& x : cdf) x /= acc;
  smdp_hgraph* g = hg_alloc(n, n - 1, false);
  if (!g) return SMDP_ERR_NOMEM;
  int64_t next = 1;
  for (int32_t i = 0; i < n; ++i) {
    g->row[i] = (int32_t)(next - 1);
    if (next < n) {
      double u = smdp::uniform01(rng_at(seed, smdp::STREAM_ZIPF, (uint64_t)i));
      int64_t k = (int64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
      if (k > cap) k = cap;
      if (k > n - next) k = n - next;
      for (int64_t j = 0; j < k; ++j) g->col[next - 1 + j] = (int32_t)(next + j);
      next += k;
    }
  }
  g->row[n] = n - 1;
  g->source = 0;
  *out = g;
  return SMDP_OK;
}

int smdp_gen_small_graph(int32_t i, int32_t vlo, int32_t vhi, double dlo, double dhi, int32_t wlo,
                         int32_t whi, uint64_t seed, smdp_hgraph** out) {
  if (!out || i < 0 || vlo < 1 || vhi < vlo || dlo < 0 || dhi < dlo || wlo > whi) return SMDP_ERR_ARG;
  double lv = std::log((double)vlo) +
              (std::log((double)vhi) - std::log((double)vlo)) *
                  smdp::uniform01(rng_at(seed, smdp::STREAM_SMALL_PARAMS, 2 * (uint64_t)i));
  int32_t n = (int32_t)std::llround(std::exp(lv));
  if (n < vlo) n = vlo;
  if (n > vhi) n = vhi;
  double deg = dlo + (dhi - dlo) * smdp::uniform01(rng_at(seed, smdp::STREAM_SMALL_PARAMS, 2 * (uint64_t)i + 1));
  int64_t m = std::llround((double)n * deg);
  if (m > INT32_MAX) return SMDP_ERR_ARG;
  uint32_t stream = smdp::STREAM_SMALL_BASE + (uint32_t)i;
  std::vector<int32_t> tail(m), head(m), wt(m);
  for (int64_t e = 0; e < m; ++e) {
    tail[e] = smdp::uniform_int(rng_at(seed, stream, 3 * (uint64_t)e), 0, n - 1);
    head[e] = smdp::uniform_int(rng_at(seed, stream, 3 * (uint64_t)e + 1), 0, n - 1);
    wt[e] = smdp::uniform_int(rng_at(seed, stream, 3 * (uint64_t)e + 2), wlo, whi);
  }
  int rc = csr_from_arcs(n, tail, head, &wt, out);
  if (rc == SMDP_OK) (*out)->source = 0;
  return rc;
}

int smdp_gen_sources(const smdp_hgraph* g, int32_t k, uint64_t seed, int32_t* srcs) {
  if (!g || !srcs || k < 0 || k > g->n) return SMDP_ERR_ARG;
  std::vector<char> taken((size_t)g->n, 0);
  int32_t got = 0;
  for (uint64_t i = 0; got < k && i < (uint64_t)64 * g->n + 1024; ++i) {
    int32_t v = smdp::uniform_int(rng_at(seed, smdp::STREAM_SOURCE, i), 0, g->n - 1);
    if (taken[v] || g->row[v + 1] == g->row[v]) continue;
    taken[v] = 1;
    srcs[got++] = v;
  }
  for (int32_t v = 0; got < k && v < g->n; ++v)  // degenerate graphs: fill in id order
    if (!taken[v]) {
      taken[v] = 1;
      srcs[got++] = v;
    }
  return SMDP_OK;
}

int smdp_gen_uniform_f32(int64_t count, float lo, float hi, uint64_t seed, uint32_t stream, float* out) {
  if (!out || count < 0) return SMDP_ERR_ARG;
  parallel_slices(count, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      // 24 random bits: k * 2^-24 is exact in FP32.
      float u = (float)(rng_at(seed, stream, (uint64_t)i) >> 40) * (1.0f / 16777216.0f);
      out[i] = lo + (hi - lo) * u;
    }
  });
  return SMDP_OK;
}

int smdp_gen_prio(int32_t n, uint64_t seed, uint32_t* out) {
  if (!out || n < 0) return SMDP_ERR_ARG;
  parallel_slices(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) out[i] = (uint32_t)(rng_at(seed, smdp::STREAM_PRIO, (uint64_t)i) >> 32);
  });
  return SMDP_OK;
}

int smdp_gen_spmv_reference_inputs(int32_t n, int32_t m, float* data, float* x) {
  // apps/SpMV/src/SpMV_wrapper.cu:105-110
  if (n < 0 || m < 0 || (m && !data) || (n && !x)) return SMDP_ERR_ARG;
  for (int32_t i = 0; i < m; ++i) data[i] = (float)(i % 13) + 0.5f;
  for (int32_t i = 0; i < n; ++i) x[i] = (float)(i % 17) + 0.5f;
  return SMDP_OK;
}

int smdp_outdegree(const smdp_hgraph* g, int32_t* outdeg) {
  if (!g || !outdeg) return SMDP_ERR_ARG;
  for (int32_t v = 0; v < g->n; ++v) outdeg[v] = g->row[v + 1] - g->row[v];
  return SMDP_OK;
}

}  // extern "C"
