// Row reductions on gfx950: CSR SpMV (y = A x) and pull PageRank.
//
// Reference: csr_spmv_kernel and its consolidated variants
// (apps/SpMV/src/SpMV_kernel.cu:81-403) and pagerank_thread_pull_kernel and
// its variants (apps/pagerank/src/pagerank_kernel.cu:72-601), driven by
// SpMV_GPU (SpMV_wrapper.cu:193-224) and pg_*_gpu
// (pagerank_wrapper.cu:199-364).  Both compute, per row r,
//     acc(r) = sum over arcs e of row r of term(e)
// with term(e) = val[e] * x[col[e]] for SpMV and term(e) = r[u]/outdeg[u]
// (u = col[e] in the in-CSR) for PageRank, followed by a per-row epilogue.
//
// Here:
//  * FP32 with compensated accumulation: every product enters as p + e with
//    e = fma(a, b, -p) its exact rounding error, every addition is a TwoSum,
//    and the errors are added back once at the end (Dot2).  The reference
//    sums in FP64 (SpMV) or in FP32 with FP64 block reductions (PageRank).
//  * Every row's sum ends in the thread that owns the row in its tile, which
//    runs the epilogue; per-tile FP64 partials (PageRank residual and
//    dangling mass) are therefore reduced in a fixed order.
//  * PageRank's dangling mass and L1 residual are reduced in FP64 on the
//    device by the last workgroup of each iteration (a ticket counter), which
//    also applies the stop rule; later iterations already enqueued return at
//    once.  The reference copies the dangling ranks to the host and sums them
//    in float every iteration, and reads a "changed" flag back every
//    iteration (pagerank_wrapper.cu:212-233).
//  * Heavy rows go to wave / workgroup LDS queues or a grid queue as the
//    solution says; in the tuned solution rows of at least `split` arcs are
//    cut into chunks that any workgroup can take, and light rows are summed
//    from coalesced LDS windows of products.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_common.h"
#include "smdp_internal.h"

namespace smdp {

constexpr int kWin = 1024;     // light-row window: products staged in LDS per pass
constexpr int kRowBlockQ = SMDP_WG;

// Device-side state of a SpMV pass or a PageRank solve.
struct RState {
  double dang[2];    // dangling mass of the ranks read by iteration k: dang[k&1]
  double residual;   // L1 residual of the last finished iteration
  int32_t done;      // iterations run when the stop rule fired (0 = running)
  int32_t ticket;    // workgroups finished in the current pass
  int32_t work_next; // dynamic work-item counter
  int32_t hlen;      // grid solution: heavy rows pushed in phase 1
  int32_t hnext;     // grid solution: heavy rows taken in phase 2
  int32_t pad[23];
};

struct RArgs {
  const int32_t* __restrict__ row;
  const int32_t* __restrict__ col;
  const float* __restrict__ val;   // SpMV values
  const float* __restrict__ x;     // SpMV x; PageRank contributions of iteration k
  float* y;                        // SpMV y; PageRank ranks written by iteration k
  const float* __restrict__ rold;  // PageRank ranks read by iteration k
  float* cnext;                    // PageRank contributions for iteration k+1
  const int32_t* __restrict__ outdeg;
  RState* st;
  Counters* cnt;
  double2* tile_part;              // per tile: (residual, dangling)
  double2* heavy_part;             // grid solution: per heavy-queue entry
  int32_t* hq;                     // grid solution: heavy rows
  const int32_t* __restrict__ mrows;   // tuned: rows cut into chunks, ascending
  const int32_t* __restrict__ mfirst;  // tuned: first chunk of each such row (nm+1)
  float2* mpart;                       // tuned: per chunk (sum, error)
  int32_t n;
  int32_t ntiles;
  int32_t nm;        // tuned: rows cut into chunks
  int32_t nchunks;   // tuned: chunks
  int32_t heavy;
  int32_t split;
  int32_t chunk;
  float damping;
  float base_rw;     // (1-d) T / N
  double tol_abs;    // stop when residual < tol_abs
  int32_t max_iter;
};

template <bool PR>
__device__ __forceinline__ void add_term(const RArgs& a, CSum& acc, int e) {
  if (PR) acc.add(a.x[a.col[e]]);
  else acc.add_prod(a.val[e], a.x[a.col[e]]);
}

// Per-row epilogue in the owning thread.
template <bool PR>
__device__ __forceinline__ void epilogue(const RArgs& a, int r, float sum, float base, double& res, double& dang) {
  if (!PR) {
    a.y[r] = sum;
    return;
  }
  const float nr = __builtin_fmaf(a.damping, sum, base);
  a.y[r] = nr;
  const int od = a.outdeg[r];
  a.cnext[r] = od > 0 ? nr / (float)od : 0.f;
  res += fabs((double)nr - (double)a.rold[r]);
  if (od == 0) dang += (double)nr;
}

__device__ __forceinline__ double2 wg_sum2(double2 v, double2* lds) {
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  const int wave = threadIdx.x / SMDP_WAVE;
  if (lane_id() == 0) lds[wave] = v;
  __syncthreads();
  double2 t = make_double2(0.0, 0.0);
  for (int k = 0; k < SMDP_WAVES_PER_WG; ++k) {
    t.x += lds[k].x;
    t.y += lds[k].y;
  }
  __syncthreads();
  return t;
}

__device__ __forceinline__ CSum wg_csum(CSum v, float2* lds) {
  v = wave_csum(v);
  const int wave = threadIdx.x / SMDP_WAVE;
  if (lane_id() == 0) lds[wave] = make_float2(v.s, v.c);
  __syncthreads();
  CSum t;
  for (int k = 0; k < SMDP_WAVES_PER_WG; ++k) {
    CSum o;
    o.s = lds[k].x;
    o.c = lds[k].y;
    t.merge(o);
  }
  __syncthreads();
  return t;
}

// Wave-cooperative sum of arcs [s, e); every lane returns the same value.
template <bool PR>
__device__ __forceinline__ CSum wave_row(const RArgs& a, int s, int e) {
  CSum acc;
  for (int k = s + lane_id(); k < e; k += SMDP_WAVE) add_term<PR>(a, acc, k);
  acc = wave_csum(acc);
  CSum out;
  out.s = __shfl(acc.s, 0);
  out.c = __shfl(acc.c, 0);
  return out;
}

// Workgroup-cooperative sum of arcs [s, e) (all threads call it).
template <bool PR>
__device__ __forceinline__ CSum wg_row(const RArgs& a, int s, int e, float2* lds) {
  CSum acc;
  for (int k = s + (int)threadIdx.x; k < e; k += SMDP_WG) add_term<PR>(a, acc, k);
  return wg_csum(acc, lds);
}

__device__ __forceinline__ float iteration_base(const RArgs& a, int k) {
  // (1-d) T / N + d * D / N with D the dangling mass of the ranks being read.
  return a.base_rw + (float)((double)a.damping * a.st->dang[k & 1] / (double)a.n);
}

// Last workgroup of a pass: signalled by the ticket, after a release fence.
__device__ __forceinline__ bool arrive_last(RState* st, int* s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    *s_flag = atomicAdd(&st->ticket, 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (*s_flag) __threadfence();
  return *s_flag;
}

// FP64 totals, the stop rule and the counter reset (last workgroup only).
template <bool PR>
__device__ void finish_pass(const RArgs& a, int k, double2 extra, double2* lds2) {
  double2 acc = extra;
  if (PR)
    for (int t = threadIdx.x; t < a.ntiles; t += SMDP_WG) {
      acc.x += a.tile_part[t].x;
      acc.y += a.tile_part[t].y;
    }
  double2 tot = wg_sum2(acc, lds2);
  if (threadIdx.x == 0) {
    RState* st = a.st;
    if (PR) {
      st->residual = tot.x;
      st->dang[(k + 1) & 1] = tot.y;
      if (tot.x < a.tol_abs || k + 1 >= a.max_iter) st->done = k + 1;
    }
    st->ticket = 0;
    st->work_next = 0;
    st->hlen = 0;
    st->hnext = 0;
    atomicMax(&a.cnt->rounds, (unsigned long long)(k + 1));
  }
}

__device__ __forceinline__ void count_work(Counters* cnt, unsigned long long edges, unsigned long long rows,
                                           unsigned long long pushes) {
  edges = wave_sum(edges);
  rows = wave_sum(rows);
  pushes = wave_sum(pushes);
  if (lane_id() == 0) {
    if (edges) atomicAdd(&cnt->edges, edges);
    if (rows) atomicAdd(&cnt->verts, rows);
    if (pushes) atomicAdd(&cnt->push[1], pushes);
  }
}

// One pass (one PageRank iteration) with one lane per row: solutions FLAT,
// NAIVE, WAVE, BLOCK and phase 1 of GRID.  Tiles of SMDP_WG rows are taken in
// grid-stride order.
template <int SOL, bool PR>
__global__ __launch_bounds__(SMDP_WG) void k_rows_lane(RArgs a, int k) {
  if (PR && a.st->done) return;
  const float base = PR ? iteration_base(a, k) : 0.f;
  const int lane = lane_id();
  __shared__ float2 s_red[SMDP_WAVES_PER_WG];
  __shared__ double2 s_red2[SMDP_WAVES_PER_WG];
  __shared__ int32_t s_bq[SOL == SMDP_SOL_BLOCK ? kRowBlockQ : 1];
  __shared__ float2 s_bsum[SOL == SMDP_SOL_BLOCK ? SMDP_WG : 1];
  __shared__ int32_t s_bqn;
  __shared__ int s_last;
  unsigned long long edges = 0, rows = 0, pushes = 0;

  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int r = tile * SMDP_WG + (int)threadIdx.x;
    const bool valid = r < a.n;
    int s = 0, e = 0;
    if (valid) {
      s = a.row[r];
      e = a.row[r + 1];
      ++rows;
    }
    const bool heavy = SOL != SMDP_SOL_FLAT && valid && (e - s) >= a.heavy;
    CSum acc;
    if (SOL == SMDP_SOL_NAIVE || SOL == SMDP_SOL_WAVE) {
      // NAIVE expands heavy rows before the light work, WAVE after it; in
      // both the whole wave reduces one heavy row at a time and the owner
      // lane keeps the result.
      if (SOL == SMDP_SOL_WAVE && !heavy)
        for (int q = s; q < e; ++q) add_term<PR>(a, acc, q);
      uint64_t hm = ballot(heavy);
      pushes += (lane == 0) ? __popcll(hm) : 0;
      while (hm) {
        const int l = first_lane(hm);
        hm &= hm - 1;
        CSum h = wave_row<PR>(a, __shfl(s, l), __shfl(e, l));
        if (lane == l) acc = h;
      }
      if (SOL == SMDP_SOL_NAIVE && !heavy)
        for (int q = s; q < e; ++q) add_term<PR>(a, acc, q);
    } else if (!heavy) {
      for (int q = s; q < e; ++q) add_term<PR>(a, acc, q);
    }
    if (SOL == SMDP_SOL_BLOCK) {
      if (threadIdx.x == 0) s_bqn = 0;
      __syncthreads();
      if (heavy) s_bq[atomicAdd(&s_bqn, 1)] = (int)threadIdx.x;
      __syncthreads();
      const int qn = s_bqn;
      pushes += threadIdx.x == 0 ? qn : 0;
      for (int j = 0; j < qn; ++j) {
        const int owner = s_bq[j];
        const int hr = tile * SMDP_WG + owner;
        CSum h = wg_row<PR>(a, a.row[hr], a.row[hr + 1], s_red);
        if (threadIdx.x == 0) s_bsum[owner] = make_float2(h.s, h.c);
      }
      __syncthreads();
      if (heavy) {
        acc.s = s_bsum[threadIdx.x].x;
        acc.c = s_bsum[threadIdx.x].y;
      }
    }
    if (SOL == SMDP_SOL_GRID && heavy) {
      const int slot = atomicAdd(&a.st->hlen, 1);
      a.hq[slot] = r;
      ++pushes;
    }
    // heavy rows of the grid solution are counted by phase 2
    if (valid && !(SOL == SMDP_SOL_GRID && heavy)) edges += e - s;
    double res = 0.0, dang = 0.0;
    if (valid && !(SOL == SMDP_SOL_GRID && heavy)) epilogue<PR>(a, r, acc.value(), base, res, dang);
    if (PR) {
      double2 t = wg_sum2(make_double2(res, dang), s_red2);
      if (threadIdx.x == 0) a.tile_part[tile] = t;
    }
  }
  count_work(a.cnt, edges, rows, pushes);
  if (SOL != SMDP_SOL_GRID && PR) {
    if (arrive_last(a.st, &s_last)) finish_pass<PR>(a, k, make_double2(0.0, 0.0), s_red2);
  }
}

// Phase 2 of the grid solution: each heavy row pushed in phase 1 is reduced
// by a whole workgroup; the last workgroup finishes the pass.
template <bool PR>
__global__ __launch_bounds__(SMDP_WG) void k_rows_grid(RArgs a, int k) {
  if (PR && a.st->done) return;
  const float base = PR ? iteration_base(a, k) : 0.f;
  __shared__ float2 s_red[SMDP_WAVES_PER_WG];
  __shared__ double2 s_red2[SMDP_WAVES_PER_WG];
  __shared__ int s_idx;
  __shared__ int s_last;
  const int H = a.st->hlen;
  unsigned long long edges = 0;
  for (;;) {
    if (threadIdx.x == 0) s_idx = atomicAdd(&a.st->hnext, 1);
    __syncthreads();
    const int idx = s_idx;
    __syncthreads();
    if (idx >= H) break;
    const int r = a.hq[idx];
    const int s = a.row[r], e = a.row[r + 1];
    CSum h = wg_row<PR>(a, s, e, s_red);
    if (threadIdx.x == 0) {
      double res = 0.0, dang = 0.0;
      epilogue<PR>(a, r, h.value(), base, res, dang);
      if (PR) a.heavy_part[idx] = make_double2(res, dang);
      edges += e - s;
    }
  }
  count_work(a.cnt, edges, 0, 0);
  if (arrive_last(a.st, &s_last)) {
    double2 extra = make_double2(0.0, 0.0);
    if (PR)
      for (int t = threadIdx.x; t < H; t += SMDP_WG) {
        extra.x += a.heavy_part[t].x;
        extra.y += a.heavy_part[t].y;
      }
    finish_pass<PR>(a, k, extra, s_red2);
  }
}

// Tuned solution, one kernel per pass.  Work items, taken from a shared
// counter: first the chunks of rows with >= split arcs (largest work first),
// then tiles of SMDP_WG consecutive rows.  In a tile, rows with >= heavy arcs
// are reduced by their wave; the others are summed by their owner thread from
// LDS windows that the whole workgroup fills with coalesced loads.  The last
// workgroup combines the chunk partials of each split row in order, runs
// their epilogues and finishes the pass.
template <bool PR>
__global__ __launch_bounds__(SMDP_WG) void k_rows_tuned(RArgs a, int k) {
  if (PR && a.st->done) return;
  const float base = PR ? iteration_base(a, k) : 0.f;
  const int lane = lane_id();
  __shared__ int s_item;
  __shared__ int s_last;
  __shared__ int s_start[SMDP_WG];
  __shared__ int s_off[SMDP_WG];
  __shared__ int s_scan[SMDP_WAVES_PER_WG];
  __shared__ float s_p[kWin];
  __shared__ float s_e[PR ? 1 : kWin];
  __shared__ float2 s_red[SMDP_WAVES_PER_WG];
  __shared__ double2 s_red2[SMDP_WAVES_PER_WG];
  unsigned long long edges = 0, rows = 0, pushes = 0;
  const int nitems = a.nchunks + a.ntiles;

  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&a.st->work_next, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= nitems) break;
    if (item < a.nchunks) {
      // ---- one chunk of a split row: binary search its row in mfirst
      int lo = 0, hi = a.nm - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.mfirst[mid] <= item) lo = mid;
        else hi = mid - 1;
      }
      const int r = a.mrows[lo];
      const int s = a.row[r] + (item - a.mfirst[lo]) * a.chunk;
      const int e = min(a.row[r + 1], s + a.chunk);
      CSum h = wg_row<PR>(a, s, e, s_red);
      if (threadIdx.x == 0) {
        a.mpart[item] = make_float2(h.s, h.c);
        edges += e - s;
      }
      continue;
    }
    // ---- a tile of rows
    const int tile = item - a.nchunks;
    const int r = tile * SMDP_WG + (int)threadIdx.x;
    const bool valid = r < a.n;
    int s = 0, deg = 0;
    if (valid) {
      s = a.row[r];
      deg = a.row[r + 1] - s;
      ++rows;
    }
    const bool split = deg >= a.split;
    const bool heavy = !split && deg >= a.heavy;
    const int ldeg = (valid && !split && !heavy) ? deg : 0;
    CSum acc;
    // medium rows: wave-cooperative
    uint64_t hm = ballot(heavy);
    pushes += lane == 0 ? __popcll(hm) : 0;
    while (hm) {
      const int l = first_lane(hm);
      hm &= hm - 1;
      CSum h = wave_row<PR>(a, __shfl(s, l), __shfl(s, l) + __shfl(deg, l));
      if (lane == l) acc = h;
    }
    if (heavy) edges += deg;
    // light rows: LDS windows of products
    s_start[threadIdx.x] = s;
    int tot;
    const int ex = wg_exclusive_scan(ldeg, s_scan, &tot);
    s_off[threadIdx.x] = ex;
    __syncthreads();
    for (int wb = 0; wb < tot; wb += kWin) {
      const int wn = min(kWin, tot - wb);
#pragma unroll
      for (int i = 0; i < kWin / SMDP_WG; ++i) {
        const int j = wb + (int)threadIdx.x + i * SMDP_WG;
        if (j < wb + wn) {
          int lo = 0, hi = SMDP_WG - 1;
          while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= j) lo = mid;
            else hi = mid - 1;
          }
          const int q = s_start[lo] + (j - s_off[lo]);
          if (PR) {
            s_p[j - wb] = a.x[a.col[q]];
          } else {
#pragma clang fp contract(off)
            const float va = a.val[q], xb = a.x[a.col[q]];
            const float p = va * xb;
            s_p[j - wb] = p;
            s_e[j - wb] = __builtin_fmaf(va, xb, -p);
          }
        }
      }
      __syncthreads();
      const int lo_q = max(ex, wb), hi_q = min(ex + ldeg, wb + wn);
      for (int q = lo_q; q < hi_q; ++q) {
        acc.add(s_p[q - wb]);
        if (!PR) acc.c += s_e[q - wb];
      }
      __syncthreads();
    }
    edges += ldeg;
    double res = 0.0, dang = 0.0;
    if (valid && !split) epilogue<PR>(a, r, acc.value(), base, res, dang);
    if (PR) {
      double2 t = wg_sum2(make_double2(res, dang), s_red2);
      if (threadIdx.x == 0) a.tile_part[tile] = t;
    }
  }
  count_work(a.cnt, edges, rows, pushes);
  if (arrive_last(a.st, &s_last)) {
    double res = 0.0, dang = 0.0;
    for (int i = threadIdx.x; i < a.nm; i += SMDP_WG) {
      CSum h;
      for (int c = a.mfirst[i]; c < a.mfirst[i + 1]; ++c) {
        CSum o;
        o.s = a.mpart[c].x;
        o.c = a.mpart[c].y;
        h.merge(o);
      }
      epilogue<PR>(a, a.mrows[i], h.value(), base, res, dang);
    }
    finish_pass<PR>(a, k, make_double2(res, dang), s_red2);
  }
}

__global__ void k_pr_init(float* rank, float* contrib, const int32_t* __restrict__ outdeg, int32_t n, float init,
                          RState* st, double dang0, Counters* cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    rank[i] = init;
    const int od = outdeg ? outdeg[i] : 0;
    contrib[i] = od > 0 ? init / (float)od : 0.f;
  }
  if (blockIdx.x == 0) {
    int32_t* p = reinterpret_cast<int32_t*>(st);
    for (int i = threadIdx.x; i < (int)(sizeof(RState) / 4); i += blockDim.x) p[i] = 0;
    unsigned long long* c = reinterpret_cast<unsigned long long*>(cnt);
    for (int i = threadIdx.x; i < (int)(sizeof(Counters) / 8); i += blockDim.x) c[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) st->dang[0] = dang0;
  }
}

// ---------------------------------------------------------------- host side

struct RowPlan {
  int32_t split = -1, chunk = -1;
  int32_t nm = 0, nchunks = 0;
  int32_t* mrows = nullptr;
  int32_t* mfirst = nullptr;
};

}  // namespace smdp

// Extra per-graph buffers of the row-reduce family, kept outside smdp_graph's
// public layout.
namespace {
struct RowBufs {
  smdp::RowPlan plan;
  smdp::RState* st = nullptr;
  double2* tile_part = nullptr;
  double2* heavy_part = nullptr;
  int32_t* hq = nullptr;
  float2* mpart = nullptr;
  int64_t mpart_cap = 0;
  float* contrib2 = nullptr;
  int64_t ndang = -1;
};
}  // namespace

namespace smdp {

static RowBufs* bufs_of(smdp_graph* g) {
  // Stored in the graph's generic extension slot.
  if (!g->rowbufs) g->rowbufs = new RowBufs();
  return static_cast<RowBufs*>(g->rowbufs);
}

void rowbufs_free(smdp_graph* g) {
  RowBufs* b = static_cast<RowBufs*>(g->rowbufs);
  if (!b) return;
  void* ptrs[] = {b->st, b->tile_part, b->heavy_part, b->hq, b->mpart, b->contrib2, b->plan.mrows, b->plan.mfirst};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete b;
  g->rowbufs = nullptr;
}

// Rows with >= split arcs and their chunk offsets: a one-time analysis per
// (graph, split, chunk), kept with the resident graph.
static int ensure_plan(smdp_graph* g, RowBufs* b, int32_t split, int32_t chunk) {
  RowPlan& p = b->plan;
  if (p.split == split && p.chunk == chunk) return SMDP_OK;
  std::vector<int32_t> row((size_t)g->n + 1);
  SMDP_HIP(hipMemcpy(row.data(), g->row, sizeof(int32_t) * row.size(), hipMemcpyDeviceToHost));
  std::vector<int32_t> mrows, mfirst(1, 0);
  for (int32_t r = 0; r < g->n; ++r) {
    int32_t d = row[r + 1] - row[r];
    if (d >= split) {
      mrows.push_back(r);
      mfirst.push_back(mfirst.back() + (d + chunk - 1) / chunk);
    }
  }
  if (p.mrows) (void)hipFree(p.mrows);
  if (p.mfirst) (void)hipFree(p.mfirst);
  p.mrows = p.mfirst = nullptr;
  p.nm = (int32_t)mrows.size();
  p.nchunks = mfirst.back();
  SMDP_CHECK(alloc_once(&p.mrows, std::max<int64_t>(1, p.nm)));
  SMDP_CHECK(alloc_once(&p.mfirst, (int64_t)p.nm + 1));
  if (p.nm) SMDP_HIP(hipMemcpy(p.mrows, mrows.data(), sizeof(int32_t) * p.nm, hipMemcpyHostToDevice));
  SMDP_HIP(hipMemcpy(p.mfirst, mfirst.data(), sizeof(int32_t) * mfirst.size(), hipMemcpyHostToDevice));
  SMDP_CHECK(ensure(&b->mpart, &b->mpart_cap, std::max<int32_t>(1, p.nchunks)));
  p.split = split;
  p.chunk = chunk;
  return SMDP_OK;
}

template <bool PR>
static void launch_pass(int sol, int grid, hipStream_t s, const RArgs& a, int k, int* launches, smdp_ctx* ctx) {
  dim3 g(grid), b(SMDP_WG);
  prof_mark(ctx, 0);
  switch (sol) {
    case SMDP_SOL_FLAT: hipLaunchKernelGGL((k_rows_lane<SMDP_SOL_FLAT, PR>), g, b, 0, s, a, k); break;
    case SMDP_SOL_NAIVE: hipLaunchKernelGGL((k_rows_lane<SMDP_SOL_NAIVE, PR>), g, b, 0, s, a, k); break;
    case SMDP_SOL_WAVE: hipLaunchKernelGGL((k_rows_lane<SMDP_SOL_WAVE, PR>), g, b, 0, s, a, k); break;
    case SMDP_SOL_BLOCK: hipLaunchKernelGGL((k_rows_lane<SMDP_SOL_BLOCK, PR>), g, b, 0, s, a, k); break;
    case SMDP_SOL_GRID: hipLaunchKernelGGL((k_rows_lane<SMDP_SOL_GRID, PR>), g, b, 0, s, a, k); break;
    default: hipLaunchKernelGGL((k_rows_tuned<PR>), g, b, 0, s, a, k); break;
  }
  prof_mark(ctx, 0);
  ++*launches;
  if (sol == SMDP_SOL_GRID) {
    prof_mark(ctx, 1);
    hipLaunchKernelGGL((k_rows_grid<PR>), g, b, 0, s, a, k);
    prof_mark(ctx, 1);
    ++*launches;
  }
}

static int prepare_args(smdp_graph* g, int app, int sol, RArgs* a, RowBufs** out) {
  if (sol < 0 || sol >= SMDP_SOL_COUNT || sol == SMDP_SOL_HIER) return SMDP_ERR_SOLUTION;
  smdp_ctx* ctx = g->ctx;
  const smdp_params& p = params_of(ctx, app);
  RowBufs* b = bufs_of(g);
  const int ntiles = (g->n + SMDP_WG - 1) / SMDP_WG;
  SMDP_CHECK(alloc_once(&b->st, 1));
  SMDP_CHECK(alloc_once(&b->tile_part, std::max(1, ntiles)));
  SMDP_CHECK(alloc_once(&b->heavy_part, std::max(1, g->n)));
  SMDP_CHECK(alloc_once(&b->hq, std::max(1, g->n)));
  if (sol == SMDP_SOL_TUNED) SMDP_CHECK(ensure_plan(g, b, std::max(p.split_threshold, p.heavy_threshold), p.chunk));
  *a = RArgs{};
  a->row = g->row;
  a->col = g->col;
  a->st = b->st;
  a->cnt = g->cnt;
  a->tile_part = b->tile_part;
  a->heavy_part = b->heavy_part;
  a->hq = b->hq;
  a->mrows = b->plan.mrows;
  a->mfirst = b->plan.mfirst;
  a->mpart = b->mpart;
  a->n = g->n;
  a->ntiles = ntiles;
  a->nm = sol == SMDP_SOL_TUNED ? b->plan.nm : 0;
  a->nchunks = sol == SMDP_SOL_TUNED ? b->plan.nchunks : 0;
  a->heavy = p.heavy_threshold;
  a->split = sol == SMDP_SOL_TUNED ? b->plan.split : INT32_MAX;
  a->chunk = p.chunk;
  *out = b;
  return SMDP_OK;
}

int run_spmv(smdp_graph* g, int sol, smdp_stats* st) {
  if (!g || !g->val || !g->x) return SMDP_ERR_ARG;
  smdp_ctx* ctx = g->ctx;
  SMDP_HIP(hipSetDevice(ctx->device));
  RArgs a;
  RowBufs* b;
  SMDP_CHECK(prepare_args(g, SMDP_APP_SPMV, sol, &a, &b));
  SMDP_CHECK(alloc_once(&g->y, std::max(1, g->n)));
  a.val = g->val;
  a.x = g->x;
  a.y = g->y;
  hipStream_t s = ctx->stream;
  prof_reset(ctx);
  const int grid = grid_size(ctx, params_of(ctx, SMDP_APP_SPMV));
  int launches = 0;
  SMDP_HIP(hipEventRecord(ctx->ev_begin, s));
  hipLaunchKernelGGL(k_pr_init, dim3(1), dim3(256), 0, s, (float*)nullptr, (float*)nullptr, (const int32_t*)nullptr, 0,
                     0.f, b->st, 0.0, g->cnt);
  ++launches;
  if (g->n > 0) launch_pass<false>(sol, grid, s, a, 0, &launches, ctx);
  SMDP_HIP(hipGetLastError());
  SMDP_HIP(hipEventRecord(ctx->ev_end, s));
  SMDP_HIP(hipEventSynchronize(ctx->ev_end));
  float ms = 0;
  SMDP_HIP(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
  g->last_res = SMDP_RES_Y;
  if (st) {
    SMDP_CHECK(fill_stats(g, st, ms, launches));
    st->rounds = 1;
    SMDP_CHECK(prof_collect(ctx, st));
  } else {
    prof_reset(ctx);
  }
  return SMDP_OK;
}

int run_pagerank(smdp_graph* g, const smdp_pr_opts* o, int sol, smdp_stats* st) {
  if (!g || !o || !g->outdeg || g->n <= 0) return SMDP_ERR_ARG;
  if (!(o->damping >= 0.f && o->damping <= 1.f) || !(o->total_rank > 0.f) || o->max_iter <= 0 || !(o->tol_l1 >= 0))
    return SMDP_ERR_ARG;
  smdp_ctx* ctx = g->ctx;
  SMDP_HIP(hipSetDevice(ctx->device));
  RArgs a;
  RowBufs* b;
  SMDP_CHECK(prepare_args(g, SMDP_APP_PAGERANK, sol, &a, &b));
  SMDP_CHECK(alloc_once(&g->rank[0], g->n));
  SMDP_CHECK(alloc_once(&g->rank[1], g->n));
  SMDP_CHECK(alloc_once(&g->contrib, g->n));
  SMDP_CHECK(alloc_once(&b->contrib2, g->n));
  if (b->ndang < 0) {
    std::vector<int32_t> od((size_t)g->n);
    SMDP_HIP(hipMemcpy(od.data(), g->outdeg, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost));
    b->ndang = std::count(od.begin(), od.end(), 0);
  }
  const float T = o->total_rank;
  const float init = o->init_zero ? 0.f : T / (float)g->n;
  const double dang0 = (double)init * (double)b->ndang;
  a.outdeg = g->outdeg;
  a.damping = o->damping;
  a.base_rw = (float)((1.0 - (double)o->damping) * (double)T / (double)g->n);
  a.tol_abs = o->tol_l1 * (double)T;
  a.max_iter = o->max_iter;
  float* cbuf[2] = {g->contrib, b->contrib2};
  hipStream_t s = ctx->stream;
  prof_reset(ctx);
  const smdp_params& p = params_of(ctx, SMDP_APP_PAGERANK);
  const int grid = grid_size(ctx, p);
  int launches = 0;
  SMDP_HIP(hipEventRecord(ctx->ev_begin, s));
  hipLaunchKernelGGL(k_pr_init, dim3(std::min<int64_t>((g->n + 255) / 256, 4096)), dim3(256), 0, s, g->rank[0],
                     cbuf[0], (const int32_t*)g->outdeg, g->n, init, b->st, dang0, g->cnt);
  ++launches;
  int k = 0;
  int batch = std::max(4, p.batch_rounds);
  int32_t done = 0;
  while (k < o->max_iter) {
    for (int j = 0; j < batch && k < o->max_iter; ++j, ++k) {
      a.x = cbuf[k & 1];
      a.cnext = cbuf[(k + 1) & 1];
      a.rold = g->rank[k & 1];
      a.y = g->rank[(k + 1) & 1];
      launch_pass<true>(sol, grid, s, a, k, &launches, ctx);
    }
    SMDP_HIP(hipGetLastError());
    SMDP_HIP(hipMemcpyAsync(ctx->pinned, &b->st->done, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SMDP_HIP(hipStreamSynchronize(s));
    done = ctx->pinned[0];
    if (done) break;
    batch = std::min(batch * 2, 32);
  }
  SMDP_HIP(hipEventRecord(ctx->ev_end, s));
  SMDP_HIP(hipEventSynchronize(ctx->ev_end));
  float ms = 0;
  SMDP_HIP(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
  if (!done) return SMDP_ERR_DEVICE;  // the last iteration always sets done
  g->rank_cur = done & 1;
  g->last_res = SMDP_RES_RANK;
  if (st) {
    SMDP_CHECK(fill_stats(g, st, ms, launches));
    RState hs;
    SMDP_HIP(hipMemcpy(&hs, b->st, sizeof hs, hipMemcpyDeviceToHost));
    st->rounds = (uint64_t)done;
    st->residual = hs.residual;
    st->dangling = hs.dang[done & 1];
    st->m_reached = (uint64_t)g->m * (uint64_t)done;
    SMDP_CHECK(prof_collect(ctx, st));
  } else {
    prof_reset(ctx);
  }
  return SMDP_OK;
}

}  // namespace smdp

using namespace smdp;

extern "C" {

int smdp_spmv_run(smdp_graph* g, int solution, smdp_stats* st) { return run_spmv(g, solution, st); }

int smdp_pagerank_run(smdp_graph* g, const smdp_pr_opts* opt, int solution, smdp_stats* st) {
  return run_pagerank(g, opt, solution, st);
}

int smdp_spmv_f32(smdp_ctx* ctx, const smdp_csr* a, const float* val, const float* x, int solution, float* y,
                  smdp_stats* st) {
  if (!ctx || !a || !val || !x || !y) return SMDP_ERR_ARG;
  double t0 = now_ms();
  smdp_graph* g = nullptr;
  smdp_csr c = *a;
  c.w = nullptr;
  SMDP_CHECK(smdp_graph_create(ctx, &c, &g));
  int rc = smdp_graph_set_values(g, val);
  if (rc == SMDP_OK) rc = smdp_graph_set_x(g, x);
  if (rc == SMDP_OK) rc = run_spmv(g, solution, st);
  if (rc == SMDP_OK) rc = smdp_graph_read(g, SMDP_RES_Y, y, g->n);
  smdp_graph_destroy(g);
  if (st) st->ms_total = now_ms() - t0;
  return rc;
}

int smdp_pagerank_f32(smdp_ctx* ctx, const smdp_csr* in_csr, const int32_t* outdeg, const smdp_pr_opts* opt,
                      int solution, float* rank, smdp_stats* st) {
  if (!ctx || !in_csr || !outdeg || !opt || !rank) return SMDP_ERR_ARG;
  double t0 = now_ms();
  smdp_graph* g = nullptr;
  smdp_csr c = *in_csr;
  c.w = nullptr;
  SMDP_CHECK(smdp_graph_create(ctx, &c, &g));
  int rc = smdp_graph_set_outdeg(g, outdeg);
  if (rc == SMDP_OK) rc = run_pagerank(g, opt, solution, st);
  if (rc == SMDP_OK) rc = smdp_graph_read(g, SMDP_RES_RANK, rank, g->n);
  smdp_graph_destroy(g);
  if (st) st->ms_total = now_ms() - t0;
  return rc;
}

}  // extern "C"
