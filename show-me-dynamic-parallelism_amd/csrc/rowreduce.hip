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
#include <cstddef>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "rowreduce.h"
#include "smdp_internal.h"

namespace smdp {

constexpr int kWin = 1024;     // light-row window: products staged in LDS per pass
constexpr int kRowBlockQ = SMDP_WG;


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
    if (threadIdx.x == 0) s_item = atomicAdd(&a.st->work_next, kGrab);
    __syncthreads();
    const int first = s_item;
    __syncthreads();
    if (first >= nitems) break;
    for (int item = first; item < min(first + kGrab, nitems); ++item) {
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
      {
        // kWin/SMDP_WG slots per thread: owners found together, then all
        // column loads, then all gathers, so the gathers overlap.
        constexpr int K = kWin / SMDP_WG;
        int j[K], own[K], q[K], c[K];
#pragma unroll
        for (int i = 0; i < K; ++i) j[i] = min(wb + (int)threadIdx.x + i * SMDP_WG, wb + wn - 1);
        tile_owners<K>(s_off, j, own);
#pragma unroll
        for (int i = 0; i < K; ++i) q[i] = s_start[own[i]] + (j[i] - s_off[own[i]]);
#pragma unroll
        for (int i = 0; i < K; ++i) c[i] = stream_load(a.col + q[i]);
        if (PR) {
          float g[K];
#pragma unroll
          for (int i = 0; i < K; ++i) g[i] = a.x[c[i]];
#pragma unroll
          for (int i = 0; i < K; ++i)
            if (wb + (int)threadIdx.x + i * SMDP_WG < wb + wn) s_p[j[i] - wb] = g[i];
        } else {
#pragma clang fp contract(off)
          float va[K], xb[K];
#pragma unroll
          for (int i = 0; i < K; ++i) {
            va[i] = stream_load(a.val + q[i]);
            xb[i] = a.x[c[i]];
          }
#pragma unroll
          for (int i = 0; i < K; ++i) {
            if (wb + (int)threadIdx.x + i * SMDP_WG < wb + wn) {
              const float p = va[i] * xb[i];
              s_p[j[i] - wb] = p;
              s_e[j[i] - wb] = __builtin_fmaf(va[i], xb[i], -p);
            }
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

// ---- column-blocked passes ---------------------------------------------------
// The sources (columns) are split into kXcd equal blocks and each block has
// its own CSR (browp/bcol/bval).  Workgroup i works only on block i % kXcd,
// so on MI355X, which hands workgroups to its 8 XCDs round-robin, every XCD
// gathers x (or the PageRank contributions) from one slice of n/8 entries
// that stays in its own L2.  Pass A writes one compensated partial sum per
// (block, row); pass B merges the kXcd partials of each row in block order
// and runs the epilogue, so results do not depend on the schedule.

struct BSrc {
  const int32_t* __restrict__ col;
  const float* __restrict__ val;
  const float* __restrict__ x;
};

template <bool PR>
__device__ __forceinline__ void bterm(const BSrc& s, CSum& acc, int e) {
  if (PR) acc.add(s.x[s.col[e]]);
  else acc.add_prod(s.val[e], s.x[s.col[e]]);
}

template <bool PR>
__device__ __forceinline__ CSum bwave_row(const BSrc& src, int s, int e) {
  CSum acc;
  for (int k = s + lane_id(); k < e; k += SMDP_WAVE) bterm<PR>(src, acc, k);
  acc = wave_csum(acc);
  CSum out;
  out.s = __shfl(acc.s, 0);
  out.c = __shfl(acc.c, 0);
  return out;
}

template <bool PR>
__device__ __forceinline__ CSum bwg_row(const BSrc& src, int s, int e, float2* lds) {
  CSum acc;
  int k = s + (int)threadIdx.x;
  for (; k + 3 * SMDP_WG < e; k += 4 * SMDP_WG) {
    int c[4];
    float g[4], va[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) c[t] = src.col[k + t * SMDP_WG];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      g[t] = src.x[c[t]];
      va[t] = PR ? 0.f : src.val[k + t * SMDP_WG];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (PR) acc.add(g[t]);
      else acc.add_prod(va[t], g[t]);
    }
  }
  for (; k < e; k += SMDP_WG) bterm<PR>(src, acc, k);
  return wg_csum(acc, lds);
}

constexpr int kRowsPerThread = 8;  // blocked pass A: consecutive rows per thread

template <bool PR>
__global__ __launch_bounds__(SMDP_WG) void k_rows_blk(RArgs a, int k) {
  if (PR && a.st->done) return;
  const int blk = blockIdx.x % kXcd;
  const int32_t* __restrict__ rp = a.browp + (int64_t)blk * (a.n + 1);
  float2* part = a.bpart + (int64_t)blk * a.n;
  const BSrc src{a.bcol, a.bval, a.x};
  const int m0 = a.bmega_off[blk];
  const int nmega = a.bmega_off[blk + 1] - m0;
  const int d0 = a.bmid_off[blk];
  const int nmid = a.bmid_off[blk + 1] - d0;
  const int nmidg = (nmid + SMDP_WAVES_PER_WG - 1) / SMDP_WAVES_PER_WG;
  constexpr int R = kRowsPerThread;
  const int rows_per_tile = SMDP_WG * R;
  const int ntiles = (a.n + rows_per_tile - 1) / rows_per_tile;
  const int nitems = nmega + nmidg + ntiles;
  __shared__ int s_item;
  __shared__ int s_start[SMDP_WG];
  __shared__ int s_off[SMDP_WG];
  __shared__ int s_scan[SMDP_WAVES_PER_WG];
  __shared__ float s_p[kWin];
  __shared__ float s_e[PR ? 1 : kWin];
  __shared__ float2 s_red[SMDP_WAVES_PER_WG];
  unsigned long long edges = 0;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&a.st->bwork[blk * 32], kGrab);
    __syncthreads();
    const int first = s_item;
    __syncthreads();
    if (first >= nitems) break;
    for (int item = first; item < min(first + kGrab, nitems); ++item) {
    if (item < nmega) {
      // a segment with >= split arcs: the whole workgroup
      const int r = a.bmega[m0 + item];
      const int s = a.bmega_pos[m0 + item];
      const int e = s + a.bmega_len[m0 + item];
      CSum h = bwg_row<PR>(src, s, e, s_red);
      if (threadIdx.x == 0) {
        part[r] = make_float2(h.s, h.c);
        edges += e - s;
      }
      continue;
    }
    if (item < nmega + nmidg) {
      // one medium segment per wave
      const int mi = (item - nmega) * SMDP_WAVES_PER_WG + (int)threadIdx.x / SMDP_WAVE;
      if (mi < nmid) {
        const int r = a.bmid[d0 + mi];
        const int s = a.bmid_pos[d0 + mi];
        const int e = s + a.bmid_len[d0 + mi];
        CSum h = bwave_row<PR>(src, s, e);
        if (lane_id() == 0) {
          part[r] = make_float2(h.s, h.c);
          edges += e - s;
        }
      }
      continue;
    }
    // R consecutive rows per thread: their arcs are one contiguous range of bcol
    const int r0 = (item - nmega - nmidg) * rows_per_tile + (int)threadIdx.x * R;
    int bound[R + 1];
#pragma unroll
    for (int i = 0; i <= R; ++i) bound[i] = rp[min(r0 + i, a.n)];
    const int s = bound[0];
    const int len = bound[R] - s;
    s_start[threadIdx.x] = s;
    int tot;
    const int ex = wg_exclusive_scan(len, s_scan, &tot);
    s_off[threadIdx.x] = ex;
    __syncthreads();
    int cur = 0;  // next row (0..R) whose partial is still open
    CSum acc;
    for (int wb = 0; wb < tot; wb += kWin) {
      const int wn = min(kWin, tot - wb);
      constexpr int K = kWin / SMDP_WG;
      int j[K], own[K], q[K], c[K];
#pragma unroll
      for (int i = 0; i < K; ++i) j[i] = min(wb + (int)threadIdx.x + i * SMDP_WG, wb + wn - 1);
      tile_owners<K>(s_off, j, own);
#pragma unroll
      for (int i = 0; i < K; ++i) q[i] = s_start[own[i]] + (j[i] - s_off[own[i]]);
#pragma unroll
      for (int i = 0; i < K; ++i) c[i] = src.col[q[i]];
      if (PR) {
        float g[K];
#pragma unroll
        for (int i = 0; i < K; ++i) g[i] = src.x[c[i]];
#pragma unroll
        for (int i = 0; i < K; ++i)
          if (wb + (int)threadIdx.x + i * SMDP_WG < wb + wn) s_p[j[i] - wb] = g[i];
      } else {
#pragma clang fp contract(off)
        float va[K], xb[K];
#pragma unroll
        for (int i = 0; i < K; ++i) {
          va[i] = src.val[q[i]];
          xb[i] = src.x[c[i]];
        }
#pragma unroll
        for (int i = 0; i < K; ++i) {
          if (wb + (int)threadIdx.x + i * SMDP_WG < wb + wn) {
            const float p = va[i] * xb[i];
            s_p[j[i] - wb] = p;
            s_e[j[i] - wb] = __builtin_fmaf(va[i], xb[i], -p);
          }
        }
      }
      __syncthreads();
      // my arcs inside this window, closing rows at their boundaries
      const int lo_q = max(ex, wb), hi_q = min(ex + len, wb + wn);
      for (int qq = lo_q; qq < hi_q; ++qq) {
        const int at = s + (qq - ex);  // arc index in bcol
        while (at >= bound[cur + 1]) {
          if (bound[cur + 1] > bound[cur]) part[r0 + cur] = make_float2(acc.s, acc.c);
          acc = CSum();
          ++cur;
        }
        acc.add(s_p[qq - wb]);
        if (!PR) acc.c += s_e[qq - wb];
      }
      __syncthreads();
    }
    // close the rows still open (all arcs consumed)
    for (; cur < R; ++cur)
      if (bound[cur + 1] > bound[cur]) {
        part[r0 + cur] = make_float2(acc.s, acc.c);
        acc = CSum();
      }
    edges += len;
    }
  }
  count_work(a.cnt, edges, 0, 0);
}

// Pass B: merge the kXcd partials of every row in block order, epilogue,
// FP64 tile partials; the last workgroup finishes the pass.
template <bool PR>
__global__ __launch_bounds__(SMDP_WG) void k_rows_blk_fin(RArgs a, int k) {
  if (PR && a.st->done) return;
  const float base = PR ? iteration_base(a, k) : 0.f;
  __shared__ double2 s_red2[SMDP_WAVES_PER_WG];
  __shared__ int s_last;
  unsigned long long rows = 0;
  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int r = tile * SMDP_WG + (int)threadIdx.x;
    double res = 0.0, dang = 0.0;
    if (r < a.n) {
      CSum acc;
      const unsigned mask = a.bmask[r];
#pragma unroll
      for (int b = 0; b < kXcd; ++b) {
        if (!(mask >> b & 1u)) continue;
        const float2 pb = a.bpart[(int64_t)b * a.n + r];
        CSum o;
        o.s = pb.x;
        o.c = pb.y;
        acc.merge(o);
      }
      epilogue<PR>(a, r, acc.value(), base, res, dang);
      ++rows;
    }
    if (PR) {
      double2 t = wg_sum2(make_double2(res, dang), s_red2);
      if (threadIdx.x == 0) a.tile_part[tile] = t;
    }
  }
  count_work(a.cnt, 0, rows, 0);
  if (arrive_last(a.st, &s_last)) finish_pass<PR>(a, k, make_double2(0.0, 0.0), s_red2);
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
    for (int i = threadIdx.x; i < (int)(kCntParts * sizeof(Counters) / 8); i += blockDim.x) c[i] = 0;
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
struct BlockPlan {
  int32_t split = -1;
  int32_t heavy = -1;
  int32_t val_version = -1;  // -1: built without values
  int32_t* browp = nullptr;
  int32_t* bcol = nullptr;
  float* bval = nullptr;
  int32_t* bmega = nullptr;
  int32_t* bmega_off = nullptr;
  int32_t* bmega_pos = nullptr;
  int32_t* bmega_len = nullptr;
  uint8_t* bmask = nullptr;
  int32_t* bmid = nullptr;
  int32_t* bmid_off = nullptr;
  int32_t* bmid_pos = nullptr;
  int32_t* bmid_len = nullptr;
  float2* bpart = nullptr;
};
struct RowBufs {
  smdp::RowPlan plan;
  BlockPlan blk;
  smdp::RState* st = nullptr;
  double2* tile_part = nullptr;
  double2* heavy_part = nullptr;
  int32_t* hq = nullptr;
  float2* mpart = nullptr;
  int64_t mpart_cap = 0;
  float* contrib2 = nullptr;
  int64_t ndang = -1;         // out-degree-0 vertices, counted for out-degrees version ndang_version
  int32_t ndang_version = -1;
  void* binned = nullptr;  // binned layout of PageRank (binned.hip)
  void* merge = nullptr;   // merge-path SpMV plan (merge.hip)
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
  void* ptrs[] = {b->st,          b->tile_part,   b->heavy_part,   b->hq,        b->mpart,
                  b->contrib2,    b->plan.mrows,  b->plan.mfirst,  b->blk.browp, b->blk.bcol,
                  b->blk.bval,    b->blk.bmega,   b->blk.bmega_off, b->blk.bpart,
                  b->blk.bmega_pos, b->blk.bmega_len, b->blk.bmask, b->blk.bmid, b->blk.bmid_off,
                  b->blk.bmid_pos, b->blk.bmid_len};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  binned_free(b->binned);
  merge_free(b->merge);
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

// The column-blocked layout: one CSR per block of n/kXcd source ids, each
// row keeping its arcs in their original order.  Built once per (graph,
// split, values) on the host and kept with the resident graph.
static int ensure_blocks(smdp_graph* g, RowBufs* rb, int32_t heavy, int32_t split, bool with_val) {
  BlockPlan& bp = rb->blk;
  const int32_t want_ver = with_val ? g->val_version : -1;
  if (bp.browp && bp.split == split && bp.heavy == heavy && bp.val_version == want_ver) return SMDP_OK;
  const int64_t n = g->n, m = g->m;
  std::vector<int32_t> row((size_t)n + 1), col((size_t)std::max<int64_t>(1, m));
  std::vector<float> val;
  SMDP_HIP(hipMemcpy(row.data(), g->row, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
  if (m) SMDP_HIP(hipMemcpy(col.data(), g->col, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  if (with_val) {
    val.resize((size_t)std::max<int64_t>(1, m));
    if (m) SMDP_HIP(hipMemcpy(val.data(), g->val, sizeof(float) * m, hipMemcpyDeviceToHost));
  }
  const int64_t width = (n + kXcd - 1) / kXcd;
  // segment lengths per (block, row)
  std::vector<int32_t> seg((size_t)kXcd * n, 0);
  for (int64_t r = 0; r < n; ++r)
    for (int32_t e = row[r]; e < row[r + 1]; ++e) seg[(size_t)(col[e] / width) * n + r]++;
  // layout in bcol: block 0 CSR (segments < split), block 0 long segments,
  // block 1 CSR, block 1 long segments, ...
  std::vector<int32_t> browp((size_t)kXcd * (n + 1));
  std::vector<int32_t> mega, mpos, mlen, moff(1, 0);
  std::vector<int32_t> mid, dpos, dlen, doff(1, 0);
  std::vector<uint8_t> mask((size_t)std::max<int64_t>(1, n), 0);
  std::vector<int32_t> wpos((size_t)kXcd * n);  // write cursor per (block, row)
  int64_t base = 0;
  for (int b = 0; b < kXcd; ++b) {
    int32_t* p = &browp[(size_t)b * (n + 1)];
    const int32_t* sl = &seg[(size_t)b * n];
    p[0] = (int32_t)base;
    for (int64_t r = 0; r < n; ++r) {
      const int32_t l = sl[r] < heavy ? sl[r] : 0;  // only light segments stay in the CSR
      wpos[(size_t)b * n + r] = p[r];
      p[r + 1] = p[r] + l;
      if (sl[r]) mask[r] |= (uint8_t)(1u << b);
    }
    base = p[n];
    std::vector<int32_t> rows, mids;
    for (int64_t r = 0; r < n; ++r) {
      if (sl[r] >= split) rows.push_back((int32_t)r);
      else if (sl[r] >= heavy) mids.push_back((int32_t)r);
    }
    std::stable_sort(rows.begin(), rows.end(), [&](int32_t x, int32_t y) { return sl[x] > sl[y]; });
    for (int32_t r : rows) {
      mega.push_back(r);
      mpos.push_back((int32_t)base);
      mlen.push_back(sl[r]);
      wpos[(size_t)b * n + r] = (int32_t)base;
      base += sl[r];
    }
    moff.push_back((int32_t)mega.size());
    for (int32_t r : mids) {
      mid.push_back(r);
      dpos.push_back((int32_t)base);
      dlen.push_back(sl[r]);
      wpos[(size_t)b * n + r] = (int32_t)base;
      base += sl[r];
    }
    doff.push_back((int32_t)mid.size());
  }
  std::vector<int32_t> bcol((size_t)std::max<int64_t>(1, m));
  std::vector<float> bval(with_val ? (size_t)std::max<int64_t>(1, m) : 0);
  for (int64_t r = 0; r < n; ++r)
    for (int32_t e = row[r]; e < row[r + 1]; ++e) {
      const int32_t at = wpos[(size_t)(col[e] / width) * n + r]++;
      bcol[at] = col[e];
      if (with_val) bval[at] = val[e];
    }
  void* old[] = {bp.browp,     bp.bcol,     bp.bval,  bp.bmega,    bp.bmega_off, bp.bmega_pos,
                 bp.bmega_len, bp.bmask,    bp.bmid,  bp.bmid_off, bp.bmid_pos,  bp.bmid_len};
  for (void* q : old)
    if (q) (void)hipFree(q);
  bp.browp = bp.bcol = bp.bmega = bp.bmega_off = bp.bmega_pos = bp.bmega_len = nullptr;
  bp.bmid = bp.bmid_off = bp.bmid_pos = bp.bmid_len = nullptr;
  bp.bval = nullptr;
  bp.bmask = nullptr;
  const int64_t nm = std::max<int64_t>(1, (int64_t)mega.size());
  SMDP_CHECK(alloc_once(&bp.browp, (int64_t)kXcd * (n + 1)));
  SMDP_CHECK(alloc_once(&bp.bcol, std::max<int64_t>(1, m)));
  if (with_val) SMDP_CHECK(alloc_once(&bp.bval, std::max<int64_t>(1, m)));
  SMDP_CHECK(alloc_once(&bp.bmega, nm));
  SMDP_CHECK(alloc_once(&bp.bmega_pos, nm));
  SMDP_CHECK(alloc_once(&bp.bmega_len, nm));
  SMDP_CHECK(alloc_once(&bp.bmega_off, kXcd + 1));
  SMDP_CHECK(alloc_once(&bp.bmask, std::max<int64_t>(1, n)));
  const int64_t nd = std::max<int64_t>(1, (int64_t)mid.size());
  SMDP_CHECK(alloc_once(&bp.bmid, nd));
  SMDP_CHECK(alloc_once(&bp.bmid_pos, nd));
  SMDP_CHECK(alloc_once(&bp.bmid_len, nd));
  SMDP_CHECK(alloc_once(&bp.bmid_off, kXcd + 1));
  if (!mid.empty()) {
    SMDP_HIP(hipMemcpy(bp.bmid, mid.data(), sizeof(int32_t) * mid.size(), hipMemcpyHostToDevice));
    SMDP_HIP(hipMemcpy(bp.bmid_pos, dpos.data(), sizeof(int32_t) * dpos.size(), hipMemcpyHostToDevice));
    SMDP_HIP(hipMemcpy(bp.bmid_len, dlen.data(), sizeof(int32_t) * dlen.size(), hipMemcpyHostToDevice));
  }
  SMDP_HIP(hipMemcpy(bp.bmid_off, doff.data(), sizeof(int32_t) * doff.size(), hipMemcpyHostToDevice));
  SMDP_CHECK(alloc_once(&bp.bpart, (int64_t)kXcd * std::max<int64_t>(1, n)));
  SMDP_HIP(hipMemcpy(bp.browp, browp.data(), sizeof(int32_t) * browp.size(), hipMemcpyHostToDevice));
  if (m) SMDP_HIP(hipMemcpy(bp.bcol, bcol.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice));
  if (with_val && m) SMDP_HIP(hipMemcpy(bp.bval, bval.data(), sizeof(float) * m, hipMemcpyHostToDevice));
  if (!mega.empty()) {
    SMDP_HIP(hipMemcpy(bp.bmega, mega.data(), sizeof(int32_t) * mega.size(), hipMemcpyHostToDevice));
    SMDP_HIP(hipMemcpy(bp.bmega_pos, mpos.data(), sizeof(int32_t) * mpos.size(), hipMemcpyHostToDevice));
    SMDP_HIP(hipMemcpy(bp.bmega_len, mlen.data(), sizeof(int32_t) * mlen.size(), hipMemcpyHostToDevice));
  }
  SMDP_HIP(hipMemcpy(bp.bmega_off, moff.data(), sizeof(int32_t) * moff.size(), hipMemcpyHostToDevice));
  if (n) SMDP_HIP(hipMemcpy(bp.bmask, mask.data(), n, hipMemcpyHostToDevice));
  bp.split = split;
  bp.heavy = heavy;
  bp.val_version = want_ver;
  return SMDP_OK;
}

// The blocked layout cuts the traffic beyond L2 about tenfold on RMAT-22
// PageRank but is not faster there on MI355X (its partial sums and per-row
// work cost as much as the misses saved), so it is used only when asked for,
// by the caller or by the autotuner.
static bool use_blocks(const smdp_params& p, int sol, int32_t) {
  return sol == SMDP_SOL_TUNED && p.layout == SMDP_LAYOUT_XCD8;
}

// The binned layout replaces every random gather of a PageRank iteration by
// streamed runs; its one-time regrouping is worth it on graphs with at least
// 2^20 arcs, where gathers miss L2.
static bool use_binned(const smdp_params& p, int sol, const smdp_graph* g) {
  if (sol != SMDP_SOL_TUNED) return false;
  return p.layout == SMDP_LAYOUT_BINNED || (p.layout == SMDP_LAYOUT_AUTO && g->m >= (1 << 20));
}

template <bool PR>
static void launch_pass(int sol, int grid, hipStream_t s, const RArgs& a, int k, int* launches, smdp_ctx* ctx) {
  dim3 g(grid), b(SMDP_WG);
  if (a.browp) {
    prof_mark(ctx, 0);
    hipLaunchKernelGGL((k_rows_blk<PR>), dim3((grid + kXcd - 1) / kXcd * kXcd), b, 0, s, a, k);
    prof_mark(ctx, 0);
    prof_mark(ctx, 1);
    hipLaunchKernelGGL((k_rows_blk_fin<PR>), g, b, 0, s, a, k);
    prof_mark(ctx, 1);
    *launches += 2;
    return;
  }
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
  if (use_blocks(p, sol, g->n)) {
    // segments are about 1/kXcd of a row: the heavy degree is scaled down
    const int32_t heavy = std::max(2, p.heavy_threshold / kXcd);
    const int32_t split = std::max(p.split_threshold / kXcd, heavy + 1);
    SMDP_CHECK(ensure_blocks(g, b, heavy, split, app == SMDP_APP_SPMV));
    a->browp = b->blk.browp;
    a->bcol = b->blk.bcol;
    a->bval = b->blk.bval;
    a->bpart = b->blk.bpart;
    a->bmega = b->blk.bmega;
    a->bmega_off = b->blk.bmega_off;
    a->bmega_pos = b->blk.bmega_pos;
    a->bmega_len = b->blk.bmega_len;
    a->bmask = b->blk.bmask;
    a->bmid = b->blk.bmid;
    a->bmid_off = b->blk.bmid_off;
    a->bmid_pos = b->blk.bmid_pos;
    a->bmid_len = b->blk.bmid_len;
    a->split = split;
  }
  *out = b;
  return SMDP_OK;
}

int run_spmv(smdp_graph* g, int sol, smdp_stats* st) {
  if (!g || !g->val || !g->x) return SMDP_ERR_ARG;
  smdp_ctx* ctx = g->ctx;
  SMDP_HIP(hipSetDevice(ctx->device));
  begin_run(g);
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
  // tuned solution on the plain CSR: merge-path SpMV (merge.hip)
  if (g->n > 0 && sol == SMDP_SOL_TUNED && !a.browp) SMDP_CHECK(merge_spmv(g, &b->merge, s, &launches));
  else if (g->n > 0) launch_pass<false>(sol, grid, s, a, 0, &launches, ctx);
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
  begin_run(g);
  RArgs a;
  RowBufs* b;
  SMDP_CHECK(prepare_args(g, SMDP_APP_PAGERANK, sol, &a, &b));
  SMDP_CHECK(alloc_once(&g->rank[0], g->n));
  SMDP_CHECK(alloc_once(&g->rank[1], g->n));
  SMDP_CHECK(alloc_once(&g->contrib, g->n));
  SMDP_CHECK(alloc_once(&b->contrib2, g->n));
  if (b->ndang < 0 || b->ndang_version != g->outdeg_version) {
    std::vector<int32_t> od((size_t)g->n);
    SMDP_HIP(hipMemcpy(od.data(), g->outdeg, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost));
    b->ndang = std::count(od.begin(), od.end(), 0);
    b->ndang_version = g->outdeg_version;
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
  const smdp_params& p = params_of(ctx, SMDP_APP_PAGERANK);
  const bool binned = use_binned(p, sol, g);
  BArgs ba{};
  if (binned) {
    int rc = binned_ensure(g, &b->binned, &ba);
    if (rc != SMDP_OK && p.layout == SMDP_LAYOUT_BINNED) return rc;
    if (rc == SMDP_OK) {
      a.tile_part = ba.bpart;
      a.ntiles = ba.nparts;
    }
  }
  const bool use_bins = binned && ba.vals != nullptr;
  prof_reset(ctx);
  const int grid = grid_size(ctx, p);
  int launches = 0;
  SMDP_HIP(hipEventRecord(ctx->ev_begin, s));
  hipLaunchKernelGGL(k_pr_init, dim3(std::min<int64_t>((g->n + 255) / 256, 4096)), dim3(256), 0, s, g->rank[0],
                     cbuf[0], (const int32_t*)g->outdeg, g->n, init, b->st, dang0, g->cnt);
  ++launches;
  // The last workgroup of every iteration applies the stop rule and reports
  // the iteration to host memory; the host keeps a few iterations queued
  // ahead of the device and stops enqueueing when one reports that it
  // stopped (iterations enqueued after it return at once).  A blocking check
  // per batch of iterations left the device idle for the round trip each
  // time.
  int k = 0;
  int32_t done = 0;
  volatile int32_t* prog = ctx->progress;
  *prog = 0;
  a.progress = ctx->progress_dev;
  auto enqueue = [&](int cnt) {
    for (int j = 0; j < cnt && k < o->max_iter; ++j, ++k) {
      a.x = cbuf[k & 1];
      a.cnext = cbuf[(k + 1) & 1];
      a.rold = g->rank[k & 1];
      a.y = g->rank[(k + 1) & 1];
      if (use_bins) binned_pass(ctx, s, a, ba, k, &launches);
      else launch_pass<true>(sol, grid, s, a, k, &launches, ctx);
    }
  };
  enqueue(kFeedFirst);
  SMDP_HIP(hipGetLastError());
  int seen = 0;
  double t_seen = now_ms();
  for (;;) {
    const int v = *prog;
    if (v < 0) {
      done = -v;
      break;
    }
    if (k >= o->max_iter) {
      // every iteration is enqueued; the last one stops at the latest
      SMDP_HIP(hipStreamSynchronize(s));
      done = *prog < 0 ? -*prog : 0;
      break;
    }
    if (v >= k - kFeedSlack) {
      enqueue(kFeed);
      SMDP_HIP(hipGetLastError());
      continue;
    }
    if (v != seen) {
      seen = v;
      t_seen = now_ms();
    } else if (now_ms() - t_seen > 100.0) {
      // no progress report for a while: a fault stops the stream
      const hipError_t e = hipStreamQuery(s);
      if (e != hipSuccess && e != hipErrorNotReady) return hip_fail(e, "PageRank iterations", __LINE__);
      t_seen = now_ms();
    }
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
