// Shared device state and helpers of the row-reduce family (CSR SpMV and
// PageRank): the pass state, the kernel argument block, the per-row epilogue,
// compensated workgroup sums and the last-workgroup finish of a pass.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "smdp_internal.h"

namespace smdp {

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
  int32_t bwork[8 * 32];  // column-blocked passes: work counter of each block, one 128-byte line each
};

constexpr int kGrab = 8;  // work items claimed per atomic

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
  int32_t* progress;               // PageRank: host memory, k + 1 after iteration k, -(k + 1) when it stopped
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
  // column-blocked layout (kXcd blocks of source ids)
  const int32_t* __restrict__ browp;     // kXcd x (n+1) offsets into bcol
  const int32_t* __restrict__ bcol;
  const float* __restrict__ bval;
  float2* bpart;                         // kXcd x n partial sums (sum, error)
  const int32_t* __restrict__ bmega;     // per block: rows whose segment has >= split arcs, longest first
  const int32_t* __restrict__ bmega_off; // kXcd + 1
  const int32_t* __restrict__ bmega_pos; // start in bcol of each such segment (stored apart from the CSR)
  const int32_t* __restrict__ bmega_len;
  const uint8_t* __restrict__ bmask;     // per row: bit b set when block b has arcs for it
  const int32_t* __restrict__ bmid;      // per block: rows whose segment has [heavy, split) arcs
  const int32_t* __restrict__ bmid_off;  // kXcd + 1
  const int32_t* __restrict__ bmid_pos;
  const int32_t* __restrict__ bmid_len;
};

// Binned layout of PageRank's in-arcs (built once per graph, host side).
// Sources are cut into chunks of 2^15 ids and rows into bins of consecutive
// rows; a segment holds the arcs of one (chunk, bin) pair.  `src` keeps the
// arcs chunk-major (the scatter pass reads it), `vals` and `dst` keep them
// bin-major (the reduce pass reads them), so each side streams.
struct BArgs {
  const int32_t* __restrict__ cm;      // nchunks x (nbins+1): start of segment (c, b) in src; [c][nbins] = end of c
  const int32_t* __restrict__ bm;      // nchunks x nbins: start of segment (c, b) in vals and dst
  const uint16_t* __restrict__ src;    // chunk-major: source id minus its chunk's first id
  const uint16_t* __restrict__ dst;    // bin-major: the arc's position in its bin, in in-CSR order
  float* vals;                         // bin-major: the source's contribution, written by the scatter pass
  const uint16_t* __restrict__ rperm;  // per row: the rows of its bin by falling degree (index in the bin)
  const int32_t* __restrict__ s_item;  // scatter items: (chunk, first bin, end bin, first arc, end arc in src)
  const int32_t* __restrict__ s_group; // scatter workgroup g takes items s_group[g] .. s_group[g+1]-1
  const int4* __restrict__ r_item;     // reduce items, two int4 each: (beg, end, r0, nr), (nwave, ngroup, part, 0);
                                       // nr < 0 marks a piece of a wide row, part indexes wpart, else bpart

  const int32_t* __restrict__ wfirst;  // nwide+1: first piece of each wide row
  const int32_t* __restrict__ wrow;    // nwide: the wide rows
  float2* wpart;                       // per wide piece: (sum, error)
  double2* bpart;                      // per bin: (residual, dangling)
  int32_t nchunks;
  int32_t nbins;
  int32_t nsgroups;                    // scatter workgroups
  int32_t nitems;                      // reduce items (claimed from RState::work_next)
  int32_t nwide;
  int32_t nparts;                      // bins of rows (entries of bpart)
};

constexpr int kXcd = 8;  // MI355X: 8 XCDs, each with its own L2; workgroup i runs on XCD i % 8

template <bool PR>
__device__ __forceinline__ void add_term(const RArgs& a, CSum& acc, int e) {
  if (PR) acc.add(a.x[a.col[e]]);
  else acc.add_prod(a.val[e], a.x[a.col[e]]);
}

// PageRank epilogue of row r with its old rank and out-degree already loaded.
__device__ __forceinline__ void pr_epilogue(const RArgs& a, int r, float sum, float base, float rold, int od,
                                            double& res, double& dang) {
  const float nr = __builtin_fmaf(a.damping, sum, base);
  a.y[r] = nr;
  a.cnext[r] = od > 0 ? nr / (float)od : 0.f;
  res += fabs((double)nr - (double)rold);
  if (od == 0) dang += (double)nr;
}

// Per-row epilogue in the owning thread.
template <bool PR>
__device__ __forceinline__ void epilogue(const RArgs& a, int r, float sum, float base, double& res, double& dang) {
  if (!PR) {
    a.y[r] = sum;
    return;
  }
  pr_epilogue(a, r, sum, base, a.rold[r], a.outdeg[r], res, dang);
}

template <int NT = SMDP_WG>
__device__ __forceinline__ double2 wg_sum2(double2 v, double2* lds) {
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  const int wave = threadIdx.x / SMDP_WAVE;
  if (lane_id() == 0) lds[wave] = v;
  __syncthreads();
  double2 t = make_double2(0.0, 0.0);
  for (int k = 0; k < NT / SMDP_WAVE; ++k) {
    t.x += lds[k].x;
    t.y += lds[k].y;
  }
  __syncthreads();
  return t;
}

template <int NT = SMDP_WG>
__device__ __forceinline__ CSum wg_csum(CSum v, float2* lds) {
  v = wave_csum(v);
  const int wave = threadIdx.x / SMDP_WAVE;
  if (lane_id() == 0) lds[wave] = make_float2(v.s, v.c);
  __syncthreads();
  CSum t;
  for (int k = 0; k < NT / SMDP_WAVE; ++k) {
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
  int k = s + lane_id();
  // four steps' loads are issued before their terms are added (same order)
  for (; k + 3 * SMDP_WAVE < e; k += 4 * SMDP_WAVE) {
    int c[4];
    float g[4], va[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) c[t] = stream_load(a.col + k + t * SMDP_WAVE);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      g[t] = a.x[c[t]];
      va[t] = PR ? 0.f : stream_load(a.val + k + t * SMDP_WAVE);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (PR) acc.add(g[t]);
      else acc.add_prod(va[t], g[t]);
    }
  }
  for (; k < e; k += SMDP_WAVE) add_term<PR>(a, acc, k);
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
  int k = s + (int)threadIdx.x;
  // four independent gathers per step, added in arc order
  for (; k + 3 * SMDP_WG < e; k += 4 * SMDP_WG) {
    int c[4];
    float g[4], va[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) c[t] = stream_load(a.col + k + t * SMDP_WG);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      g[t] = a.x[c[t]];
      va[t] = PR ? 0.f : stream_load(a.val + k + t * SMDP_WG);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (PR) acc.add(g[t]);
      else acc.add_prod(va[t], g[t]);
    }
  }
  for (; k < e; k += SMDP_WG) add_term<PR>(a, acc, k);
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
template <bool PR, int NT = SMDP_WG>
__device__ void finish_pass(const RArgs& a, int k, double2 extra, double2* lds2) {
  double2 acc = extra;
  if (PR)
    for (int t = threadIdx.x; t < a.ntiles; t += NT) {
      acc.x += a.tile_part[t].x;
      acc.y += a.tile_part[t].y;
    }
  double2 tot = wg_sum2<NT>(acc, lds2);
  if (threadIdx.x == 0) {
    RState* st = a.st;
    if (PR) {
      st->residual = tot.x;
      st->dang[(k + 1) & 1] = tot.y;
      const bool stop = tot.x < a.tol_abs || k + 1 >= a.max_iter;
      if (stop) st->done = k + 1;
      if (a.progress)
        __hip_atomic_store(a.progress, stop ? -(k + 1) : k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    st->ticket = 0;
    st->work_next = 0;
    st->hlen = 0;
    st->hnext = 0;
    for (int i = 0; i < 8; ++i) st->bwork[i * 32] = 0;
    atomicMax(&cnt_of(a.cnt)->rounds, (unsigned long long)(k + 1));
  }
}

__device__ __forceinline__ void count_work(Counters* cnt, unsigned long long edges, unsigned long long rows,
                                           unsigned long long pushes) {
  edges = wave_sum(edges);
  rows = wave_sum(rows);
  pushes = wave_sum(pushes);
  if (lane_id() == 0) {
    if (edges) atomicAdd(&cnt_of(cnt)->edges, edges);
    if (rows) atomicAdd(&cnt_of(cnt)->verts, rows);
    if (pushes) atomicAdd(&cnt_of(cnt)->push[1], pushes);
  }
}

// Host side of the binned layout (binned.hip).
int binned_ensure(smdp_graph* g, void** slot, BArgs* out);
void binned_free(void* plan);
void binned_pass(smdp_ctx* ctx, hipStream_t s, const RArgs& a, const BArgs& b, int k, int* launches);

// Merge-path SpMV (merge.hip): y = A x over g's CSR, values and x, with the
// per-graph plan kept in *slot (built on first use).
int merge_spmv(smdp_graph* g, void** slot, hipStream_t s, int* launches);
void merge_free(void* plan);

}  // namespace smdp
