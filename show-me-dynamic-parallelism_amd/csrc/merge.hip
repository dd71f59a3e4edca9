// Merge-path CSR SpMV, the tuned solution of y = A x.
//
// The rows' end offsets and the nonzeros form one merged sequence of n + m
// items (a row end comes before nonzero j when the row ends at or before j).
// Every workgroup takes kMpItems consecutive items of it, so the work per
// workgroup is the same whatever the row lengths: no row is too long for a
// workgroup and no workgroup waits on one long row.  This is the published
// merge-based SpMV (Merrill and Garland, "Merge-based parallel sparse matrix-
// vector multiplication", SC16), built here for 64-lane waves:
//
//  * the tile's nonzeros are multiplied with all x gathers of a thread in
//    flight and the products (with their rounding errors) staged in LDS,
//  * each thread walks kMpT items of the tile with a compensated sum,
//  * a row crossing threads is finished by the thread that reaches its end,
//    which merges the partial sums of the threads before it in thread order,
//  * a row crossing tiles is finished by k_merge_fixup, which merges the tile
//    partials in tile order.
//
// Every row is summed in an order fixed by the matrix alone, so results are
// identical from run to run.  Replaces csr_spmv_* (SpMV_kernel.cu:81-403) for
// the tuned solution.

#include <algorithm>
#include <vector>

#include "rowreduce.h"

namespace smdp {
namespace {

constexpr int kMpNT = 256;                // threads per workgroup
#ifndef SMDP_MERGE_ITEMS
#define SMDP_MERGE_ITEMS 16
#endif
constexpr int kMpT = SMDP_MERGE_ITEMS;    // merge items per thread
constexpr int kMpItems = kMpNT * kMpT;    // merge items per tile

#ifndef SMDP_MERGE_RELABEL
#define SMDP_MERGE_RELABEL 1
#endif

struct MergePlan {
  int32_t* tstart = nullptr;     // ntiles + 1: row ends before each tile's first item
  int32_t* pcol = nullptr;       // m: the columns renumbered by falling in-degree (ties by id)
  int32_t* perm = nullptr;       // n: renumbered column -> original column
  float* xp = nullptr;           // n: x in the renumbered order, of the columns that occur
  int32_t ncols = 0;             // columns that occur in the matrix (the first ncols renumbered ids)
  float2* carry = nullptr;       // per tile: partial sum of the row open at its end
  int32_t* carry_row = nullptr;  // per tile: that row
  float2* head = nullptr;        // per tile: its part of a row begun in an earlier tile and ended in it
  int32_t* head_row = nullptr;   // per tile: that row, or -1
  int32_t ntiles = 0;
  int32_t n = -1;
  int32_t m = -1;
};

struct MArgs {
  const int32_t* __restrict__ row;
  const int32_t* __restrict__ col;
  const float* __restrict__ val;
  const float* __restrict__ x;
  float* y;
  const int32_t* __restrict__ tstart;
  float2* carry;
  int32_t* carry_row;
  float2* head;
  int32_t* head_row;
  Counters* cnt;
  int32_t n, m, ntiles;
};

// Number of row ends among the first d merge items: the smallest i with
// end(i) > d - i - 1, searched in [max(0, d - nz), min(d, nr)].
template <typename End>
__device__ __forceinline__ int64_t merge_search(End end, int64_t d, int64_t nr, int64_t nz) {
  int64_t lo = d - nz > 0 ? d - nz : 0;
  int64_t hi = d < nr ? d : nr;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (end(mid) <= d - mid - 1) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_merge_starts(const int32_t* __restrict__ row, int32_t n, int32_t m, int32_t ntiles,
                               int32_t* tstart) {
  const int64_t total = (int64_t)n + m;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles; t += gridDim.x * blockDim.x) {
    const int64_t d = (int64_t)t * kMpItems < total ? (int64_t)t * kMpItems : total;
    tstart[t] = (int32_t)merge_search([&](int64_t i) { return (int64_t)row[i + 1]; }, d, n, m);
  }
}

__global__ __launch_bounds__(kMpNT) void k_spmv_merge(MArgs a) {
  // The tile's nr row ends and nz nonzeros fill its kMpItems merge items, so
  // one array holds both: slot k < nr the end of row i0 + k (in tile nonzero
  // positions, as int bits in .x), slot nr + j the product of nonzero j and
  // its rounding error.  At 16 items per thread, 34 KB of LDS per workgroup: 4 workgroups per CU.
  __shared__ float2 s_item[kMpItems];
  __shared__ float2 s_cs[kMpNT];       // per thread: partial sum of its open row, scanned over the row's threads
  __shared__ int32_t s_hk[kMpNT / SMDP_WAVE], s_tk[kMpNT / SMDP_WAVE];  // per wave: open row of lanes 0 and 63
  __shared__ float2 s_tv[kMpNT / SMDP_WAVE];                           // per wave: scanned sum of lane 63
  const int tid = threadIdx.x;
  const int t = blockIdx.x;
  const int64_t d0 = (int64_t)t * kMpItems;
  const int64_t dn = (int64_t)a.n + a.m;
  const int64_t d1 = d0 + kMpItems < dn ? d0 + kMpItems : dn;
  const int i0 = a.tstart[t], i1 = a.tstart[t + 1];
  const int j0 = (int)(d0 - i0), j1 = (int)(d1 - i1);
  const int nr = i1 - i0, nz = j1 - j0;
  // the tile's first row began in an earlier tile
  const bool mid_row = i0 < a.n && a.row[i0] < j0;
  if (tid == 0) {
    a.head_row[t] = (mid_row && nr > 0) ? i0 : -1;
    Counters* c = cnt_of(a.cnt);
    if (nz) atomicAdd(&c->edges, (unsigned long long)nz);
    if (nr) atomicAdd(&c->verts, (unsigned long long)nr);
  }
  // Every load of the tile is issued before the first one is used.  Loads go
  // to clamped positions inside the tile (the branches on nz and nr are the
  // same for the whole workgroup): a load under a per-lane condition would
  // make the compiler wait for it before the other lanes' default value is
  // written to the same register.
  if (nz > 0) {
    int c[kMpT], re[kMpT];
    float v[kMpT], xv[kMpT];
#pragma unroll
    for (int u = 0; u < kMpT; ++u) {
      const int k = min(tid + u * kMpNT, nz - 1);
      c[u] = stream_load(a.col + j0 + k);
      v[u] = stream_load(a.val + j0 + k);
    }
    if (nr > 0) {
#pragma unroll
      for (int u = 0; u < kMpT; ++u) re[u] = a.row[i0 + min(tid + u * kMpNT, nr - 1) + 1];
    }
#pragma unroll
    for (int u = 0; u < kMpT; ++u) xv[u] = a.x[c[u]];
    if (nr > 0) {
#pragma unroll
      for (int u = 0; u < kMpT; ++u) {
        const int k = tid + u * kMpNT;
        if (k < nr) s_item[k].x = __int_as_float(re[u] - j0);
      }
    }
#pragma unroll
    for (int u = 0; u < kMpT; ++u) {
      const int k = tid + u * kMpNT;
      if (k < nz) {
#pragma clang fp contract(off)
        const float p = v[u] * xv[u];
        s_item[nr + k] = make_float2(p, __builtin_fmaf(v[u], xv[u], -p));
      }
    }
  } else if (nr > 0) {
    int re[kMpT];
#pragma unroll
    for (int u = 0; u < kMpT; ++u) re[u] = a.row[i0 + min(tid + u * kMpNT, nr - 1) + 1];
#pragma unroll
    for (int u = 0; u < kMpT; ++u) {
      const int k = tid + u * kMpNT;
      if (k < nr) s_item[k].x = __int_as_float(re[u] - j0);
    }
  }
  __syncthreads();

  const int dl = min(tid * kMpT, nr + nz), dr = min(dl + kMpT, nr + nz);
  int ri;
  {
    int lo = max(0, dl - nz), hi = min(dl, nr);
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (__float_as_int(s_item[mid].x) <= dl - mid - 1) lo = mid + 1;
      else hi = mid;
    }
    ri = lo;
  }
  int zi = dl - ri;
  const int rfirst = ri;
  CSum acc, fpart;
  bool fdone = false;
  for (int q = dl; q < dr; ++q) {
    if (ri < nr && __float_as_int(s_item[ri].x) <= zi) {
      if (ri == rfirst) {
        fpart = acc;
        fdone = true;
      } else {
        a.y[i0 + ri] = acc.value();
      }
      acc = CSum();
      ++ri;
    } else {
      const float2 pe = s_item[nr + zi];
      acc.add(pe.x);
      acc.c += pe.y;
      ++zi;
    }
  }
  // The open rows (keys) do not decrease from thread to thread, so the
  // threads ending inside one row are consecutive.  An inclusive segmented
  // scan gives every thread the sum of its row's partials from the first of
  // those threads up to itself: six shuffle steps in the wave, then the tails
  // of at most three earlier waves.  The combination order is fixed.
  const int key = ri;
  const int lane = lane_id(), wave = tid / SMDP_WAVE;
  CSum sc = acc;
#pragma unroll
  for (int d = 1; d < SMDP_WAVE; d <<= 1) {
    const int k2 = __shfl_up(key, d);
    CSum o;
    o.s = __shfl_up(sc.s, d);
    o.c = __shfl_up(sc.c, d);
    if (lane >= d && k2 == key) {
      o.merge(sc);
      sc = o;
    }
  }
  if (lane == 0) s_hk[wave] = key;
  if (lane == SMDP_WAVE - 1) {
    s_tk[wave] = key;
    s_tv[wave] = make_float2(sc.s, sc.c);
  }
  __syncthreads();
  if (key == s_hk[wave]) {
    int v = wave;
    while (v > 0 && s_tk[v - 1] == key) {
      --v;
      if (s_hk[v] != key) break;
    }
    if (v < wave) {
      CSum pre;
      for (int u = v; u < wave; ++u) {
        CSum o;
        o.s = s_tv[u].x;
        o.c = s_tv[u].y;
        pre.merge(o);
      }
      pre.merge(sc);
      sc = pre;
    }
  }
  s_cs[tid] = make_float2(sc.s, sc.c);
  __syncthreads();

  if (fdone) {
    // thread tid - 1 ended inside rfirst: its scanned sum holds every earlier
    // thread's part of the row
    CSum tot;
    if (tid > 0) {
      tot.s = s_cs[tid - 1].x;
      tot.c = s_cs[tid - 1].y;
    }
    tot.merge(fpart);
    if (rfirst == 0 && mid_row) a.head[t] = make_float2(tot.s, tot.c);
    else a.y[i0 + rfirst] = tot.value();
  }
  if (tid == kMpNT - 1) {
    a.carry[t] = make_float2(sc.s, sc.c);
    a.carry_row[t] = i1;
  }
}

// Column renumbering.  RMAT columns are heavy-tailed: on RMAT-20 the 8192
// most frequent columns (32 KB of x) take 48.5% of the nonzeros and 48% of
// the columns never occur.  Numbered by falling in-degree, the frequent x
// values share a few cache lines that stay in each CU's L1, instead of one
// line each.  Every product and every row's order are unchanged, so y is
// bit for bit the same as with the original columns.
__global__ void k_col_degree(const int32_t* __restrict__ col, int64_t m, int32_t* deg) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    __hip_atomic_fetch_add(deg + col[e], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_renumber_cols(const int32_t* __restrict__ col, int64_t m, const int32_t* __restrict__ inv,
                                int32_t* __restrict__ pcol) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    pcol[e] = inv[col[e]];
}

// xp[i] = x[perm[i]] for the ncols columns that occur.
__global__ void k_permute_x(const float* __restrict__ x, const int32_t* __restrict__ perm, int32_t ncols,
                            float* __restrict__ xp) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncols; i += gridDim.x * blockDim.x) xp[i] = x[perm[i]];
}

// Rows that began in an earlier tile: the partials of the tiles they cross,
// merged in tile order, then the part of the tile where they end.
__global__ void k_merge_fixup(MArgs a) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < a.ntiles; t += gridDim.x * blockDim.x) {
    const int r = a.head_row[t];
    if (r < 0) continue;
    int k = t;
    while (k > 0 && a.carry_row[k - 1] == r) --k;
    CSum tot;
    for (int q = k; q < t; ++q) {
      CSum o;
      o.s = a.carry[q].x;
      o.c = a.carry[q].y;
      tot.merge(o);
    }
    CSum h;
    h.s = a.head[t].x;
    h.c = a.head[t].y;
    tot.merge(h);
    a.y[r] = tot.value();
  }
}

}  // namespace

void merge_free(void* p) {
  MergePlan* mp = static_cast<MergePlan*>(p);
  if (!mp) return;
  void* ptrs[] = {mp->tstart, mp->carry, mp->carry_row, mp->head, mp->head_row, mp->pcol, mp->perm, mp->xp};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete mp;
}

// Renumbers the columns by falling in-degree (ties by id): degrees counted
// on the device, a counting sort by degree on the host, the renumbered
// column array written on the device.  Once per matrix.
static int renumber_columns(smdp_graph* g, MergePlan* mp, hipStream_t s) {
  const int64_t n = g->n, m = g->m;
  const int grid = std::max(1, g->ctx->cus * 8);
  DeviceScratch<int32_t> tmp;
  SMDP_CHECK(alloc_once(&tmp.p, n));
  int32_t* deg = tmp.p;
  SMDP_HIP(hipMemsetAsync(deg, 0, sizeof(int32_t) * n, s));
  hipLaunchKernelGGL(k_col_degree, dim3(grid), dim3(256), 0, s, g->col, m, deg);
  std::vector<int32_t> hd((size_t)n);
  SMDP_HIP(hipMemcpyAsync(hd.data(), deg, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  SMDP_HIP(hipStreamSynchronize(s));
  int32_t dmax = 0;
  for (int64_t v = 0; v < n; ++v) dmax = std::max(dmax, hd[(size_t)v]);
  // start of each degree's run, highest degree first
  std::vector<int64_t> at((size_t)dmax + 2, 0);
  for (int64_t v = 0; v < n; ++v) ++at[(size_t)(dmax - hd[(size_t)v]) + 1];
  for (size_t d = 1; d < at.size(); ++d) at[d] += at[d - 1];
  std::vector<int32_t> perm((size_t)n), inv((size_t)n);
  for (int64_t v = 0; v < n; ++v) {
    const int64_t k = at[(size_t)(dmax - hd[(size_t)v])]++;
    perm[(size_t)k] = (int32_t)v;
    inv[(size_t)v] = (int32_t)k;
  }
  int64_t ncols = 0;
  for (int64_t v = 0; v < n; ++v) ncols += hd[(size_t)v] > 0;
  SMDP_CHECK(alloc_once(&mp->pcol, m));
  SMDP_CHECK(alloc_once(&mp->perm, n));
  SMDP_CHECK(alloc_once(&mp->xp, std::max<int64_t>(1, ncols)));
  // deg's memory holds inv while the columns are renumbered
  SMDP_HIP(hipMemcpyAsync(deg, inv.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  SMDP_HIP(hipMemcpyAsync(mp->perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_renumber_cols, dim3(grid), dim3(256), 0, s, g->col, m, deg, mp->pcol);
  SMDP_HIP(hipGetLastError());
  SMDP_HIP(hipStreamSynchronize(s));
  mp->ncols = (int32_t)ncols;
  return SMDP_OK;
}

int merge_spmv(smdp_graph* g, void** slot, hipStream_t s, int* launches) {
  const int64_t total = (int64_t)g->n + g->m;
  if (g->n <= 0 || total > (int64_t)INT32_MAX - kMpItems) return SMDP_ERR_ARG;
  MergePlan* mp = static_cast<MergePlan*>(*slot);
  if (!mp || mp->n != g->n || mp->m != g->m) {
    merge_free(mp);
    *slot = nullptr;
    mp = new MergePlan();
    *slot = mp;
    mp->ntiles = (int32_t)((total + kMpItems - 1) / kMpItems);
    const int64_t nt = std::max<int64_t>(1, mp->ntiles);
    SMDP_CHECK(alloc_once(&mp->tstart, nt + 1));
    SMDP_CHECK(alloc_once(&mp->carry, nt));
    SMDP_CHECK(alloc_once(&mp->carry_row, nt));
    SMDP_CHECK(alloc_once(&mp->head, nt));
    SMDP_CHECK(alloc_once(&mp->head_row, nt));
    hipLaunchKernelGGL(k_merge_starts, dim3((mp->ntiles + 256) / 256), dim3(256), 0, s, g->row, g->n, g->m,
                       mp->ntiles, mp->tstart);
    SMDP_HIP(hipGetLastError());
    if (SMDP_MERGE_RELABEL && g->m > 0) SMDP_CHECK(renumber_columns(g, mp, s));
    mp->n = g->n;
    mp->m = g->m;
  }
  MArgs a;
  a.row = g->row;
  a.col = g->col;
  a.val = g->val;
  a.x = g->x;
  a.y = g->y;
  a.tstart = mp->tstart;
  a.carry = mp->carry;
  a.carry_row = mp->carry_row;
  a.head = mp->head;
  a.head_row = mp->head_row;
  a.cnt = g->cnt;
  a.n = g->n;
  a.m = g->m;
  a.ntiles = mp->ntiles;
  if (mp->ntiles == 0) return SMDP_OK;
  prof_mark(g->ctx, 0);
  if (mp->pcol) {
    // x in the renumbered order (part of the timed product)
    hipLaunchKernelGGL(k_permute_x, dim3(std::max(1, std::min((mp->ncols + 255) / 256, g->ctx->cus * 8))), dim3(256),
                       0, s, g->x, mp->perm, mp->ncols, mp->xp);
    a.col = mp->pcol;
    a.x = mp->xp;
    ++*launches;
  }
  hipLaunchKernelGGL(k_spmv_merge, dim3(mp->ntiles), dim3(kMpNT), 0, s, a);
  prof_mark(g->ctx, 0);
  prof_mark(g->ctx, 1);
  hipLaunchKernelGGL(k_merge_fixup, dim3((mp->ntiles + 255) / 256), dim3(256), 0, s, a);
  prof_mark(g->ctx, 1);
  *launches += 2;
  SMDP_HIP(hipGetLastError());
  return SMDP_OK;
}

}  // namespace smdp
