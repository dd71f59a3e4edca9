// Tree Height and Tree Descendant.
//
// The reference's README.md:12-13 lists both as recursive benchmarks
// ("Tree Height (recursive)", "Tree Descendant (recursive)") but ships no
// code for them.  With dynamic parallelism each node launches a child grid
// over its children and combines their answers on return.  Here the
// recursion is unrolled into two level-synchronous passes over the tree held
// as child lists (the frontier layout of the BFS):
//  * down (k_tree_down, one launch per level): the slots of level L are
//    order[lo[L] .. lo[L+1]).  Each workgroup takes a contiguous share of the
//    level, sums its vertices' child counts, reserves that many slots of level
//    L+1 with one atomic, and writes every child into its slot, expanding the
//    share in edge-balanced tiles (a hub's 4096 children are spread over all
//    lanes of the workgroup).  A child without children of its own is a leaf:
//    it gets height 0 and no descendants on the spot and takes no slot, so no
//    later pass reads it (on the 4M-node Zipf tree 4.1M of the 4.2M vertices
//    are leaves; level 4 holds 2.48M vertices of which 32.6K have children).
//    The share's slot count is the sum of inner[u], each vertex's arcs to
//    vertices that have arcs (counted once per graph); waves take positions
//    in the reserved block from an LDS cursor.  An arc into the root, when
//    the root has arcs, keeps its slot as a hole (-1); a forest has at most
//    one.  The slots each level reserves are
//    counted per level (rel), so a launch finds where its level ends from the
//    one before without waiting for a last workgroup.
//  * up (k_tree_up, one launch per level, deepest first): every vertex of
//    level L pulls 1 + height[c] and 1 + desc[c] of its children (written by
//    the down pass for leaves, by the launch for level L+1 otherwise) into LDS
//    accumulators with integer atomics (max and add), in the same
//    edge-balanced tiles, and stores its results.
// Integer max and add do not depend on order, so results are exact and the
// same for every schedule.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "smdp_internal.h"

namespace smdp {
namespace {

constexpr int kTreeIlp = 4;  // arcs per thread per step
constexpr int kTreeCtl = 32; // control words after the level starts (own 128-byte line)

struct TArgs {
  const int32_t* __restrict__ row;
  const int32_t* __restrict__ col;
  const uint32_t* __restrict__ hasout;  // bit v set when v has children
  const int32_t* __restrict__ inner;    // inner[u]: arcs of u to vertices that have children
  int32_t* order;   // reached vertices with children (and holes, -1) level by level
  int32_t* lo;      // lo[L]: first slot of level L (written by the launch for level L - 1)
  int32_t* rel;     // rel[L]: slots reserved so far for level L + 1 (zeroed by the launch for level L - 1)
  int32_t* ctl;     // [1] number of levels, once a launch finds its level empty
  int32_t* height;
  int32_t* desc;
  Counters* cnt;
  int32_t root;
  int32_t n;
};

__global__ void k_tree_init(TArgs a) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += (int64_t)gridDim.x * blockDim.x) {
    a.height[v] = -1;
    a.desc[v] = -1;
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      a.order[0] = a.root;
      a.lo[0] = 0;
      a.lo[1] = 1;
      a.rel[0] = 0;
      a.ctl[1] = 0;
    }
    unsigned long long* k = reinterpret_cast<unsigned long long*>(a.cnt);
    for (int i = threadIdx.x; i < (int)(kCntParts * sizeof(Counters) / 8); i += blockDim.x) k[i] = 0;
  }
}

// inner[u] = arcs of u to vertices that have arcs (once per graph, inner[]
// zeroed before).  A thread takes kInnerSeg consecutive arcs, finds the
// first one's row by binary search and adds one count per row it touches, so
// a 4096-arc hub is counted by 256 threads instead of one (one thread per
// vertex took 674 us on the 4M-node Zipf tree).
constexpr int kInnerSeg = 16;
__global__ void k_tree_inner(const int32_t* __restrict__ row, const int32_t* __restrict__ col,
                             const uint32_t* __restrict__ hasout, int32_t n, int32_t* inner) {
  const int64_t m = row[n];
  const int64_t nseg = (m + kInnerSeg - 1) / kInnerSeg;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nseg; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e0 = t * kInnerSeg, e1 = min(e0 + kInnerSeg, m);
    int lo = 0, hi = n - 1;  // the last u with row[u] <= e0 (a row with arcs)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (row[mid] <= e0) lo = mid;
      else hi = mid - 1;
    }
    int u = lo, c = 0;
    int end = row[u + 1];
    for (int64_t e = e0; e < e1; ++e) {
      while (e >= end) {  // the next row with arcs
        if (c) atomicAdd(&inner[u], c);
        c = 0;
        ++u;
        end = row[u + 1];
      }
      c += (hasout[col[e] >> 5] >> (col[e] & 31)) & 1u;
    }
    if (c) atomicAdd(&inner[u], c);
  }
}

// This workgroup's share [b0, b1) of the level's slots [s, e).
__device__ __forceinline__ void level_share(int s, int e, int* b0, int* b1) {
  const int64_t len = e - s;
  *b0 = s + (int)(len * blockIdx.x / gridDim.x);
  *b1 = s + (int)(len * (blockIdx.x + 1) / gridDim.x);
}

__global__ __launch_bounds__(SMDP_WG) void k_tree_down(TArgs a, int L) {
  __shared__ int s_start[SMDP_WG];
  __shared__ int s_off[SMDP_WG];
  __shared__ int s_scan[SMDP_WAVES_PER_WG];
  __shared__ int s_base;
  __shared__ int s_cursor;
  // level L: from the end of level L - 1 for as many slots as level L - 1 reserved
  const int s = a.lo[L];
  const int e = L == 0 ? 1 : s + a.rel[L - 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.lo[L + 1] = e;  // for the launches after this one (none of this launch reads it)
    a.rel[L + 1] = 0;
    if (e == s && a.ctl[1] == 0) {
      a.ctl[1] = L;  // level L is empty: levels 0 .. L-1
      atomicMax(&cnt_of(a.cnt)->rounds, (unsigned long long)L);
    }
  }
  int b0, b1;
  level_share(s, e, &b0, &b1);
  unsigned long long edges = 0, verts = (L == 0 && blockIdx.x == 0 && threadIdx.x == 0) ? 1 : 0;
  if (b1 > b0) {
    // child count of the share, then one reservation for all of it
    int part = 0;
    for (int i = b0 + (int)threadIdx.x; i < b1; i += SMDP_WG) {
      const int u = a.order[i];
      if (u >= 0) part += a.inner[u];
    }
    int total;
    (void)wg_exclusive_scan(part, s_scan, &total);
    if (threadIdx.x == 0) {
      s_base = total ? e + atomicAdd(&a.rel[L], total) : 0;
      s_cursor = 0;
    }
    __syncthreads();
    const int base = s_base;
    for (int t0 = b0; t0 < b1; t0 += SMDP_WG) {
      const int i = t0 + (int)threadIdx.x;
      int u = -1, rs = 0, deg = 0;
      if (i < b1) u = a.order[i];
      if (u >= 0) {
        rs = a.row[u];
        deg = a.row[u + 1] - rs;
      }
      int tot;
      const int ex = wg_exclusive_scan(deg, s_scan, &tot);
      s_start[threadIdx.x] = rs;
      s_off[threadIdx.x] = ex;
      __syncthreads();
      for (int jb = 0; jb < tot; jb += kTreeIlp * SMDP_WG) {
        int j[kTreeIlp], own[kTreeIlp];
#pragma unroll
        for (int t = 0; t < kTreeIlp; ++t) j[t] = min(jb + (int)threadIdx.x + t * SMDP_WG, tot - 1);
        tile_owners<kTreeIlp>(s_off, j, own);
        int c[kTreeIlp];
        uint32_t hw[kTreeIlp];
#pragma unroll
        for (int t = 0; t < kTreeIlp; ++t) c[t] = stream_load(a.col + s_start[own[t]] + (j[t] - s_off[own[t]]));
#pragma unroll
        for (int t = 0; t < kTreeIlp; ++t) hw[t] = a.hasout[c[t] >> 5];
#pragma unroll
        for (int t = 0; t < kTreeIlp; ++t) {
          const int jj = jb + (int)threadIdx.x + t * SMDP_WG;
          const bool ok = jj < tot;
          const bool has = ok && ((hw[t] >> (c[t] & 31)) & 1u);  // the arcs inner[] counts
          const bool child = ok && c[t] != a.root;
          if (child && !has) {
            a.height[c[t]] = 0;  // a leaf
            a.desc[c[t]] = 0;
          }
          edges += ok;
          verts += child;
          const uint64_t m = ballot(has);
          if (m) {  // wave-uniform
            int p = 0;
            if (lane_id() == first_lane(m)) p = atomicAdd(&s_cursor, __popcll(m));
            p = __shfl(p, first_lane(m));
            if (has) a.order[base + p + mbcnt(m)] = child ? c[t] : -1;
          }
        }
      }
      __syncthreads();  // s_start / s_off are written again by the next tile
    }
  }
  edges = wave_sum(edges);
  verts = wave_sum(verts);
  if (lane_id() == 0) {
    if (edges) {
      atomicAdd(&cnt_of(a.cnt)->edges, edges);
      atomicAdd(&cnt_of(a.cnt)->reached, edges);
    }
    if (verts) atomicAdd(&cnt_of(a.cnt)->verts, verts);
  }
}

__global__ __launch_bounds__(SMDP_WG) void k_tree_up(TArgs a, int L) {
  __shared__ int s_start[SMDP_WG];
  __shared__ int s_off[SMDP_WG];
  __shared__ int s_scan[SMDP_WAVES_PER_WG];
  __shared__ int s_h[SMDP_WG];
  __shared__ int s_d[SMDP_WG];
  const int s = a.lo[L], e = a.lo[L + 1];
  int b0, b1;
  level_share(s, e, &b0, &b1);
  unsigned long long edges = 0;
  for (int t0 = b0; t0 < b1; t0 += SMDP_WG) {
    const int i = t0 + (int)threadIdx.x;
    int u = -1, rs = 0, deg = 0;
    if (i < b1) u = a.order[i];
    if (u >= 0) {
      rs = a.row[u];
      deg = a.row[u + 1] - rs;
    }
    int tot;
    const int ex = wg_exclusive_scan(deg, s_scan, &tot);
    s_start[threadIdx.x] = rs;
    s_off[threadIdx.x] = ex;
    s_h[threadIdx.x] = 0;
    s_d[threadIdx.x] = 0;
    __syncthreads();
    for (int jb = 0; jb < tot; jb += kTreeIlp * SMDP_WG) {
      int j[kTreeIlp], own[kTreeIlp], c[kTreeIlp];
#pragma unroll
      for (int t = 0; t < kTreeIlp; ++t) j[t] = min(jb + (int)threadIdx.x + t * SMDP_WG, tot - 1);
      tile_owners<kTreeIlp>(s_off, j, own);
#pragma unroll
      for (int t = 0; t < kTreeIlp; ++t) c[t] = stream_load(a.col + s_start[own[t]] + (j[t] - s_off[own[t]]));
      int ch[kTreeIlp], cd[kTreeIlp];
#pragma unroll
      for (int t = 0; t < kTreeIlp; ++t) {
        // clamped gathers (the root's entries are valid); skipped below
        ch[t] = a.height[c[t]];
        cd[t] = a.desc[c[t]];
      }
#pragma unroll
      for (int t = 0; t < kTreeIlp; ++t) {
        const int jj = jb + (int)threadIdx.x + t * SMDP_WG;
        if (jj < tot && c[t] != a.root) {
          atomicMax(&s_h[own[t]], 1 + ch[t]);
          atomicAdd(&s_d[own[t]], 1 + cd[t]);
          ++edges;
        }
      }
    }
    __syncthreads();
    if (u >= 0) {
      a.height[u] = s_h[threadIdx.x];
      a.desc[u] = s_d[threadIdx.x];
    }
    __syncthreads();  // the LDS arrays are written again by the next tile
  }
  edges = wave_sum(edges);
  if (lane_id() == 0 && edges) atomicAdd(&cnt_of(a.cnt)->edges, edges);
}

}  // namespace

// Runs both passes from root; heights in g->dist, descendant counts in
// g->stamp.
static int run_tree(smdp_graph* g, int32_t root, smdp_stats* st) {
  if (!g || g->n <= 0 || root < 0 || root >= g->n) return SMDP_ERR_ARG;
  smdp_ctx* ctx = g->ctx;
  SMDP_HIP(hipSetDevice(ctx->device));
  begin_run(g);
  SMDP_CHECK(ensure_forest(g));
  if (g->forest != 1) return SMDP_ERR_ARG;  // a vertex with two in-arcs: not a tree
  const int64_t n = g->n;
  SMDP_CHECK(alloc_once(&g->dist, n));
  SMDP_CHECK(alloc_once(&g->stamp, n));
  SMDP_CHECK(build_hasout(g, g->row, &g->hasout));
  if (!g->tree_inner) {
    SMDP_CHECK(alloc_once(&g->tree_inner, n));
    SMDP_HIP(hipMemsetAsync(g->tree_inner, 0, sizeof(int32_t) * n, ctx->stream));
    const int64_t nseg = ((int64_t)g->m + kInnerSeg - 1) / kInnerSeg;
    if (nseg > 0)
      hipLaunchKernelGGL(k_tree_inner, dim3((unsigned)std::min<int64_t>((nseg + 255) / 256, 8192)), dim3(256), 0,
                         ctx->stream, g->row, g->col, g->hasout, g->n, g->tree_inner);
    SMDP_HIP(hipGetLastError());
  }
  // one buffer: level starts and per-level reservations (indices up to
  // n + 2), the control line, and order[] (every reached vertex at most once
  // plus at most one hole: n + 1 slots)
  const int64_t lo_len = (n + 3 + kTreeCtl - 1) / kTreeCtl * kTreeCtl;
  SMDP_CHECK(alloc_once(&g->tree_lo, 2 * lo_len + kTreeCtl + n + 1));
  TArgs a;
  a.row = g->row;
  a.col = g->col;
  a.hasout = g->hasout;
  a.inner = g->tree_inner;
  a.lo = g->tree_lo;
  a.rel = g->tree_lo + lo_len;
  a.ctl = g->tree_lo + 2 * lo_len;
  a.order = g->tree_lo + 2 * lo_len + kTreeCtl;
  a.height = g->dist;
  a.desc = g->stamp;
  a.cnt = g->cnt;
  a.root = root;
  a.n = g->n;
  const int grid = std::max(8, ctx->cus * 4);
  hipStream_t s = ctx->stream;
  prof_reset(ctx);
  int launches = 0;
  SMDP_HIP(hipEventRecord(ctx->ev_begin, s));
  hipLaunchKernelGGL(k_tree_init, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, s, a);
  ++launches;
  // Down, in batches of levels, until a launch finds its level empty:
  // order[] holds at most n + 1 slots, so level n + 1 is empty at the latest
  // (lo and rel are indexed up to n + 2).
  // (an empty level's launch costs ~4 us; a host check between batches about
  // 100 us, so the first batch covers the usual depths)
  int L = 0, levels = 0;
  int batch = 8;
  for (;;) {
    for (int j = 0; j < batch && L <= n + 1; ++j, ++L) {
      SMDP_CHECK(prof_mark(ctx, 0));
      hipLaunchKernelGGL(k_tree_down, dim3(grid), dim3(SMDP_WG), 0, s, a, L);
      SMDP_CHECK(prof_mark(ctx, 0));
      ++launches;
    }
    SMDP_HIP(hipGetLastError());
    SMDP_HIP(hipMemcpyAsync(ctx->pinned, a.ctl + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SMDP_HIP(hipStreamSynchronize(s));
    levels = ctx->pinned[0];
    if (levels > 0) break;
    if (L > n + 1) return SMDP_ERR_OVERFLOW;  // cannot happen for a forest
    batch = std::min(batch * 2, 64);
  }
  for (int l = levels - 1; l >= 0; --l) {
    SMDP_CHECK(prof_mark(ctx, 1));
    hipLaunchKernelGGL(k_tree_up, dim3(grid), dim3(SMDP_WG), 0, s, a, l);
    SMDP_CHECK(prof_mark(ctx, 1));
    ++launches;
  }
  SMDP_HIP(hipGetLastError());
  SMDP_HIP(hipEventRecord(ctx->ev_end, s));
  SMDP_HIP(hipEventSynchronize(ctx->ev_end));
  float ms = 0;
  SMDP_HIP(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
  g->last_res = SMDP_RES_HEIGHT;
  g->dist_perm = false;
  if (st) {
    SMDP_CHECK(fill_stats(g, st, ms, launches));
    SMDP_CHECK(prof_collect(ctx, st));
  } else {
    prof_reset(ctx);
  }
  return SMDP_OK;
}

}  // namespace smdp

using namespace smdp;

extern "C" {

int smdp_tree_run(smdp_graph* g, int32_t root, smdp_stats* st) { return run_tree(g, root, st); }

int smdp_tree(smdp_ctx* ctx, const smdp_csr* h, int32_t root, int32_t* height, int32_t* desc, int32_t* tree_height,
              smdp_stats* st) {
  if (!ctx || !h) return SMDP_ERR_ARG;
  double t0 = now_ms();
  smdp_graph* g = nullptr;
  SMDP_CHECK(smdp_graph_create(ctx, h, &g));
  g->transient = true;
  int rc = run_tree(g, root, st);
  if (rc == SMDP_OK && height) rc = smdp_graph_read(g, SMDP_RES_HEIGHT, height, g->n);
  if (rc == SMDP_OK && desc) rc = smdp_graph_read(g, SMDP_RES_DESC, desc, g->n);
  if (rc == SMDP_OK && tree_height) {
    int32_t hr = 0;
    if (hipMemcpy(&hr, g->dist + root, sizeof hr, hipMemcpyDeviceToHost) != hipSuccess) rc = SMDP_ERR_DEVICE;
    *tree_height = hr;
  }
  smdp_graph_destroy(g);
  if (st) st->ms_total = now_ms() - t0;
  return rc;
}

}  // extern "C"
