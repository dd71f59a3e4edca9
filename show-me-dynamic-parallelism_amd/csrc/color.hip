// Graph colouring on gfx950.
//
// Reference rule (gclr_bitmap_kernel, apps/graph-color/src/graph_color_kernel.cu:28-49,
// rounds driven by graph_color_wrapper.cu:113-208): in round c = 1, 2, ...
// an uncoloured v takes colour c unless some u in row(v) with u < v has
// colour 0 or c.  Equivalently colour(v) = 1 + max{colour(u) : u in row(v),
// u precedes v}, which does not depend on thread timing.  With priorities,
// "u precedes v" is (prio[u], u) < (prio[v], v); without, it is u < v as in
// the reference.
//
// Here a round only looks at vertices still uncoloured (a compacted queue,
// not all N as the reference does), and each vertex resumes its scan at the
// arc that blocked it last round with the running maximum of the colours it
// already passed (resume[v], rmax[v]), so every arc is read about once plus
// one re-check per blocked round.  When the scan of v reaches the end of its
// row in round c, all its predecessors are coloured below c and the one that
// blocked it last round got colour c-1, so colour(v) = 1 + rmax = c.
// Heavy rows are scanned by a wave, a workgroup (LDS queues) or the grid
// (a second kernel), as the solution says.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "smdp_internal.h"

namespace smdp {

constexpr int kCOut = 512;     // per-wave LDS staging of blocked vertices
constexpr int kCWaveQ = 128;   // per-wave heavy queue (SOL_WAVE)
constexpr int kCBlockQ = SMDP_WG;

struct CArgs {
  const int32_t* __restrict__ row;
  const int32_t* __restrict__ col;
  const uint32_t* __restrict__ prio;  // nullptr: vertex-id priority
  int32_t* color;
  int32_t* resume;
  int32_t* rmax;
  int32_t* q0;
  int32_t* q1;
  int32_t* hq;
  Ctl* ctl;
  Counters* cnt;
  int32_t n;
  int32_t heavy;   // lane solutions: heavy degree; tuned: wave degree
  int32_t split;   // tuned: grid degree
};

// u precedes v.
__device__ __forceinline__ bool precedes(const CArgs& a, int u, uint32_t pv, int v) {
  if (!a.prio) return u < v;
  const uint32_t pu = a.prio[u];
  return pu < pv || (pu == pv && u < v);
}

struct ScanState {
  int pos;
  int mx;
};

// Serial scan by one lane from s.pos to e; returns true when it reached e.
__device__ __forceinline__ bool lane_scan(const CArgs& a, int v, uint32_t pv, int e, int c, ScanState& s,
                                          unsigned long long& edges) {
  int q = s.pos;
  for (; q < e; ++q) {
    const int u = a.col[q];
    if (!precedes(a, u, pv, v)) continue;
    const int cu = a.color[u];
    if (cu == 0 || cu == c) break;
    s.mx = max(s.mx, cu);
  }
  edges += (unsigned long long)(q - s.pos) + (q < e ? 1 : 0);
  s.pos = q;
  return q >= e;
}

// Wave-cooperative scan of one vertex (wave-uniform arguments).
__device__ __forceinline__ bool wave_scan(const CArgs& a, int v, uint32_t pv, int e, int c, ScanState& s,
                                          unsigned long long& edges) {
  const int lane = lane_id();
  for (int b = s.pos; b < e; b += SMDP_WAVE) {
    const int q = b + lane;
    bool isp = false, blk = false;
    int cu = 0;
    if (q < e) {
      const int u = a.col[q];
      isp = precedes(a, u, pv, v);
      if (isp) {
        cu = a.color[u];
        blk = (cu == 0 || cu == c);
      }
    }
    const uint64_t m = ballot(blk);
    if (m) {
      const int fb = first_lane(m);
      s.mx = max(s.mx, wave_max((isp && lane < fb) ? cu : 0));
      s.pos = b + fb;
      if (lane == 0) edges += fb + 1;
      return false;
    }
    s.mx = max(s.mx, wave_max(isp ? cu : 0));
    if (lane == 0) edges += min(SMDP_WAVE, e - b);
  }
  s.pos = e;
  return true;
}

// Workgroup-cooperative scan (all threads call it with uniform arguments).
__device__ __forceinline__ bool wg_scan(const CArgs& a, int v, uint32_t pv, int e, int c, ScanState& s,
                                        unsigned long long& edges, int* s_first, int* s_max) {
  for (int b = s.pos; b < e; b += SMDP_WG) {
    const int q = b + (int)threadIdx.x;
    bool isp = false, blk = false;
    int cu = 0;
    if (q < e) {
      const int u = a.col[q];
      isp = precedes(a, u, pv, v);
      if (isp) {
        cu = a.color[u];
        blk = (cu == 0 || cu == c);
      }
    }
    if (threadIdx.x == 0) {
      *s_first = SMDP_WG;
      *s_max = 0;
    }
    __syncthreads();
    if (blk) atomicMin(s_first, (int)threadIdx.x);
    __syncthreads();
    const int fb = *s_first;
    const int my = (isp && (int)threadIdx.x < fb) ? cu : 0;
    const int wm = wave_max(my);
    if (lane_id() == 0 && wm) atomicMax(s_max, wm);
    __syncthreads();
    s.mx = max(s.mx, *s_max);
    __syncthreads();
    if (fb < SMDP_WG) {
      s.pos = b + fb;
      if (threadIdx.x == 0) edges += fb + 1;
      return false;
    }
    if (threadIdx.x == 0) edges += min(SMDP_WG, e - b);
  }
  s.pos = e;
  return true;
}

__global__ void k_color_init(int32_t* color, int32_t n, Ctl* ctl, Counters* cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    color[i] = 0;
  if (blockIdx.x == 0) {
    int32_t* c = reinterpret_cast<int32_t*>(ctl);
    for (int i = threadIdx.x; i < 4 * 32; i += blockDim.x) c[i] = 0;
    unsigned long long* k = reinterpret_cast<unsigned long long*>(cnt);
    for (int i = threadIdx.x; i < (int)(sizeof(Counters) / 8); i += blockDim.x) k[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) ctl[0].qlen = n;  // round 0 takes every vertex, in id order
  }
}

// Finish one vertex after its scan: colour it or record where it stopped.
__device__ __forceinline__ bool settle(const CArgs& a, int v, int c, const ScanState& s, bool done) {
  if (done) {
    a.color[v] = c;  // == 1 + s.mx
    return false;
  }
  a.resume[v] = s.pos;
  a.rmax[v] = s.mx;
  return true;
}

template <int SOL>
__global__ __launch_bounds__(SMDP_WG) void k_color(CArgs a, int r) {
  if (blockIdx.x == 0 && threadIdx.x < 32) reinterpret_cast<int32_t*>(a.ctl + ((r + 2) & 3))[threadIdx.x] = 0;
  Ctl* cin = a.ctl + (r & 3);
  Ctl* cout = a.ctl + ((r + 1) & 3);
  const int L = cin->qlen;
  if (L == 0) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&a.cnt->rounds, (unsigned long long)(r + 1));
  const int c = r + 1;
  const int32_t* qin = (r & 1) ? a.q1 : a.q0;
  int32_t* qout = (r & 1) ? a.q0 : a.q1;
  const int lane = lane_id();
  const int wave = threadIdx.x / SMDP_WAVE;
  __shared__ int32_t s_out[SMDP_WAVES_PER_WG * kCOut];
  __shared__ int32_t s_q[SOL == SMDP_SOL_BLOCK ? kCBlockQ : (SOL == SMDP_SOL_WAVE ? SMDP_WAVES_PER_WG * kCWaveQ : 1)];
  __shared__ int32_t s_qn;
  __shared__ int s_first, s_max;
  WaveOut<kCOut> out;
  out.init(s_out + wave * kCOut);
  unsigned long long edges = 0, verts = 0, pushes = 0, pw = 0, pb = 0, pg = 0;
  const int heavy_deg = a.heavy;
  const int split_deg = SOL == SMDP_SOL_TUNED ? a.split : INT32_MAX;

  const int stride = gridDim.x * SMDP_WG;
  for (int base = blockIdx.x * SMDP_WG; base < L; base += stride) {
    const int i = base + (int)threadIdx.x;
    const bool valid = i < L;
    int v = 0, e = 0;
    uint32_t pv = 0;
    ScanState s{0, 0};
    if (valid) {
      v = r == 0 ? i : qin[i];
      e = a.row[v + 1];
      s.pos = r == 0 ? a.row[v] : a.resume[v];
      s.mx = r == 0 ? 0 : a.rmax[v];
      if (a.prio) pv = a.prio[v];
      ++verts;
    }
    const int rem = e - s.pos;
    const bool grid = SOL == SMDP_SOL_GRID ? (valid && rem >= heavy_deg) : (valid && rem >= split_deg);
    const bool heavy = SOL != SMDP_SOL_FLAT && valid && !grid && rem >= heavy_deg;
    bool blocked = false;
    bool settled = false;

    if (SOL == SMDP_SOL_NAIVE || SOL == SMDP_SOL_TUNED) {
      uint64_t hm = ballot(heavy);
      pw += lane == 0 ? __popcll(hm) : 0;
      while (hm) {
        const int l = first_lane(hm);
        hm &= hm - 1;
        ScanState hs{__shfl(s.pos, l), __shfl(s.mx, l)};
        const int hv = __shfl(v, l);
        const bool done = wave_scan(a, hv, (uint32_t)__shfl((int)pv, l), __shfl(e, l), c, hs, edges);
        if (lane == l) {
          s = hs;
          blocked = settle(a, v, c, s, done);
          settled = true;
        }
      }
    }
    if (valid && !heavy && !grid) {
      const bool done = lane_scan(a, v, pv, e, c, s, edges);
      blocked = settle(a, v, c, s, done);
      settled = true;
    }
    if (SOL == SMDP_SOL_WAVE) {
      const uint64_t hm = ballot(heavy);
      if (heavy) s_q[wave * kCWaveQ + mbcnt(hm)] = (int)threadIdx.x;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const int nq = __popcll(hm);
      pw += lane == 0 ? nq : 0;
      for (int j = 0; j < nq; ++j) {
        const int l = s_q[wave * kCWaveQ + j] % SMDP_WAVE;
        ScanState hs{__shfl(s.pos, l), __shfl(s.mx, l)};
        const bool done = wave_scan(a, __shfl(v, l), (uint32_t)__shfl((int)pv, l), __shfl(e, l), c, hs, edges);
        if (lane == l) {
          s = hs;
          blocked = settle(a, v, c, s, done);
          settled = true;
        }
      }
    }
    if (SOL == SMDP_SOL_BLOCK) {
      if (threadIdx.x == 0) s_qn = 0;
      __syncthreads();
      if (heavy) s_q[atomicAdd(&s_qn, 1)] = (int)threadIdx.x;
      __syncthreads();
      const int nq = s_qn;
      pb += threadIdx.x == 0 ? nq : 0;
      for (int j = 0; j < nq; ++j) {
        const int owner = s_q[j];
        const int hv = base + owner < L ? (r == 0 ? base + owner : qin[base + owner]) : 0;
        ScanState hs{r == 0 ? a.row[hv] : a.resume[hv], r == 0 ? 0 : a.rmax[hv]};
        const bool done = wg_scan(a, hv, a.prio ? a.prio[hv] : 0u, a.row[hv + 1], c, hs, edges, &s_first, &s_max);
        if ((int)threadIdx.x == owner) {
          s = hs;
          blocked = settle(a, v, c, s, done);
          settled = true;
        }
      }
    }
    if (SOL == SMDP_SOL_GRID || SOL == SMDP_SOL_TUNED) {
      const uint64_t gm = ballot(grid);
      if (gm) {
        const int leader = first_lane(gm);
        int b0 = 0;
        if (lane == leader) b0 = atomicAdd(&cin->hlen, __popcll(gm));
        b0 = __shfl(b0, leader);
        if (grid) a.hq[b0 + mbcnt(gm)] = v;
        pg += lane == leader ? __popcll(gm) : 0;
      }
    }
    (void)settled;
    pushes += out.push(blocked, v, qout, &cout->qlen);
  }
  out.flush(qout, &cout->qlen);
  edges = wave_sum(edges);
  verts = wave_sum(verts);
  pushes = wave_sum(pushes);
  pw = wave_sum(pw);
  pb = wave_sum(pb);
  pg = wave_sum(pg);
  if (lane == 0) {
    if (edges) atomicAdd(&a.cnt->edges, edges);
    if (verts) atomicAdd(&a.cnt->verts, verts);
    if (pushes) atomicAdd(&a.cnt->push[0], pushes);
    if (pw) atomicAdd(&a.cnt->push[1], pw);
    if (pb) atomicAdd(&a.cnt->push[2], pb);
    if (pg) atomicAdd(&a.cnt->push[3], pg);
  }
}

// Phase 2 of the grid solutions: each queued vertex is scanned by a whole
// workgroup.
__global__ __launch_bounds__(SMDP_WG) void k_color_grid(CArgs a, int r) {
  Ctl* cin = a.ctl + (r & 3);
  Ctl* cout = a.ctl + ((r + 1) & 3);
  const int H = cin->hlen;
  if (H == 0) return;
  const int c = r + 1;
  int32_t* qout = (r & 1) ? a.q0 : a.q1;
  __shared__ int s_idx, s_first, s_max;
  unsigned long long edges = 0, pushes = 0;
  for (;;) {
    if (threadIdx.x == 0) s_idx = atomicAdd(&cin->chunk_next, 1);
    __syncthreads();
    const int idx = s_idx;
    __syncthreads();
    if (idx >= H) break;
    const int v = a.hq[idx];
    ScanState s{r == 0 ? a.row[v] : a.resume[v], r == 0 ? 0 : a.rmax[v]};
    const bool done = wg_scan(a, v, a.prio ? a.prio[v] : 0u, a.row[v + 1], c, s, edges, &s_first, &s_max);
    if (threadIdx.x == 0) {
      if (settle(a, v, c, s, done)) {
        qout[atomicAdd(&cout->qlen, 1)] = v;
        ++pushes;
      }
    }
  }
  edges = wave_sum(edges);
  pushes = wave_sum(pushes);
  if (lane_id() == 0) {
    if (edges) {
      atomicAdd(&a.cnt->edges, edges);
      atomicAdd(&a.cnt->edges2, edges);
    }
    if (pushes) atomicAdd(&a.cnt->push[0], pushes);
  }
}

}  // namespace smdp

namespace {
struct ColorBufs {
  int32_t* resume = nullptr;
  int32_t* rmax = nullptr;
};
}  // namespace

namespace smdp {

void colorbufs_free(smdp_graph* g) {
  ColorBufs* b = static_cast<ColorBufs*>(g->colorbufs);
  if (!b) return;
  if (b->resume) (void)hipFree(b->resume);
  if (b->rmax) (void)hipFree(b->rmax);
  delete b;
  g->colorbufs = nullptr;
}

int run_color(smdp_graph* g, int use_prio, int sol, smdp_stats* st) {
  if (!g) return SMDP_ERR_ARG;
  if (sol < 0 || sol >= SMDP_SOL_COUNT || sol == SMDP_SOL_HIER) return SMDP_ERR_SOLUTION;
  if (use_prio && !g->prio) return SMDP_ERR_ARG;
  smdp_ctx* ctx = g->ctx;
  SMDP_HIP(hipSetDevice(ctx->device));
  if (!g->colorbufs) g->colorbufs = new ColorBufs();
  ColorBufs* b = static_cast<ColorBufs*>(g->colorbufs);
  const int64_t n1 = std::max<int64_t>(1, g->n);
  SMDP_CHECK(alloc_once(&g->dist, n1));
  SMDP_CHECK(alloc_once(&g->q[0], n1));
  SMDP_CHECK(alloc_once(&g->q[1], n1));
  SMDP_CHECK(alloc_once(&b->resume, n1));
  SMDP_CHECK(alloc_once(&b->rmax, n1));
  SMDP_CHECK(ensure(&g->hq, &g->hq_cap, n1 + 64));
  const smdp_params& p = params_of(ctx, SMDP_APP_COLOR);
  CArgs a;
  a.row = g->row;
  a.col = g->col;
  a.prio = use_prio ? g->prio : nullptr;
  a.color = g->dist;
  a.resume = b->resume;
  a.rmax = b->rmax;
  a.q0 = g->q[0];
  a.q1 = g->q[1];
  a.hq = reinterpret_cast<int32_t*>(g->hq);
  a.ctl = g->ctl;
  a.cnt = g->cnt;
  a.n = g->n;
  a.heavy = p.heavy_threshold;
  a.split = std::max(p.split_threshold, p.heavy_threshold);
  const int grid = grid_size(ctx, p);
  const bool two_phase = sol == SMDP_SOL_GRID || sol == SMDP_SOL_TUNED;
  hipStream_t s = ctx->stream;
  prof_reset(ctx);
  int launches = 0;
  SMDP_HIP(hipEventRecord(ctx->ev_begin, s));
  hipLaunchKernelGGL(k_color_init, dim3(std::min<int64_t>((n1 + 255) / 256, 4096)), dim3(256), 0, s, g->dist, g->n,
                     g->ctl, g->cnt);
  ++launches;
  int r = 0;
  int batch = std::max(1, p.batch_rounds);
  if (g->n > 0) {
    for (;;) {
      for (int j = 0; j < batch; ++j, ++r) {
        dim3 gd(grid), bd(SMDP_WG);
        prof_mark(ctx, 0);
        switch (sol) {
          case SMDP_SOL_FLAT: hipLaunchKernelGGL(k_color<SMDP_SOL_FLAT>, gd, bd, 0, s, a, r); break;
          case SMDP_SOL_NAIVE: hipLaunchKernelGGL(k_color<SMDP_SOL_NAIVE>, gd, bd, 0, s, a, r); break;
          case SMDP_SOL_WAVE: hipLaunchKernelGGL(k_color<SMDP_SOL_WAVE>, gd, bd, 0, s, a, r); break;
          case SMDP_SOL_BLOCK: hipLaunchKernelGGL(k_color<SMDP_SOL_BLOCK>, gd, bd, 0, s, a, r); break;
          case SMDP_SOL_GRID: hipLaunchKernelGGL(k_color<SMDP_SOL_GRID>, gd, bd, 0, s, a, r); break;
          default: hipLaunchKernelGGL(k_color<SMDP_SOL_TUNED>, gd, bd, 0, s, a, r); break;
        }
        prof_mark(ctx, 0);
        ++launches;
        if (two_phase) {
          prof_mark(ctx, 1);
          hipLaunchKernelGGL(k_color_grid, gd, bd, 0, s, a, r);
          prof_mark(ctx, 1);
          ++launches;
        }
      }
      SMDP_HIP(hipGetLastError());
      SMDP_HIP(hipMemcpyAsync(ctx->pinned, &g->ctl[r & 3].qlen, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      SMDP_HIP(hipStreamSynchronize(s));
      if (ctx->pinned[0] == 0) break;
      if (r >= p.max_rounds || r >= g->n + 1) return SMDP_ERR_OVERFLOW;
      batch = std::min(batch * 2, 32);
    }
  }
  SMDP_HIP(hipEventRecord(ctx->ev_end, s));
  SMDP_HIP(hipEventSynchronize(ctx->ev_end));
  float ms = 0;
  SMDP_HIP(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
  g->last_res = SMDP_RES_COLOR;
  if (st) {
    SMDP_CHECK(fill_stats(g, st, ms, launches));
    st->m_reached = g->m;
    SMDP_CHECK(prof_collect(ctx, st));
  } else {
    prof_reset(ctx);
  }
  return SMDP_OK;
}

}  // namespace smdp

using namespace smdp;

extern "C" {

int smdp_color_run(smdp_graph* g, int use_prio, int solution, smdp_stats* st) {
  return run_color(g, use_prio, solution, st);
}

int smdp_color(smdp_ctx* ctx, const smdp_csr* h, const uint32_t* prio, int solution, int32_t* color,
               smdp_stats* st) {
  if (!ctx || !h || !color) return SMDP_ERR_ARG;
  double t0 = now_ms();
  smdp_graph* g = nullptr;
  smdp_csr c = *h;
  c.w = nullptr;
  SMDP_CHECK(smdp_graph_create(ctx, &c, &g));
  int rc = prio ? smdp_graph_set_prio(g, prio) : SMDP_OK;
  if (rc == SMDP_OK) rc = run_color(g, prio != nullptr, solution, st);
  if (rc == SMDP_OK) rc = smdp_graph_read(g, SMDP_RES_COLOR, color, g->n);
  smdp_graph_destroy(g);
  if (st) st->ms_total = now_ms() - t0;
  return rc;
}

}  // extern "C"
