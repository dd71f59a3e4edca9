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
//
// The tuned solution computes the same colours by topological layering
// (below): colour(v) is the length of the longest chain of preceding
// neighbours ending at v, so Kahn's algorithm over the "precedes" order gives
// it with every arc read twice per run, where the round rule reads the
// blocking arc of every blocked vertex once per round.  On a symmetrised
// RMAT-23 graph with random priorities that is 2373 rounds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device_common.h"
#include "grid_barrier.h"
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
  int32_t heavy;   // heavy degree (wave, workgroup or grid consolidation)
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
    for (int i = threadIdx.x; i < 4 * kCtlInts; i += blockDim.x) c[i] = 0;
    unsigned long long* k = reinterpret_cast<unsigned long long*>(cnt);
    for (int i = threadIdx.x; i < (int)(kCntParts * sizeof(Counters) / 8); i += blockDim.x) k[i] = 0;
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
  ctl_clear(a.ctl, (r + 2) & 3);
  Ctl* cin = a.ctl + (r & 3);
  Ctl* cout = a.ctl + ((r + 1) & 3);
  const int L = cin->qlen;
  if (L == 0) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&cnt_of(a.cnt)->rounds, (unsigned long long)(r + 1));
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
    const bool grid = SOL == SMDP_SOL_GRID && valid && rem >= heavy_deg;
    const bool heavy = SOL != SMDP_SOL_FLAT && valid && !grid && rem >= heavy_deg;
    bool blocked = false;
    bool settled = false;

    if (SOL == SMDP_SOL_NAIVE) {
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
    if (SOL == SMDP_SOL_GRID) {
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
    const int queued = out.push(blocked, v, qout, &cout->qlen);  // the same on every lane
    if (lane == 0) pushes += queued;
  }
  out.flush(qout, &cout->qlen);
  edges = wave_sum(edges);
  verts = wave_sum(verts);
  pushes = wave_sum(pushes);
  pw = wave_sum(pw);
  pb = wave_sum(pb);
  pg = wave_sum(pg);
  if (lane == 0) {
    if (edges) atomicAdd(&cnt_of(a.cnt)->edges, edges);
    if (verts) atomicAdd(&cnt_of(a.cnt)->verts, verts);
    if (pushes) atomicAdd(&cnt_of(a.cnt)->push[0], pushes);
    if (pw) atomicAdd(&cnt_of(a.cnt)->push[1], pw);
    if (pb) atomicAdd(&cnt_of(a.cnt)->push[2], pb);
    if (pg) atomicAdd(&cnt_of(a.cnt)->push[3], pg);
  }
}

// Phase 2 of the grid solutions: each queued vertex is scanned by a whole
// workgroup.
__global__ __launch_bounds__(SMDP_WG) void k_color_grid(CArgs a, int r) {
  Ctl* cin = a.ctl + (r & 3);
  Ctl* cout = a.ctl + ((r + 1) & 3);
  const int H = cin->hlen;
  if ((int)blockIdx.x >= H) return;
  const int c = r + 1;
  int32_t* qout = (r & 1) ? a.q0 : a.q1;
  __shared__ int s_idx, s_first, s_max;
  unsigned long long edges = 0, pushes = 0;
  int part = 0;  // thread 0: claim cursor
  for (;;) {
    if (threadIdx.x == 0) {
      int f = 0, e = 0;
      s_idx = claim_items(cin->chunks, H, 1, part, &f, &e) ? f : -1;
    }
    __syncthreads();
    const int idx = s_idx;
    __syncthreads();
    if (idx < 0) break;
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
      atomicAdd(&cnt_of(a.cnt)->edges, edges);
      atomicAdd(&cnt_of(a.cnt)->edges2, edges);
    }
    if (pushes) atomicAdd(&cnt_of(a.cnt)->push[0], pushes);
  }
}

// ---- tuned solution: topological layering --------------------------------
// pend[v] counts the arcs of row(v) that point to a preceding vertex.  The
// vertices with pend 0 get colour 1.  In round r every vertex u of colour r+1
// decrements pend[v] once for each arc of row(v) that points to u, found
// through the transposed CSR (trow/tcol), when u precedes v; the v whose
// count reaches zero gets colour r+2: all its preceding neighbours are
// coloured and the last one, u, has the largest colour among them.  A vertex
// enters a round's queue once, when its count reaches zero.

struct LArgs {
  const int32_t* __restrict__ row;
  const int32_t* __restrict__ col;
  const int32_t* __restrict__ trow;
  const int32_t* __restrict__ tcol;
  const uint32_t* __restrict__ prio;  // nullptr: vertex-id priority
  // with priorities: prio of each arc's vertex, aligned with col and tcol
  // (built once per graph and priority upload), so an arc's priority comes
  // with the arc instead of a dependent gather
  const uint32_t* __restrict__ cprio;
  const uint32_t* __restrict__ tprio;
  int32_t* color;
  int32_t* pend;
  int32_t* q0;
  int32_t* q1;
  Chunk* hq;
  Ctl* ctl;
  Counters* cnt;
  int32_t n;
  int32_t heavy;  // rows of at least this many arcs are counted by a wave
  int32_t split;  // rows of at least this many arcs go to grid chunks
  int32_t chunk;  // arcs per grid chunk
  int32_t grid;   // workgroups of k_layer_expand (the round's tile shape depends on it)
  // Layer indices kept on the device, so that a launch can run several layers:
  // rd[0] the layer k_layer_expand / k_layer_drain run next, rd[1] the layer
  // the last drain ran, rd[2] set once a layer is empty (done), rd[3] set
  // when a barrier between short layers gave up waiting.
  int32_t* rd;
  uint32_t* bar;  // flags of k_layer_short's barrier (two parities of 64-bit flags), zeroed before each launch
};

constexpr int kLayerIlp = 4;
constexpr int kLayerCountSlot = 3;  // control slot of the counting chunks (cleared when layer r = 1 starts)

__global__ void k_layer_init(Ctl* ctl, Counters* cnt, int32_t* rd) {
  if (threadIdx.x < 4) rd[threadIdx.x] = 0;
  int32_t* c = reinterpret_cast<int32_t*>(ctl);
  for (int i = threadIdx.x; i < 4 * kCtlInts; i += blockDim.x) c[i] = 0;
  unsigned long long* k = reinterpret_cast<unsigned long long*>(cnt);
  for (int i = threadIdx.x; i < (int)(kCntParts * sizeof(Counters) / 8); i += blockDim.x) k[i] = 0;
}

// Arcs of row(v) [s, e) pointing to a vertex that precedes v.
__device__ __forceinline__ int count_preceding(const LArgs& a, int v, uint32_t pv, int s, int e, int step, int first) {
  int c = 0;
  for (int q = s + first; q < e; q += step) {
    const int u = a.col[q];
    bool p;
    if (!a.prio) {
      p = u < v;
    } else {
      const uint32_t pu = a.cprio[q];
      p = pu < pv || (pu == pv && u < v);
    }
    c += p;
  }
  return c;
}

// pend[v] for every vertex; rows of >= split arcs are cut into chunks that
// k_layer_count_chunks adds.
__global__ __launch_bounds__(SMDP_WG) void k_layer_count(LArgs a) {
  const int lane = lane_id();
  Ctl* cs = a.ctl + kLayerCountSlot;
  unsigned long long edges = 0;
  for (int base = blockIdx.x * SMDP_WG; base < a.n; base += gridDim.x * SMDP_WG) {
    const int v = base + (int)threadIdx.x;
    const bool valid = v < a.n;
    int s = 0, e = 0;
    uint32_t pv = 0;
    if (valid) {
      s = a.row[v];
      e = a.row[v + 1];
      if (a.prio) pv = a.prio[v];
    }
    const int deg = e - s;
    const bool big = valid && deg >= a.split;
    const bool heavy = valid && !big && deg >= a.heavy;
    int c = 0;
    if (valid && !big && !heavy) c = count_preceding(a, v, pv, s, e, 1, 0);
    uint64_t hm = ballot(heavy);
    while (hm) {
      const int l = first_lane(hm);
      hm &= hm - 1;
      const int hv = __shfl(v, l);
      const int hc = wave_sum(count_preceding(a, hv, (uint32_t)__shfl((int)pv, l), __shfl(s, l), __shfl(e, l),
                                              SMDP_WAVE, lane));
      if (lane == l) c = hc;
    }
    const int nch = big ? (deg + a.chunk - 1) / a.chunk : 0;
    const int incl = wave_inclusive_scan(nch);
    const int wtot = __shfl(incl, SMDP_WAVE - 1);
    if (wtot > 0) {
      int b0 = 0;
      if (lane == 0) b0 = atomicAdd(&cs->hlen, wtot);
      b0 = __shfl(b0, 0);
      for (int k = 0; k < nch; ++k) a.hq[b0 + incl - nch + k] = Chunk{v, s + k * a.chunk};
    }
    if (valid) {
      a.pend[v] = c;  // chunks of a big row add theirs in the next kernel
      if (!big) edges += deg;
    }
  }
  edges = wave_sum(edges);
  if (lane == 0 && edges) atomicAdd(&cnt_of(a.cnt)->edges, edges);
}

__global__ __launch_bounds__(SMDP_WG) void k_layer_count_chunks(LArgs a) {
  Ctl* cs = a.ctl + kLayerCountSlot;
  const int H = cs->hlen;
  if ((int)blockIdx.x >= H) return;
  __shared__ int s_idx;
  unsigned long long edges = 0;
  int part = 0;  // thread 0: claim cursor
  for (;;) {
    if (threadIdx.x == 0) {
      int f = 0, e = 0;
      s_idx = claim_items(cs->chunks, H, 1, part, &f, &e) ? f : -1;
    }
    __syncthreads();
    const int idx = s_idx;
    __syncthreads();
    if (idx < 0) break;
    const Chunk ch = a.hq[idx];
    const int e = min(a.row[ch.v + 1], ch.first + a.chunk);
    const int c = wave_sum(count_preceding(a, ch.v, a.prio ? a.prio[ch.v] : 0u, ch.first, e, SMDP_WG,
                                           (int)threadIdx.x));
    if (lane_id() == 0 && c) atomicAdd(&a.pend[ch.v], c);
    if (threadIdx.x == 0) edges += e - ch.first;
  }
  edges = wave_sum(edges);
  if (lane_id() == 0 && edges) atomicAdd(&cnt_of(a.cnt)->edges, edges);
}

// Colour 1 for the vertices without preceding neighbours: the queue of round 0.
// A vertex that no row lists (no transposed arcs) has nobody to notify and is
// coloured without being queued; RMAT-23's first layer is mostly such.
__global__ __launch_bounds__(SMDP_WG) void k_layer_seed(LArgs a) {
  __shared__ int32_t s_out[SMDP_WAVES_PER_WG * kCOut];
  WaveOut<kCOut> out;
  out.init(s_out + (threadIdx.x / SMDP_WAVE) * kCOut);
  // colour 1 exists whenever n > 0, even if none of its vertices is queued
  if (blockIdx.x == 0 && threadIdx.x == 0 && a.n > 0) atomicMax(&cnt_of(a.cnt)->rounds, 1ull);
  for (int base = blockIdx.x * SMDP_WG; base < a.n; base += gridDim.x * SMDP_WG) {
    const int v = base + (int)threadIdx.x;
    bool ready = false, queue = false;
    if (v < a.n) {
      ready = a.pend[v] == 0;
      a.color[v] = ready ? 1 : 0;
      queue = ready && a.trow[v + 1] > a.trow[v];
    }
    out.push(queue, v, a.q0, &a.ctl[0].qlen);
  }
  out.flush(a.q0, &a.ctl[0].qlen);
}

// Arcs u[t] -> v[t] of the transposed CSR, K per lane (ok[t]: the arc exists;
// v[t] is a valid vertex for every t).  An arc whose source precedes its
// target decrements pend[v]; the lane that brings it to zero colours v with
// newc and queues it.  pv[t]: the priority of v[t], loaded with the arc from
// tprio (unused without priorities).
// The decrements are packed to the low lanes through the wave's LDS area
// (s_pack: K * 64 ints) and issued one per lane: K conditional atomics ran
// one after another, each waiting for its result (about half of the arcs
// decrement).  Unconditional atomics that subtract 0 were measured 38x
// slower, since the padded slots of a step's last pass all hit one vertex.
// Returns the vertices this wave queued on lane 0, 0 on the other lanes (so a
// wave_sum of the running totals counts each once).
template <int K>
__device__ __forceinline__ int notify_packed(const LArgs& a, const int* u, const uint32_t* pu, const int* v,
                                             const uint32_t* pv, const bool* ok, int newc, int32_t* s_pack,
                                             WaveOut<kCOut>& out, int32_t* qout, int32_t* qlen) {
  bool want[K];
  if (a.prio) {
#pragma unroll
    for (int t = 0; t < K; ++t) want[t] = ok[t] && (pu[t] < pv[t] || (pu[t] == pv[t] && u[t] < v[t]));
  } else {
#pragma unroll
    for (int t = 0; t < K; ++t) want[t] = ok[t] && u[t] < v[t];
  }
  int total = 0;
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const uint64_t m = ballot(want[t]);
    if (want[t]) s_pack[total + mbcnt(m)] = v[t];
    total += __popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int pushes = 0;
  for (int r0 = 0; r0 < total; r0 += SMDP_WAVE) {
    const int i = r0 + lane_id();
    int vv = 0;
    bool won = false;
    if (i < total) {
      vv = s_pack[i];
      won = atomicSub(&a.pend[vv], 1) == 1;
      if (won) a.color[vv] = newc;
    }
    pushes += out.push(won, vv, qout, qlen);
  }
  // s_pack is written again by the next step
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return lane_id() == 0 ? pushes : 0;
}

// Shape of round r's work, from its queue length L and the grid (the expand
// and drain kernels of a round run on the same grid, so both derive the same
// values).  Tiles hold 256 queue entries, halved while the tiles still number
// at most one per workgroup: a short layer (RMAT-23 has ~2300 layers of about
// 1200 vertices) then spreads over the whole grid instead of a few
// workgroups.  In such short rounds rows are cut into chunks of at most one
// pass of the workgroup (kLayerIlp * 256 arcs).
struct LayerShape {
  int tile, split, chunk;
};
constexpr int kLayerShortChunk = kLayerIlp * SMDP_WG;
__device__ __forceinline__ LayerShape layer_shape(const LArgs& a, int L, int grid) {
  LayerShape sh;
  sh.tile = SMDP_WG;
  while (sh.tile > 1 && (L + sh.tile / 2 - 1) / (sh.tile / 2) <= grid) sh.tile >>= 1;
  const bool short_round = sh.tile < SMDP_WG;
  sh.split = short_round ? min(a.split, kLayerShortChunk) : a.split;
  sh.chunk = short_round ? min(a.chunk, kLayerShortChunk) : a.chunk;
  return sh;
}

// Layers of at most kShortL vertices (most of RMAT-23's 2373) are expanded
// arc-balanced: G = min(grid, max(L / kShortPerWG, kShortMinWG)) workgroups
// each stage the whole layer (vertex, first arc, arc offset) in LDS and take
// an equal share of its arcs, so a layer's hubs are split over the group
// without grid chunks and the round needs no drain.
constexpr int kShortL = 2048;
constexpr int kShortPerWG = 8;
constexpr int kShortMinWG = 128;

// Expands short layer r (L <= kShortL queued vertices) on workgroups
// 0 .. G-1 of the launch, G = min(workgroups, max(L / kShortPerWG,
// kShortMinWG)); the others return at once.  s_raw: 3 * kShortL ints.
template <bool COH>
__device__ __forceinline__ void short_layer(const LArgs& a, int r, int L, int32_t* s_raw, int* s_scan,
                                            int32_t* s_pack, WaveOut<kCOut>& out, unsigned long long& edges,
                                            unsigned long long& verts, unsigned long long& pushes) {
  const int G = min((int)gridDim.x, max((L + kShortPerWG - 1) / kShortPerWG, kShortMinWG));
  if ((int)blockIdx.x >= G) return;
  Ctl* cout = a.ctl + ((r + 1) & 3);
  const int32_t* qin = (r & 1) ? a.q1 : a.q0;
  int32_t* qout = (r & 1) ? a.q0 : a.q1;
  const int newc = r + 2;
  int* s_v = s_raw;
  int* s_st = s_raw + kShortL;
  int* s_of = s_raw + 2 * kShortL;
  constexpr int E = kShortL / SMDP_WG;  // consecutive entries staged per thread
  int vv[E], st[E], en[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = min((int)threadIdx.x * E + k, L - 1);
    vv[k] = COH ? __hip_atomic_load(qin + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : qin[i];
  }
#pragma unroll
  for (int k = 0; k < E; ++k) {
    st[k] = a.trow[vv[k]];
    en[k] = a.trow[vv[k] + 1];
  }
  int sum = 0;
#pragma unroll
  for (int k = 0; k < E; ++k) sum += (int)threadIdx.x * E + k < L ? en[k] - st[k] : 0;
  int A;
  int run = wg_exclusive_scan(sum, s_scan, &A);
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int i = (int)threadIdx.x * E + k;
    if (i < L) {
      s_v[i] = vv[k];
      s_st[i] = st[k];
      s_of[i] = run;
      run += en[k] - st[k];
    } else {
      s_of[i] = A;  // never the owner of an arc j < A
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) verts += L;
  const int a0 = (int)((int64_t)A * blockIdx.x / G), a1 = (int)((int64_t)A * (blockIdx.x + 1) / G);
  for (int jb = a0; jb < a1; jb += kLayerIlp * SMDP_WG) {
    int j[kLayerIlp], own[kLayerIlp], v[kLayerIlp], uu[kLayerIlp], k[kLayerIlp];
    uint32_t pu[kLayerIlp], pv[kLayerIlp];
    bool ok[kLayerIlp];
#pragma unroll
    for (int t = 0; t < kLayerIlp; ++t) j[t] = min(jb + (int)threadIdx.x + t * SMDP_WG, a1 - 1);
    tile_owners<kLayerIlp, kShortL>(s_of, j, own);
#pragma unroll
    for (int t = 0; t < kLayerIlp; ++t) {
      uu[t] = s_v[own[t]];
      k[t] = s_st[own[t]] + (j[t] - s_of[own[t]]);
      v[t] = stream_load(a.tcol + k[t]);
      ok[t] = jb + (int)threadIdx.x + t * SMDP_WG < a1 && v[t] != uu[t];
      edges += jb + (int)threadIdx.x + t * SMDP_WG < a1;
    }
    if (a.prio) {
#pragma unroll
      for (int t = 0; t < kLayerIlp; ++t) {
        pv[t] = stream_load(a.tprio + k[t]);
        pu[t] = a.prio[uu[t]];
      }
    } else {
#pragma unroll
      for (int t = 0; t < kLayerIlp; ++t) pu[t] = pv[t] = 0u;
    }
    pushes += notify_packed<kLayerIlp>(a, uu, pu, v, pv, ok, newc,
                                       s_pack + (threadIdx.x / SMDP_WAVE) * kLayerIlp * SMDP_WAVE, out, qout,
                                       &cout->qlen);
  }
}

// Round r: the queue holds the vertices of colour r+1.  Short layers as
// above; longer ones in edge-balanced tiles of queue entries (layer_shape),
// with transposed rows of >= split arcs cut into grid chunks.
__global__ __launch_bounds__(SMDP_WG) void k_layer_expand(LArgs a) {
  const int r = a.rd[0];  // written by the kernels before this one
  ctl_clear(a.ctl, (r + 2) & 3);
  Ctl* cin = a.ctl + (r & 3);
  Ctl* cout = a.ctl + ((r + 1) & 3);
  const int L = cin->qlen;
  if (L == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) a.rd[2] = 1;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&cnt_of(a.cnt)->rounds, (unsigned long long)(r + 1));
  const int32_t* qin = (r & 1) ? a.q1 : a.q0;
  int32_t* qout = (r & 1) ? a.q0 : a.q1;
  const int newc = r + 2;
  const int lane = lane_id();
  __shared__ int32_t s_raw[3 * kShortL];  // short layers: vertex, first arc, arc offset; tiles: see below
  __shared__ int s_tile;
  __shared__ int s_scan[SMDP_WAVES_PER_WG];
  __shared__ int32_t s_out[SMDP_WAVES_PER_WG * kCOut];
  __shared__ int32_t s_pack[SMDP_WAVES_PER_WG * kLayerIlp * SMDP_WAVE];  // notify_packed
  WaveOut<kCOut> out;
  out.init(s_out + (threadIdx.x / SMDP_WAVE) * kCOut);
  unsigned long long edges = 0, verts = 0, pushes = 0, pg = 0;

  if (L <= kShortL) {
    short_layer<false>(a, r, L, s_raw, s_scan, s_pack, out, edges, verts, pushes);
  } else {
  const LayerShape sh = layer_shape(a, L, a.grid);  // launched with a.grid workgroups
  const int ntiles = (L + sh.tile - 1) / sh.tile;
  // workgroups past the tile count would only contend for the counter
  if ((int)blockIdx.x >= ntiles) return;
  // at most one tile per workgroup: tile = workgroup, no claims
  const bool one_each = ntiles <= (int)gridDim.x;
  int* s_start = s_raw;
  int* s_u = s_raw + SMDP_WG;
  uint32_t* s_pu = reinterpret_cast<uint32_t*>(s_raw + 2 * SMDP_WG);
  int* s_off = s_raw + 3 * SMDP_WG;
  int part = 0;  // thread 0: claim cursor
  for (int pass = 0;; ++pass) {
    int tile;
    if (one_each) {
      if (pass) break;
      tile = blockIdx.x;
    } else {
      if (threadIdx.x == 0) {
        int f = 0, e = 0;
        s_tile = claim_items(cin->tiles, ntiles, 1, part, &f, &e) ? f : -1;
      }
      __syncthreads();
      tile = s_tile;
      __syncthreads();
      if (tile < 0) break;
    }
    const int i = tile * sh.tile + (int)threadIdx.x;
    const bool valid = (int)threadIdx.x < sh.tile && i < L;
    int u = 0, s = 0, deg = 0;
    uint32_t pu = 0;
    if (valid) {
      u = qin[i];
      s = a.trow[u];
      deg = a.trow[u + 1] - s;
      if (a.prio) pu = a.prio[u];
      ++verts;
    }
    const bool big = deg >= sh.split;
    const int nch = big ? (deg + sh.chunk - 1) / sh.chunk : 0;
    const int incl = wave_inclusive_scan(nch);
    const int wtot = __shfl(incl, SMDP_WAVE - 1);
    if (wtot > 0) {
      int b0 = 0;
      if (lane == 0) b0 = atomicAdd(&cin->hlen, wtot);
      b0 = __shfl(b0, 0);
      for (int k = 0; k < nch; ++k) a.hq[b0 + incl - nch + k] = Chunk{u, s + k * sh.chunk};
      pg += nch;
    }
    s_start[threadIdx.x] = s;
    s_u[threadIdx.x] = u;
    s_pu[threadIdx.x] = pu;
    int tot;
    const int ex = wg_exclusive_scan(big ? 0 : deg, s_scan, &tot);
    s_off[threadIdx.x] = ex;
    __syncthreads();
    for (int jb = 0; jb < tot; jb += kLayerIlp * SMDP_WG) {
      int j[kLayerIlp], own[kLayerIlp], v[kLayerIlp], uu[kLayerIlp], k[kLayerIlp];
      uint32_t pu[kLayerIlp], pv[kLayerIlp];
      bool ok[kLayerIlp];
#pragma unroll
      for (int t = 0; t < kLayerIlp; ++t) j[t] = min(jb + (int)threadIdx.x + t * SMDP_WG, tot - 1);
      tile_owners<kLayerIlp>(s_off, j, own);
#pragma unroll
      for (int t = 0; t < kLayerIlp; ++t) {
        k[t] = s_start[own[t]] + (j[t] - s_off[own[t]]);
        v[t] = stream_load(a.tcol + k[t]);
        uu[t] = s_u[own[t]];
        pu[t] = s_pu[own[t]];
        ok[t] = jb + (int)threadIdx.x + t * SMDP_WG < tot && v[t] != uu[t];
        edges += jb + (int)threadIdx.x + t * SMDP_WG < tot;
      }
      if (a.prio) {
#pragma unroll
        for (int t = 0; t < kLayerIlp; ++t) pv[t] = stream_load(a.tprio + k[t]);
      } else {
#pragma unroll
        for (int t = 0; t < kLayerIlp; ++t) pv[t] = 0u;
      }
      pushes += notify_packed<kLayerIlp>(a, uu, pu, v, pv, ok, newc,
                                         s_pack + (threadIdx.x / SMDP_WAVE) * kLayerIlp * SMDP_WAVE, out, qout,
                                         &cout->qlen);
    }
    __syncthreads();
  }
  }
  out.flush(qout, &cout->qlen);
  edges = wave_sum(edges);
  verts = wave_sum(verts);
  pushes = wave_sum(pushes);
  pg = wave_sum(pg);
  if (lane == 0) {
    if (edges) atomicAdd(&cnt_of(a.cnt)->edges, edges);
    if (verts) atomicAdd(&cnt_of(a.cnt)->verts, verts);
    if (pushes) atomicAdd(&cnt_of(a.cnt)->push[0], pushes);
    if (pg) atomicAdd(&cnt_of(a.cnt)->push[3], pg);
  }
}

// Runs short layers one after another in one launch: layers rd[1] + 1,
// rd[1] + 2, ... while they hold at most kShortL vertices, with a flag
// barrier (below) after each (a layer's queue is complete only when every workgroup
// has pushed).  Then rd[0] = the first layer not run, for k_layer_expand.
// RMAT-23 has ~2300 such layers; one expand and one drain launch per layer
// cost about 18 us, mostly launch and ramp.
constexpr int kFlagGrid = 64;  // workgroups the flag barrier serves (one flag per lane of a wave)
static_assert(2 * kFlagGrid * 2 <= kBarLines * 32, "two parities of 64-bit flags fit the barrier lines");

// Flag barrier between the short layers (at most kFlagGrid workgroups).
// Workgroup b stores (epoch << 32 | entries it queued in this layer) in its
// flag of the epoch's parity; wave 0 reads every flag, one per lane, until
// all carry the epoch, and the sum of their entries is the next layer's
// length, returned to the whole workgroup (-1 after about a second of
// polling: not every workgroup is resident).  A workgroup writes the flags
// of one parity again only at epoch e + 2, after every workgroup stored its
// epoch e + 1 flag, which each does after it read all of epoch e: no read of
// epoch e sees a later count.  queued: this wave's entries (read on lane 0).
//
// What crosses workgroups between layers is stored and read device-coherent
// (queue entries, the cleared control slot, the flags: global_store /
// global_load with sc1), and each wave waits for its stores to complete
// before the flag is written, so no agent-scope release (an L2 write-back)
// or acquire (an L2 invalidate) is needed.  Colours are read only after the
// kernel; pend[] changes only by atomics, which are performed beyond L2.
// Against the two-level atomic barrier with agent-scope fences, RMAT-23's
// colouring went from 35.4 to 33.9 ms (flags with the fences kept: 34.6).
__device__ __forceinline__ int flag_barrier(uint64_t* flags, uint32_t epoch, int queued, int* s_cnt, int* s_len) {
  if (lane_id() == 0 && queued) atomicAdd(s_cnt, queued);
  __builtin_amdgcn_s_waitcnt(0);  // this wave's coherent stores are complete
  __syncthreads();
  uint64_t* f = flags + (epoch & 1) * kFlagGrid;
  if (threadIdx.x < SMDP_WAVE) {
    const int lane = threadIdx.x;
    const int G = gridDim.x;
    if (lane == 0) {
      const uint64_t mine = ((uint64_t)epoch << 32) | (uint32_t)*s_cnt;
      *s_cnt = 0;  // the next adds come after the __syncthreads below
      __hip_atomic_store(f + blockIdx.x, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    uint64_t v = 0;
    bool ok = true;
    for (int spins = 0;;) {
      v = lane < G ? __hip_atomic_load(f + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   : (uint64_t)epoch << 32;
      if (ballot((uint32_t)(v >> 32) < epoch) == 0) break;
      __builtin_amdgcn_s_sleep(2);
      if (++spins > kBarSpinLimit) {
        ok = false;
        break;
      }
    }
    const unsigned long long tot = wave_sum((unsigned long long)(lane < G ? (uint32_t)v : 0u));
    if (lane == 0) *s_len = ok ? (int)tot : -1;
  }
  __syncthreads();
  return *s_len;
}

// ctl_clear with device-coherent stores (no L2 write-back publishes them).
__device__ __forceinline__ void ctl_clear_coherent(Ctl* ctl, int slot) {
  if (blockIdx.x != 0) return;
  int32_t* c = reinterpret_cast<int32_t*>(ctl + slot);
  for (int i = threadIdx.x; i < kCtlInts; i += blockDim.x)
    __hip_atomic_store(c + i, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(SMDP_WG) void k_layer_short(LArgs a) {
  __shared__ int32_t s_raw[3 * kShortL];
  __shared__ int s_scan[SMDP_WAVES_PER_WG];
  __shared__ int32_t s_out[SMDP_WAVES_PER_WG * kCOut];
  __shared__ int32_t s_pack[SMDP_WAVES_PER_WG * kLayerIlp * SMDP_WAVE];  // notify_packed
  __shared__ int s_cnt, s_len;
  const int r0 = a.rd[1];
  if (a.ctl[r0 & 3].qlen == 0) return;  // the colouring is complete (every workgroup sees it)
  if (threadIdx.x == 0) s_cnt = 0;
  WaveOut<kCOut> out;
  out.init(s_out + (threadIdx.x / SMDP_WAVE) * kCOut, INT32_MAX, true);
  unsigned long long edges = 0, verts = 0, pushes = 0;
  uint64_t* flags = reinterpret_cast<uint64_t*>(a.bar);
  int r = r0 + 1;
  int L = a.ctl[r & 3].qlen;  // written by the kernels before this one
  for (uint32_t epoch = 1;; ++epoch) {
    if (L == 0 || L > kShortL) break;
    ctl_clear_coherent(a.ctl, (r + 2) & 3);  // workgroup 0; no one uses that slot until layer r + 1
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&cnt_of(a.cnt)->rounds, (unsigned long long)(r + 1));
    const unsigned long long before = pushes;
    short_layer<true>(a, r, L, s_raw, s_scan, s_pack, out, edges, verts, pushes);
    out.flush((r & 1) ? a.q0 : a.q1, &a.ctl[(r + 1) & 3].qlen);
    L = flag_barrier(flags, epoch, (int)(pushes - before), &s_cnt, &s_len);
    if (L < 0) {
      if (threadIdx.x == 0) a.rd[3] = 1;
      break;
    }
    ++r;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) a.rd[0] = r;
  edges = wave_sum(edges);
  verts = wave_sum(verts);
  pushes = wave_sum(pushes);
  if (lane_id() == 0) {
    if (edges) atomicAdd(&cnt_of(a.cnt)->edges, edges);
    if (verts) atomicAdd(&cnt_of(a.cnt)->verts, verts);
    if (pushes) atomicAdd(&cnt_of(a.cnt)->push[0], pushes);
  }
}

// Grid chunks of round r: a workgroup takes one chunk of a transposed row.
__global__ __launch_bounds__(SMDP_WG) void k_layer_drain(LArgs a) {
  const int r = a.rd[0];
  Ctl* cin = a.ctl + (r & 3);
  Ctl* cout = a.ctl + ((r + 1) & 3);
  if (blockIdx.x == 0) {
    // for k_layer_short, which runs next: the layer done here, and fresh
    // barrier flags (no other kernel is using them now)
    if (threadIdx.x == 0) a.rd[1] = r;
    for (int i = threadIdx.x; i < kBarLines * 32; i += blockDim.x) a.bar[i] = 0;
  }
  const int H = cin->hlen;
  if ((int)blockIdx.x >= H) return;
  const int chunk = layer_shape(a, cin->qlen, a.grid).chunk;
  int32_t* qout = (r & 1) ? a.q0 : a.q1;
  const int newc = r + 2;
  __shared__ int s_idx;
  __shared__ int32_t s_out[SMDP_WAVES_PER_WG * kCOut];
  __shared__ int32_t s_pack[SMDP_WAVES_PER_WG * kLayerIlp * SMDP_WAVE];  // notify_packed
  WaveOut<kCOut> out;
  out.init(s_out + (threadIdx.x / SMDP_WAVE) * kCOut);
  unsigned long long edges = 0, pushes = 0;
  int part = 0;  // thread 0: claim cursor
  for (;;) {
    if (threadIdx.x == 0) {
      int f = 0, e = 0;
      s_idx = claim_items(cin->chunks, H, 1, part, &f, &e) ? f : -1;
    }
    __syncthreads();
    const int idx = s_idx;
    __syncthreads();
    if (idx < 0) break;
    const Chunk ch = a.hq[idx];
    const int e = min(a.trow[ch.v + 1], ch.first + chunk);
    const uint32_t pu1 = a.prio ? a.prio[ch.v] : 0u;
    for (int b = ch.first; b < e; b += kLayerIlp * SMDP_WG) {
      int v[kLayerIlp], uu[kLayerIlp];
      uint32_t pu[kLayerIlp], pv[kLayerIlp];
      bool ok[kLayerIlp];
#pragma unroll
      for (int t = 0; t < kLayerIlp; ++t) {
        const int k = b + (int)threadIdx.x + t * SMDP_WG;
        v[t] = stream_load(a.tcol + min(k, e - 1));
        pv[t] = a.prio ? stream_load(a.tprio + min(k, e - 1)) : 0u;
        uu[t] = ch.v;
        pu[t] = pu1;
        ok[t] = k < e && v[t] != ch.v;
        edges += k < e;
      }
      pushes += notify_packed<kLayerIlp>(a, uu, pu, v, pv, ok, newc,
                                         s_pack + (threadIdx.x / SMDP_WAVE) * kLayerIlp * SMDP_WAVE, out, qout,
                                         &cout->qlen);
    }
  }
  out.flush(qout, &cout->qlen);
  edges = wave_sum(edges);
  pushes = wave_sum(pushes);
  if (lane_id() == 0) {
    if (edges) {
      atomicAdd(&cnt_of(a.cnt)->edges, edges);
      atomicAdd(&cnt_of(a.cnt)->edges2, edges);
    }
    if (pushes) atomicAdd(&cnt_of(a.cnt)->push[0], pushes);
  }
}

// ---- transposed CSR (built once per resident graph) ----
__global__ void k_tdeg(const int32_t* __restrict__ col, int64_t m, int32_t* tdeg) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&tdeg[col[e]], 1);
}

// Each arc finds its row by binary search over row[] (consecutive arcs take
// the same path, which stays in cache) and appends its row to tcol.
__global__ void k_tfill(const int32_t* __restrict__ row, const int32_t* __restrict__ col, int32_t n, int64_t m,
                        int32_t* tcur, int32_t* tcol) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = n - 1;  // the last v with row[v] <= e
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (row[mid] <= e) lo = mid;
      else hi = mid - 1;
    }
    tcol[atomicAdd(&tcur[col[e]], 1)] = lo;
  }
}

// prio of each arc's vertex: out[e] = prio[col[e]].
__global__ void k_arc_prio(const int32_t* __restrict__ col, int64_t m, const uint32_t* __restrict__ prio,
                           uint32_t* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = prio[col[e]];
}

}  // namespace smdp

namespace {
struct ColorBufs {
  int32_t* resume = nullptr;
  int32_t* rmax = nullptr;
  int32_t* pend = nullptr;   // tuned: preceding arcs not yet coloured
  int32_t* trow = nullptr;   // tuned: transposed CSR, built once per graph
  int32_t* tcol = nullptr;
  int32_t* rd = nullptr;     // tuned: device layer indices (LArgs::rd)
  uint32_t* bar = nullptr;   // tuned: barrier flags of k_layer_short
  uint32_t* cprio = nullptr;  // tuned, with priorities: prio[col[e]] per arc
  uint32_t* tprio = nullptr;  // tuned, with priorities: prio[tcol[e]] per arc
  uint32_t prio_version = 0;  // the graph's priority upload cprio/tprio hold (0: none)
};
}  // namespace

namespace smdp {

void colorbufs_free(smdp_graph* g) {
  ColorBufs* b = static_cast<ColorBufs*>(g->colorbufs);
  if (!b) return;
  if (b->resume) (void)hipFree(b->resume);
  if (b->rmax) (void)hipFree(b->rmax);
  if (b->pend) (void)hipFree(b->pend);
  if (b->trow) (void)hipFree(b->trow);
  if (b->tcol) (void)hipFree(b->tcol);
  if (b->rd) (void)hipFree(b->rd);
  if (b->bar) (void)hipFree(b->bar);
  if (b->cprio) (void)hipFree(b->cprio);
  if (b->tprio) (void)hipFree(b->tprio);
  delete b;
  g->colorbufs = nullptr;
}

// The transposed CSR: trow from a count of each vertex's appearances in col
// (prefix sum on the host), tcol filled by one atomic cursor per vertex.
static int build_transpose(smdp_graph* g, ColorBufs* b) {
  const int64_t n = g->n, m = g->m;
  hipStream_t s = g->ctx->stream;
  SMDP_CHECK(alloc_once(&b->trow, n + 1));
  SMDP_CHECK(alloc_once(&b->tcol, std::max<int64_t>(1, m)));
  DeviceScratch<int32_t> cur;
  SMDP_CHECK(alloc_once(&cur.p, std::max<int64_t>(1, n)));
  int32_t* tcur = cur.p;
  SMDP_HIP(hipMemsetAsync(tcur, 0, sizeof(int32_t) * n, s));
  const int grid = std::max(1, g->ctx->cus * 8);
  if (m) hipLaunchKernelGGL(k_tdeg, dim3(grid), dim3(256), 0, s, g->col, m, tcur);
  SMDP_HIP(hipGetLastError());
  std::vector<int32_t> deg((size_t)n), off((size_t)n + 1);
  SMDP_HIP(hipMemcpyAsync(deg.data(), tcur, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  SMDP_HIP(hipStreamSynchronize(s));
  off[0] = 0;
  for (int64_t v = 0; v < n; ++v) off[(size_t)v + 1] = off[(size_t)v] + deg[(size_t)v];
  SMDP_HIP(hipMemcpyAsync(b->trow, off.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
  SMDP_HIP(hipMemcpyAsync(tcur, off.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  if (m) hipLaunchKernelGGL(k_tfill, dim3(grid), dim3(256), 0, s, g->row, g->col, g->n, m, tcur, b->tcol);
  SMDP_HIP(hipGetLastError());
  SMDP_HIP(hipStreamSynchronize(s));
  return SMDP_OK;
}

// A build that failed part of the way leaves nothing behind, so that a later
// call does not take half-filled arrays for the transposed CSR.
static int ensure_transpose(smdp_graph* g, ColorBufs* b) {
  if (b->trow) return SMDP_OK;
  const int rc = build_transpose(g, b);
  if (rc != SMDP_OK) {
    if (b->trow) (void)hipFree(b->trow);
    if (b->tcol) (void)hipFree(b->tcol);
    b->trow = nullptr;
    b->tcol = nullptr;
  }
  return rc;
}

// Tuned solution: topological layering (see the kernels above).
static int run_color_layers(smdp_graph* g, int use_prio, smdp_stats* st) {
  smdp_ctx* ctx = g->ctx;
  ColorBufs* b = static_cast<ColorBufs*>(g->colorbufs);
  const smdp_params& p = params_of(ctx, SMDP_APP_COLOR);
  const int64_t n1 = std::max<int64_t>(1, g->n);
  SMDP_CHECK(alloc_once(&g->dist, n1));
  SMDP_CHECK(alloc_once(&g->q[0], n1));
  SMDP_CHECK(alloc_once(&g->q[1], n1));
  SMDP_CHECK(alloc_once(&b->pend, n1));
  // chunks of one round (or of the count) never exceed n + m/chunk, with the
  // chunk of short rounds at most kLayerShortChunk arcs
  SMDP_CHECK(ensure(&g->hq, &g->hq_cap,
                    n1 + (int64_t)g->m / std::max(1, std::min(p.chunk, kLayerShortChunk)) + 64));
  SMDP_CHECK(ensure_transpose(g, b));
  SMDP_CHECK(alloc_once(&b->rd, 4));
  SMDP_CHECK(alloc_once(&b->bar, kBarLines * 32));
  if (use_prio && b->prio_version != g->prio_version) {
    // the arcs' priorities, aligned with col and tcol: once per priority upload
    const int64_t m1 = std::max<int64_t>(1, g->m);
    SMDP_CHECK(alloc_once(&b->cprio, m1));
    SMDP_CHECK(alloc_once(&b->tprio, m1));
    const int pg = std::max(1, ctx->cus * 8);
    if (g->m) {
      hipLaunchKernelGGL(k_arc_prio, dim3(pg), dim3(256), 0, ctx->stream, g->col, g->m, g->prio, b->cprio);
      hipLaunchKernelGGL(k_arc_prio, dim3(pg), dim3(256), 0, ctx->stream, b->tcol, g->m, g->prio, b->tprio);
      SMDP_HIP(hipGetLastError());
    }
    b->prio_version = g->prio_version;
  }
  LArgs a;
  a.row = g->row;
  a.col = g->col;
  a.trow = b->trow;
  a.tcol = b->tcol;
  a.prio = use_prio ? g->prio : nullptr;
  a.cprio = use_prio ? b->cprio : nullptr;
  a.tprio = use_prio ? b->tprio : nullptr;
  a.color = g->dist;
  a.pend = b->pend;
  a.q0 = g->q[0];
  a.q1 = g->q[1];
  a.hq = g->hq;
  a.ctl = g->ctl;
  a.cnt = g->cnt;
  a.n = g->n;
  a.heavy = p.heavy_threshold;
  a.split = std::max(p.split_threshold, p.heavy_threshold);
  a.chunk = p.chunk;
  const int grid = grid_size(ctx, p);
  const int drain_grid = std::max(1, std::min(grid, ctx->cus));
  a.grid = grid;
  a.rd = b->rd;
  a.bar = reinterpret_cast<uint32_t*>(b->bar);
  // k_layer_short: its workgroups meet at barriers between layers, so all
  // must be resident at once.  The grid is a quarter of the CUs (its barrier costs
  // more per workgroup than a short layer gains from more of them: on
  // RMAT-23, 256/128/64/32/16/8 workgroups took 48.6/43.0/39.1/42.6/56.0/
  // 83.8 ms per colouring), capped at what the device holds, and launched
  // after the stream's previous kernels finished, so it fits on the idle
  // device; a barrier that waits too long (another process holding CUs)
  // makes the run return SMDP_ERR_DEVICE instead of hanging.  A plain launch
  // rather than a cooperative one: rocprofv3 ended with a segmentation fault
  // in its exit handlers after cooperative launches.
  int per_cu = 0;
  SMDP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_layer_short),
                                                          SMDP_WG, 0));
  if (per_cu < 1) return SMDP_ERR_DEVICE;
  const int short_grid = std::min({std::max(8, ctx->cus / 4), per_cu * ctx->cus, kFlagGrid});
  hipStream_t s = ctx->stream;
  prof_reset(ctx);
  int launches = 0;
  SMDP_HIP(hipEventRecord(ctx->ev_begin, s));
  hipLaunchKernelGGL(k_layer_init, dim3(1), dim3(256), 0, s, g->ctl, g->cnt, a.rd);
  hipLaunchKernelGGL(k_layer_count, dim3(grid), dim3(SMDP_WG), 0, s, a);
  hipLaunchKernelGGL(k_layer_count_chunks, dim3(grid), dim3(SMDP_WG), 0, s, a);
  hipLaunchKernelGGL(k_layer_seed, dim3(std::min<int64_t>((n1 + SMDP_WG - 1) / SMDP_WG, grid)), dim3(SMDP_WG), 0, s, a);
  launches += 4;
  // A step is one layer on the whole grid (expand, then the drain of its
  // chunks) followed by the run of short layers after it in one launch.  The
  // layer indices live on the device (rd), so steps are enqueued in batches
  // and the host only checks rd between batches.  A step enqueued after the
  // last one with work still costs ~20 us (three launches),
  // so batches stop growing at 8: on RMAT-23 (at most 148 steps with work)
  // caps of 64, 16 and 8 enqueued 188, 156 and 148 steps and took 38.9, 38.1
  // and 38.0 ms; sizing batches from the next layer's decay per step
  // overshot (203 steps).
  int batch = std::max(1, p.batch_rounds);
  for (;;) {
    for (int j = 0; j < batch; ++j) {
      SMDP_CHECK(prof_mark(ctx, 0));
      hipLaunchKernelGGL(k_layer_expand, dim3(grid), dim3(SMDP_WG), 0, s, a);
      SMDP_CHECK(prof_mark(ctx, 0));
      SMDP_CHECK(prof_mark(ctx, 1));
      // one workgroup per CU claims the round's chunks: most rounds have none,
      // and an empty launch of the full grid cost 4.3 us per round
      hipLaunchKernelGGL(k_layer_drain, dim3(drain_grid), dim3(SMDP_WG), 0, s, a);
      SMDP_CHECK(prof_mark(ctx, 1));
      SMDP_CHECK(prof_mark(ctx, 0));
      hipLaunchKernelGGL(k_layer_short, dim3(short_grid), dim3(SMDP_WG), 0, s, a);
      SMDP_CHECK(prof_mark(ctx, 0));
      launches += 3;
    }
    SMDP_HIP(hipGetLastError());
    SMDP_HIP(hipMemcpyAsync(ctx->pinned, a.rd, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SMDP_HIP(hipStreamSynchronize(s));
    if (ctx->pinned[3]) return SMDP_ERR_DEVICE;  // a barrier between short layers gave up waiting
    if (ctx->pinned[2]) break;
    if (ctx->pinned[0] >= p.max_rounds || ctx->pinned[0] >= g->n + 1) return SMDP_ERR_OVERFLOW;
    batch = std::min(batch * 2, 8);
  }
  SMDP_HIP(hipEventRecord(ctx->ev_end, s));
  SMDP_HIP(hipEventSynchronize(ctx->ev_end));
  float ms = 0;
  SMDP_HIP(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
  g->last_res = SMDP_RES_COLOR;
  g->dist_perm = false;
  if (st) {
    SMDP_CHECK(fill_stats(g, st, ms, launches));
    st->m_reached = g->m;
    SMDP_CHECK(prof_collect(ctx, st));
  } else {
    prof_reset(ctx);
  }
  return SMDP_OK;
}

int run_color(smdp_graph* g, int use_prio, int sol, smdp_stats* st) {
  if (!g) return SMDP_ERR_ARG;
  if (sol < 0 || sol >= SMDP_SOL_COUNT || sol == SMDP_SOL_HIER) return SMDP_ERR_SOLUTION;
  if (use_prio && !g->prio) return SMDP_ERR_ARG;
  smdp_ctx* ctx = g->ctx;
  SMDP_HIP(hipSetDevice(ctx->device));
  begin_run(g);
  if (!g->colorbufs) g->colorbufs = new ColorBufs();
  ColorBufs* b = static_cast<ColorBufs*>(g->colorbufs);
  if (sol == SMDP_SOL_TUNED && g->n > 0) return run_color_layers(g, use_prio, st);
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
  const int grid = grid_size(ctx, p);
  const bool two_phase = sol == SMDP_SOL_GRID;
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
          default: hipLaunchKernelGGL(k_color<SMDP_SOL_GRID>, gd, bd, 0, s, a, r); break;
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
  g->dist_perm = false;
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
