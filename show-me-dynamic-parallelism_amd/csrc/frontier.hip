// Frontier expansion for SSSP and BFS on gfx950.
//
// Reference algorithm (apps/sssp/src/sssp_kernel.cu:40-61, :496-506 and
// apps/sssp/src/sssp_wrapper.cu:147-302): unordered worklist Bellman-Ford.
// Every round takes the queue of vertices whose cost dropped, and for each
// arc (u,v,w) of a queued u does "alt = cost[u] + w; if (cost[v] > alt)
// atomicMin(cost[v], alt) and mark v".  Marked vertices form the next queue.
// BFS levels (apps/bfs-rec/src/bfs_rec_kernel.cu:42-141) are the same
// relaxation with w = 1 and the 0x7FFFFFFF sentinel.
//
// What is different here:
//  * The next queue is built while relaxing: the lane whose atomicMin lowers
//    cost[v] claims v for the next round through stamp[v] (one atomicExch per
//    improvement), and claims are compacted with one atomic per wave
//    (ballot + mbcnt).  The reference instead rescans all N update flags every
//    round (unorder_generateQueue_kernel).
//  * Heavy vertices (degree >= threshold) are never handed to a child grid.
//    They go to a per-wave LDS queue, a per-workgroup LDS queue or a grid-wide
//    HBM queue, according to the solution, and are drained by waves,
//    workgroups or the whole grid (phase 2 kernel).
//  * Rounds are launched back to back with the round number as an argument;
//    each kernel reads its queue length from device memory and returns at
//    once when it is zero, so the host only reads the queue length back once
//    per batch of rounds instead of once per round.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "frontier.h"
#include "smdp_internal.h"

namespace smdp {

constexpr int kWaveQ = 256;   // per-wave LDS queue capacity (SMDP_SOL_WAVE)
constexpr int kBlockQ = 1024; // per-workgroup LDS queue capacity (SMDP_SOL_BLOCK)
constexpr int kIlp = 4;       // arcs per thread per step in the tuned kernels
constexpr int kTileGrab = 1;  // frontier tiles claimed per atomic
constexpr int kSmallSplit = 32;  // split degree of the tuned solution when the queue is short
constexpr int kDrainTile = 8;    // grid-queue entries per drain tile (a power of two, at most 64)

struct FArgs {
  const int32_t* __restrict__ row;
  const int32_t* __restrict__ col;
  const int32_t* __restrict__ w;
  int32_t* dist;
  int32_t* stamp;
  int32_t* q0;
  int32_t* q1;
  Chunk* hq;
  Ctl* ctl;
  Counters* cnt;
  int32_t heavy;  // lane solutions: heavy degree; tuned: split degree
  int32_t chunk;  // grid-queue chunk in edges; 0 = whole vertex
  int32_t small_q;  // tuned: a queue shorter than this splits vertices of >= kSmallSplit arcs
  int32_t forest;   // tuned BFS: no vertex has two in-arcs, so levels are stored without atomics
  // Tuned solution: bit v set when row v has arcs.  A vertex without arcs is
  // never queued (there is nothing to relax from it); on RMAT-style inputs
  // and trees most reached vertices are such.  nullptr: queue every claim.
  const uint32_t* __restrict__ hasout;
  unsigned long long* rlog;  // profiling: per round (queue length, arcs), or nullptr
  int32_t* progress;         // tuned: host memory, r + 1 when round r starts with work, -(r + 1) when it has none
};

// Profiling log: round r's queue length and arcs (one atomic per wave).
__device__ __forceinline__ void log_round(const FArgs& a, int r, int L, unsigned long long wave_arcs) {
  if (!a.rlog || r >= kRoundLogCap) return;
  if (blockIdx.x == 0 && threadIdx.x == 0 && L >= 0) a.rlog[2 * r] = (unsigned long long)L;
  if (lane_id() == 0 && wave_arcs) atomicAdd(&a.rlog[2 * r + 1], wave_arcs);
}

template <bool BFS>
__device__ __forceinline__ int arc_weight(const FArgs& a, int e) {
  return BFS ? 1 : stream_load(a.w + e);
}

// Relax one arc with candidate `alt`.  Returns true when this lane lowered
// dist[v] and is the first in this round to claim v for the next queue.
// relax_seen() takes a dist[v] read earlier (so several reads can be in
// flight together); relax() reads it itself.
// A plain read of dist[v] may be stale (larger) when another XCD lowered it;
// the atomic decides, so a stale read only costs an extra atomic.
// BFS: every lane of round r offers level r+1, so only the lane that finds v
// unvisited lowers it; that atomic alone makes the claim unique and the stamp
// is not needed.
// BFS on a forest (a.forest): v has a single in-arc, so exactly one lane ever
// offers v a level and a plain store is enough.
__device__ __forceinline__ bool has_arcs(const FArgs& a, int v) {
  return !a.hasout || ((a.hasout[v >> 5] >> (v & 31)) & 1u);
}

template <bool BFS>
__device__ __forceinline__ bool relax_seen(const FArgs& a, int v, int alt, int seen, int label, Sink& out) {
  if (alt >= seen) return false;
  ++out.tries;
  if (BFS && a.forest) {
    a.dist[v] = alt;
    ++out.wins;
    return has_arcs(a, v);
  }
  int old = atomicMin(&a.dist[v], alt);
  // read while the atomic is in flight: neither depends on its result, and
  // the stamp is only a filter (the exchange decides)
  const bool arcs = has_arcs(a, v);
  const int stamp = BFS ? 0 : a.stamp[v];
  if (alt >= old) return false;
  ++out.wins;
  if (!arcs) return false;
  if (BFS) return true;
  if (stamp == label) return false;
  return atomicExch(&a.stamp[v], label) != label;
}

template <bool BFS>
__device__ __forceinline__ bool relax(const FArgs& a, int v, int alt, int label, Sink& out) {
  return relax_seen<BFS>(a, v, alt, a.dist[v], label, out);
}

// K candidate relaxations per lane; want[t] says the plain read found v[t]
// improvable.  About 4% of SSSP's arcs get that far, yet with 64 lanes each
// of the K slots has a candidate in most steps, and a slot's atomics (min,
// stamp read, exchange) run in a branch that waits for each of them before
// the next slot: up to 3K dependent round trips per step.  The candidates
// are packed instead to the low lanes through the wave's LDS area (s_pair:
// 2 * K * 64 ints), so the wave relaxes them one per lane, normally in a
// single round, and the winners are queued from there.
template <bool BFS, int K>
__device__ __forceinline__ void relax_packed(const FArgs& a, const int* v, const int* alt, const bool* want,
                                             int label, int32_t* s_pair, Sink& out) {
  const int lane = lane_id();
  int total = 0;
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const uint64_t m = ballot(want[t]);
    if (want[t]) {
      const int i = total + mbcnt(m);
      s_pair[2 * i] = v[t];
      s_pair[2 * i + 1] = alt[t];
    }
    total += __popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int r0 = 0; r0 < total; r0 += SMDP_WAVE) {
    const int i = r0 + lane;
    int vv = 0;
    bool p = false;
    if (i < total) {
      vv = s_pair[2 * i];
      p = relax_seen<BFS>(a, vv, s_pair[2 * i + 1], INT32_MAX, label, out);
    }
    out.push(p, vv);
  }
  // s_pair is written again by the next step
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void clear_slot(Ctl* ctl, int r) { ctl_clear(ctl, (r + 2) & 3); }

// Wave-cooperative expansion of arcs [s, e) of one vertex with cost du.
template <bool BFS>
__device__ __forceinline__ void wave_expand(const FArgs& a, int s, int e, int du, int label,
                                            unsigned long long& edges, Sink& out) {
  const int lane = lane_id();
  for (int b = s; b < e; b += SMDP_WAVE) {
    int k = b + lane;
    int v = 0;
    bool p = false;
    if (k < e) {
      v = a.col[k];
      p = relax<BFS>(a, v, du + arc_weight<BFS>(a, k), label, out);
      ++edges;
    }
    out.push(p, v);
  }
}

// Workgroup-cooperative expansion (all SMDP_WG threads call it).
template <bool BFS>
__device__ __forceinline__ void wg_expand(const FArgs& a, int s, int e, int du, int label,
                                          unsigned long long& edges, Sink& out) {
  for (int b = s; b < e; b += kIlp * SMDP_WG) {
    int k[kIlp], v[kIlp], alt[kIlp];
#pragma unroll
    for (int t = 0; t < kIlp; ++t) k[t] = min(b + (int)threadIdx.x + t * SMDP_WG, e - 1);
#pragma unroll
    for (int t = 0; t < kIlp; ++t) v[t] = stream_load(a.col + k[t]);
#pragma unroll
    for (int t = 0; t < kIlp; ++t) alt[t] = du + arc_weight<BFS>(a, k[t]);
    int seen[kIlp];
#pragma unroll
    for (int t = 0; t < kIlp; ++t) seen[t] = a.dist[v[t]];
#pragma unroll
    for (int t = 0; t < kIlp; ++t) {
      const bool ok = b + (int)threadIdx.x + t * SMDP_WG < e;
      const bool p = ok && relax_seen<BFS>(a, v[t], alt[t], seen[t], label, out);
      edges += ok;
      out.push(p, v[t]);
    }
  }
}

__global__ void k_init_frontier(int32_t* dist, int32_t* stamp, int32_t n, int32_t inf, int32_t src,
                                int32_t* q0, Ctl* ctl, Counters* cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    dist[i] = (i == src) ? 0 : inf;
    stamp[i] = (i == src) ? 0 : -1;
  }
  if (blockIdx.x == 0) {
    int32_t* c = reinterpret_cast<int32_t*>(ctl);
    for (int i = threadIdx.x; i < 4 * kCtlInts; i += blockDim.x) c[i] = 0;
    unsigned long long* k = reinterpret_cast<unsigned long long*>(cnt);
    for (int i = threadIdx.x; i < (int)(kCntParts * sizeof(Counters) / 8); i += blockDim.x) k[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
      q0[0] = src;
      ctl[0].qlen = 1;
    }
  }
}

// One round with one lane per queued vertex (solutions FLAT..GRID).
template <int SOL, bool BFS>
__global__ __launch_bounds__(SMDP_WG) void k_expand_lane(FArgs a, int r) {
  clear_slot(a.ctl, r);
  Ctl* cin = a.ctl + (r & 3);
  Ctl* cout = a.ctl + ((r + 1) & 3);
  const int L = cin->qlen;
  if (L == 0) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&cnt_of(a.cnt)->rounds, (unsigned long long)(r + 1));
  const int32_t* qin = (r & 1) ? a.q1 : a.q0;
  int32_t* qout = (r & 1) ? a.q0 : a.q1;
  const int label = r + 1;
  const int lane = lane_id();
  const int wave = threadIdx.x / SMDP_WAVE;

  constexpr int kLdsQ = SOL == SMDP_SOL_BLOCK ? kBlockQ : (SOL == SMDP_SOL_WAVE ? SMDP_WAVES_PER_WG * kWaveQ : 1);
  __shared__ int32_t s_q[kLdsQ];
  __shared__ int32_t s_qn;
  __shared__ int32_t s_out[SMDP_WAVES_PER_WG * kOutCap];
  Sink out;
  out.init(s_out + wave * kOutCap, qout, &cout->qlen);
  int wq_n = 0;  // SOL_WAVE: entries in this wave's LDS queue (wave-uniform)
  if (SOL == SMDP_SOL_BLOCK) {
    if (threadIdx.x == 0) s_qn = 0;
    __syncthreads();
  }

  unsigned long long edges = 0, verts = 0, pw = 0, pb = 0, pg = 0;
  const int stride = gridDim.x * SMDP_WG;
  for (int base = blockIdx.x * SMDP_WG; base < L; base += stride) {
    const int i = base + (int)threadIdx.x;
    const bool valid = i < L;
    int u = 0, s = 0, e = 0, du = 0;
    if (valid) {
      u = qin[i];
      s = a.row[u];
      e = a.row[u + 1];
      du = BFS ? r : a.dist[u];  // BFS: queue r holds exactly the vertices of level r
      ++verts;
    }
    const bool heavy = SOL != SMDP_SOL_FLAT && valid && (e - s) >= a.heavy;

    if (SOL == SMDP_SOL_NAIVE) {
      // Analogue of the per-thread child launch: each heavy vertex is
      // expanded right away by the wave that found it.
      uint64_t hm = ballot(heavy);
      while (hm) {
        int l = first_lane(hm);
        hm &= hm - 1;
        ++pw;
        wave_expand<BFS>(a, __shfl(s, l), __shfl(e, l), __shfl(du, l), label, edges, out);
      }
    }

    {
      // Light vertices, one per lane.  The loop runs to the wave's largest
      // light degree so every lane reaches the wave-wide push together.
      const int deg = heavy ? 0 : e - s;
      const int iters = wave_max(deg);
      for (int j = 0; j < iters; ++j) {
        int v = 0;
        bool p = false;
        if (j < deg) {
          const int k = s + j;
          v = a.col[k];
          p = relax<BFS>(a, v, du + arc_weight<BFS>(a, k), label, out);
          ++edges;
        }
        out.push(p, v);
      }
    }

    if (SOL == SMDP_SOL_WAVE) {
      uint64_t hm = ballot(heavy);
      if (heavy) s_q[wave * kWaveQ + wq_n + mbcnt(hm)] = u;
      // Lanes read entries other lanes wrote: order the LDS traffic.
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      wq_n += __popcll(hm);
      pw += (lane == 0) ? __popcll(hm) : 0;
      // Drain when another iteration could overflow, and at the end.
      bool last = base + stride >= L;
      if (wq_n > kWaveQ - SMDP_WAVE || (last && wq_n > 0)) {
        for (int j = 0; j < wq_n; ++j) {
          int hu = s_q[wave * kWaveQ + j];
          wave_expand<BFS>(a, a.row[hu], a.row[hu + 1], a.dist[hu], label, edges, out);
        }
        wq_n = 0;
      }
    }

    if (SOL == SMDP_SOL_BLOCK) {
      uint64_t hm = ballot(heavy);
      if (hm) {
        int b0 = 0;
        if (lane == first_lane(hm)) b0 = atomicAdd(&s_qn, __popcll(hm));
        b0 = __shfl(b0, first_lane(hm));
        if (heavy) s_q[b0 + mbcnt(hm)] = u;
        pb += (lane == first_lane(hm)) ? __popcll(hm) : 0;
      }
      __syncthreads();
      const int qn = s_qn;
      // Every wave must read the count before any wave appends again.
      __syncthreads();
      const bool last = base + stride >= L;
      if (qn > kBlockQ - SMDP_WG || (last && qn > 0)) {
        for (int j = 0; j < qn; ++j) {
          int hu = s_q[j];
          wg_expand<BFS>(a, a.row[hu], a.row[hu + 1], a.dist[hu], label, edges, out);
        }
        __syncthreads();
        if (threadIdx.x == 0) s_qn = 0;
        __syncthreads();
      }
    }

    if (SOL == SMDP_SOL_GRID) {
      uint64_t hm = ballot(heavy);
      if (hm) {
        int leader = first_lane(hm);
        int b0 = 0;
        if (lane == leader) b0 = atomicAdd(&cin->hlen, __popcll(hm));
        b0 = __shfl(b0, leader);
        if (heavy) a.hq[b0 + mbcnt(hm)] = Chunk{u, s};
        pg += (lane == leader) ? __popcll(hm) : 0;
      }
    }
  }
  log_round(a, r, L, wave_sum(edges));
  flush_counters(a.cnt, edges, verts, out.pushes, pw, pb, pg, out);
}

// One round of the tuned solution: the workgroup takes tiles of SMDP_WG queue
// entries from a shared counter, stages (first arc, degree, cost) in LDS,
// scans the degrees and lets every thread take consecutive arcs of the tile,
// so loads of col/w are coalesced and lanes stay balanced whatever the degree
// mix.  Vertices of degree >= split are cut into chunks on the grid queue.
template <bool BFS>
__global__ __launch_bounds__(SMDP_WG) void k_expand_tile(FArgs a, int r) {
  clear_slot(a.ctl, r);
  Ctl* cin = a.ctl + (r & 3);
  Ctl* cout = a.ctl + ((r + 1) & 3);
  const int L = cin->qlen;
  if (blockIdx.x == 0 && threadIdx.x == 0 && a.progress)
    __hip_atomic_store(a.progress, L > 0 ? r + 1 : -(r + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  if (L == 0) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&cnt_of(a.cnt)->rounds, (unsigned long long)(r + 1));
  // workgroups past the tile count would only contend for the counter
  if ((int)blockIdx.x * SMDP_WG * kTileGrab >= L) return;
  const int32_t* qin = (r & 1) ? a.q1 : a.q0;
  int32_t* qout = (r & 1) ? a.q0 : a.q1;
  const int label = r + 1;
  const int lane = lane_id();

  __shared__ int s_tile, s_tend;
  __shared__ int s_start[SMDP_WG];
  __shared__ int s_du[SMDP_WG];
  __shared__ int s_off[SMDP_WG];
  __shared__ int s_scan[SMDP_WAVES_PER_WG];
  __shared__ int32_t s_out[SMDP_WAVES_PER_WG * kOutCap];
  __shared__ int32_t s_pair[SMDP_WAVES_PER_WG * 2 * kIlp * SMDP_WAVE];  // relax_packed
  Sink out;
  out.init(s_out + (threadIdx.x / SMDP_WAVE) * kOutCap, qout, &cout->qlen);

  unsigned long long edges = 0, verts = 0, pg = 0;
  const int ntiles = (L + SMDP_WG - 1) / SMDP_WG;
  int part = 0;  // thread 0: claim cursor
  for (;;) {
    if (threadIdx.x == 0) {
      int f = 0, e = 0;
      s_tile = claim_items(cin->tiles, ntiles, kTileGrab, part, &f, &e) ? f : -1;
      s_tend = e;
    }
    __syncthreads();
    const int first = s_tile, tend = s_tend;
    __syncthreads();
    if (first < 0) break;
    for (int tile = first; tile < tend; ++tile) {
    const int base = tile * SMDP_WG;
    const int i = base + (int)threadIdx.x;
    const bool valid = i < L;
    int u = 0, s = 0, deg = 0, du = 0;
    if (valid) {
      u = qin[i];
      s = a.row[u];
      deg = a.row[u + 1] - s;
      du = BFS ? r : a.dist[u];  // BFS: queue r holds exactly the vertices of level r
      ++verts;
    }
    // A short queue has too few tiles to keep every workgroup busy, and its
    // vertices tend to be hubs: most of their arcs then go to the grid queue,
    // which spreads them over all workgroups.
    const bool split = deg >= (L < a.small_q ? min(a.heavy, kSmallSplit) : a.heavy);
    const int nch = split ? (deg + a.chunk - 1) / a.chunk : 0;
    const int incl = wave_inclusive_scan(nch);
    const int wtot = __shfl(incl, SMDP_WAVE - 1);
    if (wtot > 0) {
      int b0 = 0;
      if (lane == 0) b0 = atomicAdd(&cin->hlen, wtot);
      b0 = __shfl(b0, 0);
      int off = b0 + incl - nch;
      for (int c = 0; c < nch; ++c) a.hq[off + c] = Chunk{u, s + c * a.chunk};
      pg += nch;
    }
    s_start[threadIdx.x] = s;
    s_du[threadIdx.x] = du;
    int tot;
    const int ex = wg_exclusive_scan(split ? 0 : deg, s_scan, &tot);
    s_off[threadIdx.x] = ex;
    __syncthreads();
    // Each thread takes kIlp arcs per step (slots j, j+WG, ...): owners are
    // found together and all loads are issued before any atomic, so every
    // thread keeps kIlp gathers in flight.
    for (int jb = 0; jb < tot; jb += kIlp * SMDP_WG) {
      int j[kIlp], own[kIlp], k[kIlp], v[kIlp], alt[kIlp];
#pragma unroll
      for (int t = 0; t < kIlp; ++t) j[t] = min(jb + (int)threadIdx.x + t * SMDP_WG, tot - 1);
      tile_owners<kIlp>(s_off, j, own);
#pragma unroll
      for (int t = 0; t < kIlp; ++t) k[t] = s_start[own[t]] + (j[t] - s_off[own[t]]);
#pragma unroll
      for (int t = 0; t < kIlp; ++t) v[t] = stream_load(a.col + k[t]);
#pragma unroll
      for (int t = 0; t < kIlp; ++t) alt[t] = s_du[own[t]] + arc_weight<BFS>(a, k[t]);
      int seen[kIlp];
#pragma unroll
      for (int t = 0; t < kIlp; ++t) seen[t] = a.dist[v[t]];
      bool want[kIlp];
#pragma unroll
      for (int t = 0; t < kIlp; ++t) {
        const bool ok = jb + (int)threadIdx.x + t * SMDP_WG < tot;
        want[t] = ok && alt[t] < seen[t];
        edges += ok;
      }
      relax_packed<BFS, kIlp>(a, v, alt, want, label, s_pair + (threadIdx.x / SMDP_WAVE) * 2 * kIlp * SMDP_WAVE,
                              out);
    }
    __syncthreads();
    }
  }
  log_round(a, r, L, wave_sum(edges));
  flush_counters(a.cnt, edges, verts, out.pushes, 0, 0, pg, out);
}

// Phase 2 of the grid solutions: workgroups claim tiles of kDrainTile
// grid-queue entries and expand them edge-balanced, like k_expand_tile: wave
// 0 scans the entries' lengths and every thread takes kIlp arcs per step.  A
// short round's entries are whole vertices of a few dozen arcs, so one
// workgroup pass then covers several of them instead of one each.
template <bool BFS>
__global__ __launch_bounds__(SMDP_WG) void k_drain(FArgs a, int r) {
  Ctl* cin = a.ctl + (r & 3);
  Ctl* cout = a.ctl + ((r + 1) & 3);
  const int H = cin->hlen;
  // entries per tile: kDrainTile, halved while the tiles still number at most
  // one per workgroup (a few hub chunks then go to as many workgroups)
  int ept = kDrainTile;
  while (ept > 1 && (H + ept / 2 - 1) / (ept / 2) <= (int)gridDim.x) ept >>= 1;
  const int ntiles = (H + ept - 1) / ept;
  if ((int)blockIdx.x >= ntiles) return;
  int32_t* qout = (r & 1) ? a.q0 : a.q1;
  const int label = r + 1;
  __shared__ int s_tile, s_tot;
  __shared__ int s_start[kDrainTile];
  __shared__ int s_du[kDrainTile];
  __shared__ int s_off[kDrainTile];
  __shared__ int32_t s_out[SMDP_WAVES_PER_WG * kOutCap];
  __shared__ int32_t s_pair[SMDP_WAVES_PER_WG * 2 * kIlp * SMDP_WAVE];  // relax_packed
  Sink out;
  out.init(s_out + (threadIdx.x / SMDP_WAVE) * kOutCap, qout, &cout->qlen);
  unsigned long long edges = 0;
  int part = 0;  // thread 0: claim cursor
  for (;;) {
    if (threadIdx.x == 0) {
      int f = 0, e = 0;
      s_tile = claim_items(cin->chunks, ntiles, 1, part, &f, &e) ? f : -1;
    }
    __syncthreads();
    const int tile = s_tile;
    if (tile < 0) break;
    if (threadIdx.x < SMDP_WAVE) {
      const int idx = tile * ept + (int)threadIdx.x;
      int len = 0, st = 0, du = 0;
      if ((int)threadIdx.x < ept && idx < H) {
        const Chunk c = a.hq[idx];
        const int vend = a.row[c.v + 1];
        const int end = (a.chunk > 0 && vend - c.first > a.chunk) ? c.first + a.chunk : vend;
        len = end - c.first;
        st = c.first;
        du = BFS ? r : a.dist[c.v];
      }
      const int incl = wave_inclusive_scan(len);
      if ((int)threadIdx.x < kDrainTile) {
        s_off[threadIdx.x] = incl - len;
        s_start[threadIdx.x] = st;
        s_du[threadIdx.x] = du;
      }
      if (threadIdx.x == kDrainTile - 1) s_tot = incl;
    }
    __syncthreads();
    const int tot = s_tot;
    for (int jb = 0; jb < tot; jb += kIlp * SMDP_WG) {
      int j[kIlp], own[kIlp], k[kIlp], v[kIlp], alt[kIlp];
#pragma unroll
      for (int t = 0; t < kIlp; ++t) j[t] = min(jb + (int)threadIdx.x + t * SMDP_WG, tot - 1);
      tile_owners<kIlp, kDrainTile>(s_off, j, own);
#pragma unroll
      for (int t = 0; t < kIlp; ++t) k[t] = s_start[own[t]] + (j[t] - s_off[own[t]]);
#pragma unroll
      for (int t = 0; t < kIlp; ++t) v[t] = stream_load(a.col + k[t]);
#pragma unroll
      for (int t = 0; t < kIlp; ++t) alt[t] = s_du[own[t]] + arc_weight<BFS>(a, k[t]);
      int seen[kIlp];
#pragma unroll
      for (int t = 0; t < kIlp; ++t) seen[t] = a.dist[v[t]];
      bool want[kIlp];
#pragma unroll
      for (int t = 0; t < kIlp; ++t) {
        const bool ok = jb + (int)threadIdx.x + t * SMDP_WG < tot;
        want[t] = ok && alt[t] < seen[t];
        edges += ok;
      }
      relax_packed<BFS, kIlp>(a, v, alt, want, label, s_pair + (threadIdx.x / SMDP_WAVE) * 2 * kIlp * SMDP_WAVE,
                              out);
    }
    __syncthreads();
  }
  log_round(a, r, -1, wave_sum(edges));
  flush_counters(a.cnt, edges, 0, out.pushes, 0, 0, 0, out);
  unsigned long long e2 = wave_sum(edges);
  if (lane_id() == 0 && e2) atomicAdd(&cnt_of(a.cnt)->edges2, e2);
}

// Out-arcs of reached vertices (the GTEPS numerator).
__global__ void k_reached(const int32_t* __restrict__ row, const int32_t* __restrict__ dist, int32_t n,
                          int32_t inf, Counters* cnt) {
  // 8 vertices per thread with their loads issued at once
  constexpr int K = 8;
  unsigned long long acc = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += K * stride) {
    int d[K], r0[K], r1[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t i = min(i0 + k * stride, (int64_t)n - 1);
      d[k] = dist[i];
      r0[k] = row[i];
      r1[k] = row[i + 1];
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (i0 + k * stride < n && d[k] != inf) acc += (unsigned long long)(r1[k] - r0[k]);
  }
  acc = wave_sum(acc);
  if (lane_id() == 0 && acc) atomicAdd(&cnt_of(cnt)->reached, acc);
}

template <bool BFS>
static void launch_round(int sol, int grid, hipStream_t st, const FArgs& a, int r, int* launches) {
  dim3 g(grid), b(SMDP_WG);
  switch (sol) {
    case SMDP_SOL_FLAT: hipLaunchKernelGGL((k_expand_lane<SMDP_SOL_FLAT, BFS>), g, b, 0, st, a, r); break;
    case SMDP_SOL_NAIVE:
    case SMDP_SOL_HIER: hipLaunchKernelGGL((k_expand_lane<SMDP_SOL_NAIVE, BFS>), g, b, 0, st, a, r); break;
    case SMDP_SOL_WAVE: hipLaunchKernelGGL((k_expand_lane<SMDP_SOL_WAVE, BFS>), g, b, 0, st, a, r); break;
    case SMDP_SOL_BLOCK: hipLaunchKernelGGL((k_expand_lane<SMDP_SOL_BLOCK, BFS>), g, b, 0, st, a, r); break;
    case SMDP_SOL_GRID: hipLaunchKernelGGL((k_expand_lane<SMDP_SOL_GRID, BFS>), g, b, 0, st, a, r); break;
    default: hipLaunchKernelGGL((k_expand_tile<BFS>), g, b, 0, st, a, r); break;
  }
  ++*launches;
}

template <bool BFS>
static void launch_drain(int grid, hipStream_t st, const FArgs& a, int r, int* launches) {
  hipLaunchKernelGGL((k_drain<BFS>), dim3(grid), dim3(SMDP_WG), 0, st, a, r);
  ++*launches;
}

static int ensure_frontier(smdp_graph* g, int32_t chunk) {
  SMDP_CHECK(alloc_once(&g->dist, g->n));
  SMDP_CHECK(alloc_once(&g->stamp, g->n));
  SMDP_CHECK(alloc_once(&g->q[0], g->n));
  SMDP_CHECK(alloc_once(&g->q[1], g->n));
  // Each vertex enters a round's queue at most once (stamp), so a round pushes
  // at most n whole-vertex entries or n + m/chunk chunk entries.
  int64_t need = (int64_t)g->n + (chunk > 0 ? (int64_t)g->m / chunk : 0) + 64;
  return ensure(&g->hq, &g->hq_cap, need);
}

// Copies the per-round log and pairs it with the recorded launch events
// (phase 1 and, for two-phase solutions, phase 2 of each round).
static int collect_round_log(smdp_ctx* ctx, int rounds, bool two_phase) {
  const int k = std::min(rounds, kRoundLogCap);
  std::vector<unsigned long long> h(2 * (size_t)k);
  if (k) SMDP_HIP(hipMemcpy(h.data(), ctx->rlog, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
  ctx->log_qlen.assign(k, 0);
  ctx->log_arcs.assign(k, 0);
  ctx->log_ms.assign(k, 0.0);
  const size_t per_round = two_phase ? 4 : 2;  // events per round
  for (int i = 0; i < k; ++i) {
    ctx->log_qlen[i] = (int64_t)h[2 * i];
    ctx->log_arcs[i] = (int64_t)h[2 * i + 1];
    for (size_t e = 0; e + 1 < per_round; e += 2) {
      const size_t at = i * per_round + e;
      if (at + 1 >= ctx->prof_ev.size()) break;
      float t = 0;
      SMDP_HIP(hipEventElapsedTime(&t, ctx->prof_ev[at], ctx->prof_ev[at + 1]));
      ctx->log_ms[i] += t;
    }
  }
  return SMDP_OK;
}

// ---- renumbered copy for the tuned solution (SMDP_LAYOUT_HOT) --------------
// Vertices are renumbered by falling in-degree (ties by id), rows keep their
// arcs in order.  The distances that most arcs gather then share the first
// few MB of dist and stay in each XCD's L2: RMAT-22 SSSP went from 4.55 to
// 4.09 ms (median of 11).  The solve runs on the copy and results are
// gathered back to the caller's numbering when read.


__global__ void k_unpermute(const int32_t* __restrict__ d, const int32_t* __restrict__ newid, int64_t count,
                            int32_t* out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < count; v += (int64_t)gridDim.x * blockDim.x)
    out[v] = d[newid[v]];
}

void relabel_free(smdp_graph* g) {
  Relabel* rl = static_cast<Relabel*>(g->relabel);
  if (!rl) return;
  void* ptrs[] = {rl->row, rl->col, rl->w, rl->newid, rl->hasout, rl->rec, rl->cw, rl->phase,
                  rl->rrow, rl->rcw, rl->hasin};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (rl->hctl) (void)hipHostFree(rl->hctl);
  delete rl;
  g->relabel = nullptr;
}

int relabel_unpermute(smdp_graph* g, int32_t* out_dev, int64_t count) {
  Relabel* rl = static_cast<Relabel*>(g->relabel);
  if (!rl || !out_dev || count > g->n) return SMDP_ERR_ARG;
  if (count > 0)
    hipLaunchKernelGGL(k_unpermute, dim3((unsigned)std::min<int64_t>((count + 255) / 256, 4096)), dim3(256), 0,
                       g->ctx->stream, g->dist, rl->newid, count, out_dev);
  SMDP_HIP(hipGetLastError());
  return SMDP_OK;
}

static int ensure_relabel(smdp_graph* g, Relabel** out) {
  if (g->relabel) {
    *out = static_cast<Relabel*>(g->relabel);
    return SMDP_OK;
  }
  const int64_t n = g->n, m = g->m;
  std::vector<int32_t> row((size_t)n + 1), col((size_t)m), w(g->w ? (size_t)m : 0);
  SMDP_HIP(hipMemcpy(row.data(), g->row, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
  if (m) SMDP_HIP(hipMemcpy(col.data(), g->col, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  if (m && g->w) SMDP_HIP(hipMemcpy(w.data(), g->w, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  std::vector<int32_t> indeg((size_t)n, 0);
  for (int64_t e = 0; e < m; ++e) ++indeg[(size_t)col[(size_t)e]];
  int32_t minw = 0;
  if (g->w && m) minw = *std::min_element(w.begin(), w.end());
  int32_t maxd = 0;
  for (int32_t d : indeg) maxd = std::max(maxd, d);
  // counting sort by falling in-degree, stable in the vertex id
  std::vector<int64_t> start((size_t)maxd + 2, 0);
  for (int32_t d : indeg) ++start[(size_t)(maxd - d) + 1];
  for (size_t k = 1; k < start.size(); ++k) start[k] += start[k - 1];
  std::vector<int32_t> order((size_t)n);
  std::vector<int32_t> newid((size_t)n);
  for (int64_t v = 0; v < n; ++v) order[(size_t)start[(size_t)(maxd - indeg[(size_t)v])]++] = (int32_t)v;
  for (int64_t i = 0; i < n; ++i) newid[(size_t)order[(size_t)i]] = (int32_t)i;
  std::vector<int32_t> row2((size_t)n + 1);
  row2[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t u = order[(size_t)i];
    row2[(size_t)i + 1] = row2[(size_t)i] + (row[(size_t)u + 1] - row[(size_t)u]);
  }
  std::vector<int32_t> col2((size_t)m), w2(g->w ? (size_t)m : 0);
  parallel_ranges(n, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) {
      const int32_t u = order[(size_t)i];
      int64_t o = row2[(size_t)i];
      for (int64_t e = row[(size_t)u]; e < row[(size_t)u + 1]; ++e, ++o) {
        col2[(size_t)o] = newid[(size_t)col[(size_t)e]];
        if (g->w) w2[(size_t)o] = w[(size_t)e];
      }
    }
  });
  // attached to the graph only once complete (relabel_free releases a partial one)
  g->relabel = new Relabel();
  Relabel* rl = static_cast<Relabel*>(g->relabel);
  rl->h_newid = std::move(newid);
  rl->minw = minw;
  // in-row offsets of the copy, for the bucketed solve's pull settle
  rl->h_rrow.assign((size_t)n + 1, 0);
  for (int64_t i = 0; i < n; ++i)
    rl->h_rrow[(size_t)i + 1] = rl->h_rrow[(size_t)i] + indeg[(size_t)order[(size_t)i]];
  int rc = alloc_once(&rl->row, n + 1);
  if (rc == SMDP_OK) rc = alloc_once(&rl->col, std::max<int64_t>(1, m));
  if (rc == SMDP_OK) rc = alloc_once(&rl->newid, std::max<int64_t>(1, n));
  if (rc == SMDP_OK && g->w) rc = alloc_once(&rl->w, std::max<int64_t>(1, m));
  if (rc == SMDP_OK &&
      (hipMemcpy(rl->row, row2.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
       (m && hipMemcpy(rl->col, col2.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice) != hipSuccess) ||
       (m && g->w && hipMemcpy(rl->w, w2.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice) != hipSuccess) ||
       hipMemcpy(rl->newid, rl->h_newid.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess))
    rc = SMDP_ERR_DEVICE;
  if (rc != SMDP_OK) {
    relabel_free(g);
    return rc;
  }
  *out = rl;
  return SMDP_OK;
}

// Bit v of words[v / 32] set when row v has arcs.
__global__ void k_hasout(const int32_t* __restrict__ row, int32_t n, uint32_t* words) {
  const int32_t nw = (n + 31) / 32;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += gridDim.x * blockDim.x) {
    uint32_t b = 0;
    for (int k = 0; k < 32; ++k) {
      const int32_t v = i * 32 + k;
      if (v < n && row[v + 1] > row[v]) b |= 1u << k;
    }
    words[i] = b;
  }
}

int build_hasout(smdp_graph* g, const int32_t* row, uint32_t** out) {
  if (*out) return SMDP_OK;
  const int64_t nw = std::max<int64_t>(1, ((int64_t)g->n + 31) / 32);
  SMDP_CHECK(alloc_once(out, nw));
  hipLaunchKernelGGL(k_hasout, dim3((unsigned)std::min<int64_t>((nw + 255) / 256, 4096)), dim3(256), 0,
                     g->ctx->stream, row, g->n, *out);
  SMDP_HIP(hipGetLastError());
  return SMDP_OK;
}

// Whether every vertex has at most one in-arc, checked once per resident graph.
__global__ void k_indeg_at_most_one(const int32_t* __restrict__ col, int64_t m, int32_t* cnt, int32_t* multi) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    if (atomicAdd(&cnt[col[e]], 1) >= 1) *multi = 1;
}

int ensure_forest(smdp_graph* g) {
  if (g->forest >= 0) return SMDP_OK;
  int32_t* cnt = nullptr;
  SMDP_CHECK(alloc_once(&cnt, (int64_t)g->n + 1));
  hipStream_t s = g->ctx->stream;
  int rc = SMDP_OK;
  if (hipMemsetAsync(cnt, 0, sizeof(int32_t) * ((size_t)g->n + 1), s) != hipSuccess) rc = SMDP_ERR_DEVICE;
  if (rc == SMDP_OK && g->m > 0) {
    hipLaunchKernelGGL(k_indeg_at_most_one, dim3(std::max(1, g->ctx->cus * 8)), dim3(256), 0, s, g->col,
                       (int64_t)g->m, cnt, cnt + g->n);
    if (hipGetLastError() != hipSuccess) rc = SMDP_ERR_DEVICE;
  }
  int32_t multi = 1;
  if (rc == SMDP_OK && (hipMemcpyAsync(&multi, cnt + g->n, sizeof multi, hipMemcpyDeviceToHost, s) != hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess))
    rc = SMDP_ERR_DEVICE;
  (void)hipFree(cnt);
  if (rc == SMDP_OK) g->forest = multi ? 0 : 1;
  return rc;
}

// The renumbered copy is built on the host once per resident graph; a graph
// uploaded by a drop-in call for one solve never pays for it (RMAT-22: a
// drop-in solve took 21 ms without it and 575 ms with it), so it runs the
// unordered rounds on the graph as given.
static bool use_relabel(const smdp_params& p, const smdp_graph* g) {
  if (g->transient) return false;
  return p.layout == SMDP_LAYOUT_HOT || p.layout == SMDP_LAYOUT_BUCKETS ||
         (p.layout == SMDP_LAYOUT_AUTO && !g->transient && g->m >= (1 << 20));
}

int run_delta(smdp_graph* g, Relabel* rl, int32_t src, const smdp_params& p, int grid, float* ms_out,
              int* launches_out, bool* reached_counted);

// Common end of a frontier run: out-arcs of reached vertices, the result kind,
// the per-round log (log_rounds >= 0: rounds logged by the unordered kernels)
// and the statistics.
static int finish_frontier(smdp_graph* g, const int32_t* row, int32_t inf, bool bfs, bool relabel, float ms,
                           int launches, int log_rounds, bool two_phase, smdp_stats* st, bool reached_counted = false) {
  smdp_ctx* ctx = g->ctx;
  if (!reached_counted) {
    ctx->counters_staged = false;  // k_reached adds to the counters
    hipLaunchKernelGGL(k_reached, dim3(std::min<int64_t>((g->n + 255) / 256, 2048)), dim3(256), 0, ctx->stream, row,
                       g->dist, g->n, inf, g->cnt);
    SMDP_HIP(hipGetLastError());
  }
  if (log_rounds >= 0) SMDP_CHECK(collect_round_log(ctx, log_rounds, two_phase));
  if (st) {
    SMDP_CHECK(fill_stats(g, st, ms, launches));
    SMDP_CHECK(prof_collect(ctx, st));
  } else {
    prof_reset(ctx);
  }
  g->last_res = bfs ? SMDP_RES_LEVEL : SMDP_RES_DIST;
  g->dist_perm = relabel;
  return SMDP_OK;
}

// Runs SSSP (bfs = false) or BFS from src with the given solution; results stay
// in g->dist.
int run_frontier(smdp_graph* g, int32_t src, int sol, bool bfs, smdp_stats* st) {
  if (!g) return SMDP_ERR_ARG;
  if (g->n <= 0 || src < 0 || src >= g->n) return SMDP_ERR_ARG;
  if (sol < 0 || sol >= SMDP_SOL_COUNT) return SMDP_ERR_SOLUTION;
  if (sol == SMDP_SOL_HIER && !bfs) return SMDP_ERR_SOLUTION;
  if (!bfs && g->m > 0 && !g->w) return SMDP_ERR_ARG;
  smdp_ctx* ctx = g->ctx;
  SMDP_HIP(hipSetDevice(ctx->device));
  begin_run(g);
  const smdp_params& p = params_of(ctx, bfs ? SMDP_APP_BFS : SMDP_APP_SSSP);
  const bool tuned = sol == SMDP_SOL_TUNED;
  SMDP_CHECK(ensure_frontier(g, tuned ? p.chunk : 0));
  const bool relabel = tuned && use_relabel(p, g);
  const int32_t* row = g->row;
  const int32_t* col = g->col;
  const int32_t* w = g->w;
  if (relabel) {
    Relabel* rl = nullptr;
    SMDP_CHECK(ensure_relabel(g, &rl));
    row = rl->row;
    col = rl->col;
    w = rl->w;
    src = rl->h_newid[(size_t)src];
  }

  FArgs a;
  a.row = row;
  a.col = col;
  a.w = w;
  a.dist = g->dist;
  a.stamp = g->stamp;
  a.q0 = g->q[0];
  a.q1 = g->q[1];
  a.hq = g->hq;
  a.ctl = g->ctl;
  a.cnt = g->cnt;
  a.heavy = tuned ? p.split_threshold : (sol == SMDP_SOL_HIER ? 1 : p.heavy_threshold);
  a.chunk = tuned ? p.chunk : 0;
  a.rlog = ctx->profiling ? ctx->rlog : nullptr;
  if (a.rlog) SMDP_HIP(hipMemsetAsync(a.rlog, 0, sizeof(unsigned long long) * 2 * kRoundLogCap, ctx->stream));
  a.progress = tuned ? ctx->progress_dev : nullptr;
  const int grid = grid_size(ctx, p);
  a.small_q = grid * SMDP_WG;
  a.hasout = nullptr;
  if (tuned) {
    uint32_t** slot = relabel ? &static_cast<Relabel*>(g->relabel)->hasout : &g->hasout;
    SMDP_CHECK(build_hasout(g, row, slot));
    a.hasout = *slot;
  }
  a.forest = 0;
  if (tuned && bfs && !g->transient) {
    SMDP_CHECK(ensure_forest(g));
    a.forest = g->forest;
  }
  const int32_t inf = bfs ? SMDP_BFS_UNDEF : SMDP_SSSP_INF;
  const bool two_phase = sol == SMDP_SOL_GRID || tuned;
  hipStream_t s = ctx->stream;
  // SSSP on the renumbered copy runs in distance buckets unless asked for
  // unordered rounds or a weight is negative (buckets need w >= 0)
  const bool buckets = relabel && !bfs && p.layout != SMDP_LAYOUT_HOT &&
                       static_cast<Relabel*>(g->relabel)->minw >= 0;
  if (buckets) {
    float ms = 0;
    int launches = 0;
    bool counted = false;
    SMDP_CHECK(run_delta(g, static_cast<Relabel*>(g->relabel), src, p, grid, &ms, &launches, &counted));
    if (!st) ctx->counters_staged = false;
    return finish_frontier(g, row, inf, false, relabel, ms, launches, -1, false, st, counted);
  }
  prof_reset(ctx);

  int launches = 0;
  SMDP_HIP(hipEventRecord(ctx->ev_begin, s));
  hipLaunchKernelGGL(k_init_frontier, dim3(std::min<int64_t>((g->n + 255) / 256, 4096)), dim3(256), 0, s,
                     g->dist, g->stamp, g->n, inf, src, g->q[0], g->ctl, g->cnt);
  ++launches;
  int r = 0;
  auto enqueue = [&](int k) -> int {
    for (int j = 0; j < k; ++j, ++r) {
      SMDP_CHECK(prof_mark(ctx, 0));
      if (bfs) launch_round<true>(sol, grid, s, a, r, &launches);
      else launch_round<false>(sol, grid, s, a, r, &launches);
      SMDP_CHECK(prof_mark(ctx, 0));
      if (two_phase) {
        SMDP_CHECK(prof_mark(ctx, 1));
        if (bfs) launch_drain<true>(grid, s, a, r, &launches);
        else launch_drain<false>(grid, s, a, r, &launches);
        SMDP_CHECK(prof_mark(ctx, 1));
      }
    }
    SMDP_HIP(hipGetLastError());
    return SMDP_OK;
  };
  if (tuned) {
    // The tile kernel reports each round it starts in host memory, so the
    // host keeps a few rounds queued ahead of the device and stops at the
    // round that finds the queue empty (rounds enqueued after it return at
    // once).  A blocking check per batch of rounds left the device idle for
    // the round trip each time.
    volatile int32_t* prog = ctx->progress;
    *prog = 0;
    SMDP_CHECK(enqueue(std::min(kFeedFirst, p.max_rounds)));
    int seen = 0;
    double t_seen = now_ms();
    for (;;) {
      const int v = *prog;
      if (v < 0) break;
      if (v >= r - kFeedSlack) {
        if (r >= p.max_rounds) {
          // the cap is enqueued: done only if its last round found no work
          SMDP_HIP(hipStreamSynchronize(s));
          if (*prog < 0) break;
          return SMDP_ERR_OVERFLOW;
        }
        SMDP_CHECK(enqueue(std::min(kFeed, p.max_rounds - r)));
        continue;
      }
      if (v != seen) {
        seen = v;
        t_seen = now_ms();
      } else if (now_ms() - t_seen > 100.0) {
        // no progress report for a while: a fault stops the stream
        const hipError_t e = hipStreamQuery(s);
        if (e != hipSuccess && e != hipErrorNotReady) return hip_fail(e, "tuned frontier rounds", __LINE__);
        t_seen = now_ms();
      }
    }
  } else {
    RoundBatcher rb(p.batch_rounds, 16);
    for (;;) {
      SMDP_CHECK(enqueue(rb.batch));
      SMDP_HIP(hipMemcpyAsync(ctx->pinned, &g->ctl[r & 3].qlen, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      SMDP_HIP(hipStreamSynchronize(s));
      const int32_t q = ctx->pinned[0];
      if (q == 0) break;
      if (r >= p.max_rounds) return SMDP_ERR_OVERFLOW;
      rb.advance(q, r);
    }
  }
  SMDP_HIP(hipEventRecord(ctx->ev_end, s));
  SMDP_HIP(hipEventSynchronize(ctx->ev_end));
  float ms = 0;
  SMDP_HIP(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
  return finish_frontier(g, row, inf, bfs, relabel, ms, launches, a.rlog ? r : -1, two_phase, st);
}

// ---- SSSP on a caller's device CSR with no allocation -------------------------
// Every working buffer is carved from the library's fixed per-GPU array
// (scratch_ptr), 256-byte aligned, in this order.
namespace noalloc {
constexpr int64_t kMaxN = SMDP_NOALLOC_MAX_N;
constexpr int64_t kMaxM = SMDP_NOALLOC_MAX_M;
constexpr int64_t kHqCap = kMaxN + kMaxM / 64 + 64;  // chunk >= 64 (smdp_set_params)
constexpr int64_t align(int64_t b) { return (b + 255) / 256 * 256; }
constexpr int64_t kCtl = 0;
constexpr int64_t kCnt = kCtl + align(4 * (int64_t)sizeof(Ctl));
constexpr int64_t kFlag = kCnt + align(kCntParts * (int64_t)sizeof(Counters));
constexpr int64_t kStamp = kFlag + 256;
constexpr int64_t kQ0 = kStamp + align(4 * kMaxN);
constexpr int64_t kQ1 = kQ0 + align(4 * kMaxN);
constexpr int64_t kHasout = kQ1 + align(4 * kMaxN);
constexpr int64_t kHq = kHasout + align(4 * ((kMaxN + 31) / 32));
constexpr int64_t kEnd = kHq + align((int64_t)sizeof(Chunk) * kHqCap);
static_assert(kEnd <= kNoAllocScratchBytes, "the working buffers fit the fixed array");
}  // namespace noalloc

// flag = 1 when row is not non-decreasing from row[0] = 0 to row[n] = m or a
// column is outside [0, n).  Reads row[0..n] and col[0..m) only.
__global__ void k_check_csr(const int32_t* __restrict__ row, const int32_t* __restrict__ col, int32_t n, int32_t m,
                            int32_t* flag) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = row[i];
    if ((i == 0 && r != 0) || (i == n && r != m) || (i < n && row[i + 1] < r)) bad = true;
  }
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    if ((uint32_t)col[e] >= (uint32_t)n) bad = true;
  if (bad) atomicOr(flag, 1);
}

__global__ void k_clear_flag(int32_t* flag) { *flag = 0; }

int sssp_noalloc(smdp_ctx* ctx, int32_t n, int32_t m, const int32_t* d_row, const int32_t* d_col,
                 const int32_t* d_w, int32_t src, int sol, int32_t* d_dist, smdp_stats* st) {
  if (!ctx || !d_row || !d_dist || n <= 0 || m < 0 || (m > 0 && (!d_col || !d_w))) return SMDP_ERR_ARG;
  if (n > noalloc::kMaxN || m > noalloc::kMaxM) return SMDP_ERR_ARG;  // beyond the fixed array
  if (src < 0 || src >= n) return SMDP_ERR_ARG;
  if (sol < 0 || sol >= SMDP_SOL_COUNT || sol == SMDP_SOL_HIER) return SMDP_ERR_SOLUTION;
  SMDP_HIP(hipSetDevice(ctx->device));
  char* base = static_cast<char*>(scratch_ptr(ctx->device));
  if (!base) return SMDP_ERR_DEVICE;
  hipStream_t s = ctx->stream;
  int32_t* flag = reinterpret_cast<int32_t*>(base + noalloc::kFlag);
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, flag);
  hipLaunchKernelGGL(k_check_csr, dim3((unsigned)std::min<int64_t>(((int64_t)std::max(n + 1, m) + 255) / 256, 2048)),
                     dim3(256), 0, s, d_row, d_col, n, m, flag);
  SMDP_HIP(hipGetLastError());
  SMDP_HIP(hipMemcpyAsync(ctx->pinned, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  SMDP_HIP(hipStreamSynchronize(s));
  if (ctx->pinned[0]) return SMDP_ERR_ARG;

  smdp_graph g;  // every buffer is the caller's or the fixed array's: nothing to free
  g.ctx = ctx;
  g.n = n;
  g.m = m;
  g.row = const_cast<int32_t*>(d_row);
  g.col = const_cast<int32_t*>(d_col);
  g.w = const_cast<int32_t*>(d_w);
  g.dist = d_dist;
  g.stamp = reinterpret_cast<int32_t*>(base + noalloc::kStamp);
  g.q[0] = reinterpret_cast<int32_t*>(base + noalloc::kQ0);
  g.q[1] = reinterpret_cast<int32_t*>(base + noalloc::kQ1);
  g.hq = reinterpret_cast<Chunk*>(base + noalloc::kHq);
  g.hq_cap = noalloc::kHqCap;
  g.ctl = reinterpret_cast<Ctl*>(base + noalloc::kCtl);
  g.cnt = reinterpret_cast<Counters*>(base + noalloc::kCnt);
  g.transient = true;  // no renumbered copy
  g.forest = 0;
  if (sol == SMDP_SOL_TUNED) {
    // the has-arcs bits are rebuilt every call (the array keeps the last call's contents)
    g.hasout = reinterpret_cast<uint32_t*>(base + noalloc::kHasout);
    const int64_t nw = ((int64_t)n + 31) / 32;
    hipLaunchKernelGGL(k_hasout, dim3((unsigned)std::min<int64_t>((nw + 255) / 256, 4096)), dim3(256), 0, s,
                       d_row, n, g.hasout);
    SMDP_HIP(hipGetLastError());
  }
  // k_init_frontier clears the control blocks and counters and sets stamp; the
  // queues and the grid queue are written before they are read
  const int prof = ctx->profiling;
  ctx->profiling = 0;  // profiling would create events
  const int rc = run_frontier(&g, src, sol, false, st);
  ctx->profiling = prof;
  return rc;
}

// Host-array drop-in: upload, run, copy back (SSSP_GPU / BFS_REC_GPU shape).
int frontier_dropin(smdp_ctx* ctx, const smdp_csr* h, int32_t src, int sol, bool bfs, int32_t* out,
                    smdp_stats* st) {
  if (!ctx || !h || !out) return SMDP_ERR_ARG;
  double t0 = now_ms();
  smdp_graph* g = nullptr;
  // one solve: a renumbered copy would not pay for itself, and the buffers
  // come from the context's arena
  SMDP_CHECK(graph_create_transient(ctx, h, &g));
  int rc = run_frontier(g, src, sol, bfs, st);
  if (rc == SMDP_OK) rc = smdp_graph_read(g, bfs ? SMDP_RES_LEVEL : SMDP_RES_DIST, out, g->n);
  smdp_graph_destroy(g);
  if (st) st->ms_total = now_ms() - t0;
  return rc;
}

}  // namespace smdp

using namespace smdp;

extern "C" {

int smdp_sssp_run(smdp_graph* g, int32_t src, int solution, smdp_stats* st) {
  return run_frontier(g, src, solution, false, st);
}

int smdp_bfs_run(smdp_graph* g, int32_t src, int solution, smdp_stats* st) {
  return run_frontier(g, src, solution, true, st);
}

int smdp_sssp(smdp_ctx* ctx, const smdp_csr* g, int32_t src, int solution, int32_t* dist, smdp_stats* st) {
  return frontier_dropin(ctx, g, src, solution, false, dist, st);
}

int smdp_bfs(smdp_ctx* ctx, const smdp_csr* g, int32_t src, int solution, int32_t* level, smdp_stats* st) {
  return frontier_dropin(ctx, g, src, solution, true, level, st);
}

int smdp_sssp_device_noalloc(smdp_ctx* ctx, int32_t n, int32_t m, const int32_t* d_row, const int32_t* d_col,
                             const int32_t* d_w, int32_t src, int solution, int32_t* d_dist, smdp_stats* st) {
  return sssp_noalloc(ctx, n, m, d_row, d_col, d_w, src, solution, d_dist, st);
}

}  // extern "C"
