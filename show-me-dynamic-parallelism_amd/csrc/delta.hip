// Bucketed SSSP (delta-stepping) for the tuned solution on gfx950.
//
// The reference relaxes in unordered rounds: every queued vertex relaxes
// every arc, and a vertex is queued again each time its cost drops
// (apps/sssp/src/sssp_kernel.cu:40-61, :496-506; apps/sssp/src/
// sssp_wrapper.cu:147-302).  On RMAT-22 with weights in [1,255] that relaxes
// about 3.3 arcs per reachable arc, most of them from hubs reached first by
// long paths and expanded again when a shorter one arrives.  Distances are
// the same for every relaxation order, so here vertices are taken in order of
// distance buckets of width delta = 2^shift (Meyer and Sanders' delta-stepping)
// and each arc is relaxed about once.
//
// A bucket b (distances in [b*delta, (b+1)*delta)) is processed in phases:
//   light sweep  every vertex is read; those in bucket b relax the light
//                arcs of their row (w < delta) whose candidate stays in
//                bucket b;
//   light        vertices lowered within bucket b by the previous phase are
//                expanded the same way from a queue (one entry per vertex per
//                phase, deduplicated by a phase stamp), until a phase lowers
//                none;
//   settle       every vertex is read again; those in bucket b now have their
//                final distance and relax, once, every arc of their row whose
//                candidate lies beyond bucket b.  The pass also records the
//                smallest distance beyond bucket b (distances it reads and
//                candidates that won their atomicMin), which names the next
//                bucket, so empty buckets cost nothing.  A settle can instead
//                pull: every vertex beyond bucket b reads its in-arcs (from an
//                in-arc copy of the graph) and takes the smallest candidate
//                from a source below the bucket's end.  No such candidate is
//                below the end (the light phases relaxed those), so the
//                result is the push settle's.  The light passes count the in-
//                and out-degrees of the vertices they bring below the bucket's
//                end; the settle pulls when the in-arcs beyond the bucket are
//                fewer than the bucket's out-arcs.  On RMAT-22 bucket 0's
//                settle reads 4.8M in-arcs instead of relaxing 57.5M out-arcs
//                (0.65 -> 0.13 ms); on a grid the bucket's out-arcs stay fewer
//                and every settle pushes.
// When a settle pass sees few vertices left beyond its bucket (at most
// n >> kTailShift, and at most 1/kTailRatio of those settled so far: a small
// frontier near the source is not a tail), the rest of the solve is one
// open-ended bucket: its light phases relax whole rows with no upper limit
// (unordered rounds, as in the reference) until no distance changes.  The
// tail of a skewed graph spreads its last vertices over many sparse buckets,
// each of which would cost a sweep, a light phase and a settle.  On RMAT-22
// (n/256, 1/32) started the tail at bucket 5 and took 2.03 ms a solve;
// (n/16, 1/4) 1.92 ms; (n/8, 1/2), at bucket 2, 1.84 ms with the same arcs
// relaxed per reachable arc (1.115); a uniform graph of 4M vertices and
// degree 16 took 3.01, 2.95 and 2.95 ms.
//
// A vertex is expanded only while its distance is in the current bucket, and
// it relaxes each arc in at most one of the light and settle passes with its
// final distance (plus light re-expansions inside the bucket), so on RMAT-22
// the arcs relaxed fall to about 1.06 per reachable arc.
//
// Each row is stored in weight bands: kBands bands of width delta / kBands
// for the light arcs, lightest first, then the heavy arcs, with the band
// starts in a table of kBands ints per row.  A vertex at distance du relaxes
// in a light pass only arcs with w < lim - du and in the settle pass only
// arcs with w >= lim - du (lim = the bucket's end), so each pass reads the
// bands that hold its arcs and filters one partial band.  Reading whole light
// parts instead read about four arcs per relaxed one in the light phases
// near the source of an RMAT graph, where hubs are expanded again each time
// their distance falls.
//
// Phases are launched back to back; each kernel computes its own phase from
// the previous phase's descriptor and counters (values no kernel of the
// current phase writes), so the host enqueues phases blindly and reads the
// state back once per batch.  A phase's first pass runs per wave, with no
// workgroup barrier (a compute unit's waves wait on their memory steps
// independently): a wave reads 128 entries, lays their arc ranges end to end
// and relaxes them 256 at a time.  Vertices with many arcs are cut into
// chunks on the grid queue and drained by a second kernel per phase, as in
// the unordered tuned solution.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "frontier.h"
#include "grid_barrier.h"
#include "smdp_internal.h"

namespace smdp {

#ifndef SMDP_DELTA_ILP
#define SMDP_DELTA_ILP 4
#endif
constexpr int kDIlp = SMDP_DELTA_ILP;  // arcs per thread per step
#ifndef SMDP_DELTA_SMALL_SPLIT
#define SMDP_DELTA_SMALL_SPLIT 256
#endif
constexpr int kDSmallSplit = SMDP_DELTA_SMALL_SPLIT;  // split degree when the light queue is short
#ifndef SMDP_DELTA_DRAIN_TILE
#define SMDP_DELTA_DRAIN_TILE 8
#endif
constexpr int kDDrainTile = SMDP_DELTA_DRAIN_TILE;  // grid-queue entries per drain tile (a power of two, at most 64)
#ifndef SMDP_DELTA_WAVE_GROUPS
#define SMDP_DELTA_WAVE_GROUPS 2
#endif
constexpr int kWaveGroups = SMDP_DELTA_WAVE_GROUPS;  // groups of 64 entries a wave reads at once (wave passes)
constexpr int kRunPhases = 1024;  // phases per persistent launch (the host checks the solve between launches)
#ifndef SMDP_DELTA_TAIL_SHIFT
#define SMDP_DELTA_TAIL_SHIFT 3
#endif
#ifndef SMDP_DELTA_TAIL_RATIO
#define SMDP_DELTA_TAIL_RATIO 2
#endif
constexpr int kTailShift = SMDP_DELTA_TAIL_SHIFT;  // a settle that leaves at most n >> kTailShift vertices ends in one open bucket
constexpr int kTailRatio = SMDP_DELTA_TAIL_RATIO;  // ... if at least kTailRatio times as many are settled

enum : int32_t { kSweepLight = 0, kLight = 1, kSettle = 2, kDone = 3 };

// Per-row record of the bucketed layout, one 64-byte line: the fields a pass
// reads for one vertex come in one line fill.
constexpr int kRec = 16;
constexpr int kRecEnd = kBands + 1;    // [0] row start, [1 .. kBands] band starts (kBands: the heavy part), then row end
constexpr int kRecIn = kBands + 2;     // in-row start (pull settles)
constexpr int kRecInEnd = kBands + 3;  // in-row end
constexpr int kRecDeg = kBands + 4;    // out-degree, then in-degree (the census)
static_assert(kRecDeg + 2 <= kRec && kRecDeg % 2 == 0, "record fields");
typedef int int2v __attribute__((ext_vector_type(2)));  // {vertex, weight} of an arc
// A light-queue entry with this bit set is a vertex its candidate brought from
// at or beyond the bucket's end to below it: its expansion adds its degrees
// to the census (from the record line it loads anyway).
constexpr int32_t kJoinFlag = INT32_MIN;
// How a settle pass reaches the arcs that leave its bucket (smdp_params.settle).
enum : int32_t { kSettleAuto = 0, kSettlePush = 1, kSettlePull = 2 };

// Descriptor of one phase; phase p's lives in slot p & 3 (one 128-byte line).
struct DPhase {
  int32_t mode;
  int32_t bucket;
  int32_t nin;   // entries: queue length (light) or vertices (sweeps)
  int32_t tail;  // 1: the bucket has no upper limit (the last one)
  int32_t settled;  // vertices expanded by the settle passes before this phase
  int32_t pull;     // settle: the vertices beyond the bucket read their in-arcs
  int32_t resweep;  // a sweep of a bucket already swept once (its vertices are counted)
  int32_t pad0;
  int64_t sin;   // census: in-degrees of the vertices found below the bucket's end so far
  int64_t bout;  // census: out-degrees of the vertices found in the bucket
  int32_t pad[20];
};
static_assert(sizeof(DPhase) == 128, "one line per descriptor");
constexpr int kPhaseInts = (int)(sizeof(DPhase) / sizeof(int32_t));

// Phase after `prev`: qlen = light-queue entries pushed by prev, nb = the
// largest next-bucket code stored by prev (INT32_MAX - bucket, 0 = none),
// rem = the vertices prev's settle pass saw beyond its bucket, sel = those it
// expanded, jin / jout = the in- and out-degrees of the vertices prev found
// newly below its bucket's end.  A bucket's settle pulls when the in-arcs
// of every vertex beyond the bucket (m - sin) are fewer than the out-arcs
// of the bucket's vertices (bout), or always or never by `settle`.
__host__ __device__ inline DPhase next_phase(const DPhase& prev, int32_t qlen, int32_t nb, int32_t rem, int32_t sel,
                                             int64_t jin, int64_t jout, int32_t n, int64_t m, int32_t settle) {
  DPhase d = prev;
  d.resweep = 0;
  switch (prev.mode) {
    case kSweepLight:
    case kLight:
      d.sin = prev.sin + jin;
      d.bout = prev.bout + jout;
      if (qlen > n) {
        // more pushes than the queue holds (a vertex can be queued twice in a
        // race): its entries are incomplete, so the bucket is swept again
        d.mode = kSweepLight;
        d.nin = n;
        d.resweep = 1;
      } else if (qlen > 0) {
        d.mode = kLight;
        d.nin = qlen;
      } else if (prev.tail) {
        d.mode = kDone;
        d.nin = 0;
      } else {
        d.mode = kSettle;
        d.nin = n;
        d.pull = settle == kSettlePull || (settle == kSettleAuto && m - d.sin < d.bout);
      }
      break;
    case kSettle:
      d.settled = prev.settled + sel;
      d.pull = 0;
      d.bout = 0;
      if (nb > 0) {
        d.mode = kSweepLight;
        d.bucket = INT32_MAX - nb;
        d.nin = n;
        d.tail = rem <= (n >> kTailShift) && (int64_t)rem * kTailRatio <= d.settled;
      } else {
        d.mode = kDone;
        d.nin = 0;
      }
      break;
    default:
      d.mode = kDone;
      d.nin = 0;
      break;
  }
  return d;
}

struct DArgs {
  const int32_t* __restrict__ row;
  const int32_t* __restrict__ rec;     // kRec ints per row: start, band starts, end, in-row, degrees
  const int2v* __restrict__ cw;        // arcs in band order: {target, weight}
  const uint32_t* __restrict__ hasout; // bit v set when row v has arcs
  int32_t* dist;
  int32_t* stamp;  // phase label of the last light-queue push
  int32_t* q0;
  int32_t* q1;
  Chunk* hq;
  Ctl* ctl;
  DPhase* ph;
  Counters* cnt;
  // pull settle: in-arcs {source, weight} of the in-rows the records name,
  // bit v set when v has in-arcs; null when settle is kSettlePush
  const int2v* __restrict__ rcw;
  const uint32_t* __restrict__ hasin;
  int64_t m;
  int32_t settle;   // kSettleAuto, kSettlePush or kSettlePull
  int32_t small_q;  // a light queue shorter than this cuts entries of >= kDSmallSplit arcs into chunks
  unsigned long long* rlog;  // profiling: per phase (entries, arcs relaxed), or nullptr
  int32_t* progress;         // host memory: p + 1 when phase p starts, -(p + 1) when it is the kDone phase
  int32_t n;
  int32_t shift;    // delta = 1 << shift
  int32_t bshift;   // band width = 1 << bshift (delta / kBands, at least 1)
  int32_t split;    // an entry with at least this many arcs is cut into grid chunks
  int32_t chunk;    // arcs per grid chunk
};

__device__ __forceinline__ bool d_has_arcs(const DArgs& a, int v) { return (a.hasout[v >> 5] >> (v & 31)) & 1u; }
__device__ __forceinline__ bool d_has_in(const DArgs& a, int v) { return (a.hasin[v >> 5] >> (v & 31)) & 1u; }

// First arc of band j (0..kBands) of row u; band kBands is the heavy part.
__device__ __forceinline__ int band_start(const DArgs& a, int u, int j) { return a.rec[(int64_t)u * kRec + j]; }
__device__ __forceinline__ int rec_field(const DArgs& a, int u, int f) { return a.rec[(int64_t)u * kRec + f]; }
// {out-degree, in-degree} of v
__device__ __forceinline__ int2v rec_degrees(const DArgs& a, int v) {
  return *reinterpret_cast<const int2v*>(a.rec + (int64_t)v * kRec + kRecDeg);
}

// A vertex at distance du in a bucket ending at lim relaxes, in a light pass,
// the arcs with w < T = lim - du (0 < T <= delta): bands [0, ceil(T / width)),
// the last of them only in part.  Returns that range's end.
__device__ __forceinline__ int light_end(const DArgs& a, int u, int du, int lim) {
  const int T = lim - du;
  const int c = min(kBands, max(0, (T + (1 << a.bshift) - 1) >> a.bshift));
  return band_start(a, u, c);
}

// ... and in the settle pass the arcs with w >= T: from band floor(T / width)
// (in part) to the end of the row.  Returns that range's start.
__device__ __forceinline__ int settle_start(const DArgs& a, int u, int du, int lim) {
  return band_start(a, u, min(kBands, max(0, (lim - du) >> a.bshift)));
}

// Phase p, computed by every workgroup of phase p from phase p-1's descriptor
// (written by workgroup 0 of phase p-1) and from slot p & 3, which phase p-1
// appended to and no kernel of phase p writes.
__device__ __forceinline__ DPhase phase_of(const DArgs& a, int p) {
  if (p == 0) {
    DPhase d;
    d.mode = kSweepLight;
    d.bucket = 0;  // the source, at distance 0
    d.nin = a.n;
    d.tail = 0;
    d.settled = 0;
    d.pull = 0;
    d.resweep = 0;
    d.sin = 0;
    d.bout = 0;
    return d;
  }
  // the descriptor and every counter the next phase may need are read at
  // once (one round of loads, whatever prev's mode turns out to be)
  const DPhase prev = a.ph[(p - 1) & 3];
  const Ctl* c = a.ctl + (p & 3);
  const int qlen = c->qlen;
  int nb = 0, rem = 0, sel = 0;
  int64_t jin = 0, jout = 0;
#pragma unroll
  for (int k = 0; k < kClaimParts; ++k) {
    nb = max(nb, c->nbk[k * kClaimStride]);
    rem += c->rem[k * kClaimStride];
    sel += c->sel[k * kClaimStride];
    jin += c->jin[k * kClaimStride];
    jout += c->jout[k * kClaimStride];
  }
  if (prev.mode == kSettle) jin = jout = 0;
  return next_phase(prev, qlen, nb, rem, sel, jin, jout, a.n, a.m, a.settle);
}

// Candidate relaxations of one step, packed to the low lanes through the
// wave's LDS area (s_pair: 2 * K * 64 ints) so that each lane issues one
// atomic.  Every candidate is below the distance its thread read before, so
// the vertex's distance falls in this phase, through this candidate or a
// lower one: the atomicMin returns nothing (the wave does not wait for it),
// and every candidate counts as a winner.  Light phases queue the vertex for
// the next light phase (once per phase, by a device-coherent stamp check; a
// rare race queues it twice, which costs one more expansion); the settle
// phase keeps the smallest candidate distance of a vertex with arcs.
template <int K>
__device__ __forceinline__ void relax_packed_d(const DArgs& a, const int* v, const int* alt, const bool* want,
                                               bool settle, int label, int32_t* s_pair, Sink& out, int& nbmin,
                                               int& nrem) {
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
    bool push = false;
    if (i < total) {
      vv = s_pair[2 * i];  // may carry kJoinFlag, which goes into the queue with it
      const int x = vv & ~kJoinFlag;
      const int aa = s_pair[2 * i + 1];
      ++out.tries;
      ++out.wins;
      __hip_atomic_fetch_min(&a.dist[x], aa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (d_has_arcs(a, x)) {
        if (settle) {
          nbmin = min(nbmin, aa);
          ++nrem;
        } else if (__hip_atomic_load(&a.stamp[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != label) {
          // device-coherent read and write: a plain read could see another
          // compute unit's stale cached stamp and queue the vertex again
          __hip_atomic_store(&a.stamp[x], label, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          push = true;
        }
      }
    }
    out.push(push, vv);
  }
  // s_pair is written again by the next step
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The settle pass's smallest distance beyond the bucket and its count of
// vertices seen beyond it, over the workgroup, into the kClaimParts copies of
// slot `c` (copy blockIdx % kClaimParts); clears the thread's values (a
// persistent launch stores after each of a phase's two passes).  All
// threads; s_red needs 3 * SMDP_WAVES_PER_WG ints.
__device__ __forceinline__ void store_next_bucket(Ctl* c, int& nbmin, int& nrem, int& nsel, int shift, int* s_red) {
  constexpr int W = SMDP_WAVES_PER_WG;
  int m = nbmin;
#pragma unroll
  for (int off = SMDP_WAVE / 2; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off));
  const int r = wave_sum(nrem), q = wave_sum(nsel);
  nbmin = INT32_MAX;
  nrem = nsel = 0;
  if (lane_id() == 0) {
    s_red[threadIdx.x / SMDP_WAVE] = m;
    s_red[W + threadIdx.x / SMDP_WAVE] = r;
    s_red[2 * W + threadIdx.x / SMDP_WAVE] = q;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int rr = s_red[W], qq = s_red[2 * W];
    for (int k = 1; k < W; ++k) {
      m = min(m, s_red[k]);
      rr += s_red[W + k];
      qq += s_red[2 * W + k];
    }
    const int part = (blockIdx.x % kClaimParts) * kClaimStride;
    if (m != INT32_MAX) atomicMax(&c->nbk[part], INT32_MAX - (m >> shift));
    if (rr) atomicAdd(&c->rem[part], rr);
    if (qq) atomicAdd(&c->sel[part], qq);
  }
  __syncthreads();
}

// The census of one sweep or light pass over the workgroup (the in- and
// out-degrees of the vertices it found newly below the bucket's end), added
// to copy blockIdx % kClaimParts of slot `c`; clears the thread's sums.  All
// threads; s_red needs 2 * SMDP_WAVES_PER_WG words.
__device__ __forceinline__ void store_joins(Ctl* c, unsigned long long& jin, unsigned long long& jout,
                                            unsigned long long* s_red) {
  constexpr int W = SMDP_WAVES_PER_WG;
  const unsigned long long i = wave_sum(jin), o = wave_sum(jout);
  jin = jout = 0;
  if (lane_id() == 0) {
    s_red[threadIdx.x / SMDP_WAVE] = i;
    s_red[W + threadIdx.x / SMDP_WAVE] = o;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long si = 0, so = 0;
    for (int k = 0; k < W; ++k) {
      si += s_red[k];
      so += s_red[W + k];
    }
    const int part = (blockIdx.x % kClaimParts) * kClaimStride;
    if (si) atomicAdd(&c->jin[part], (uint32_t)si);
    if (so) atomicAdd(&c->jout[part], (uint32_t)so);
  }
  __syncthreads();
}

// Profiling log of phase p: word 0 = mode << 60 | bucket << 32 | entries
// (written by the step kernel, d != nullptr; a pull settle logs mode 4),
// word 1 += arcs relaxed.
__device__ __forceinline__ void log_phase(const DArgs& a, int p, const DPhase* d, unsigned long long wave_arcs) {
  if (!a.rlog || p >= kRoundLogCap) return;
  if (blockIdx.x == 0 && threadIdx.x == 0 && d)
    a.rlog[2 * p] = (unsigned long long)(d->mode == kSettle && d->pull ? 4 : d->mode) << 60 |
                    (unsigned long long)(uint32_t)d->bucket << 32 | (uint32_t)d->nin;
  if (lane_id() == 0 && wave_arcs) atomicAdd(&a.rlog[2 * p + 1], wave_arcs);
}

// flags[0]: set when a grid barrier gave up waiting; flags[1]: the first
// phase a persistent launch did not run.
__global__ void k_delta_init(int32_t* dist, int32_t* stamp, int32_t n, int32_t src, Ctl* ctl, Counters* cnt,
                             uint32_t* bar, int32_t* flags) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    dist[i] = (i == src) ? 0 : SMDP_SSSP_INF;
    stamp[i] = -1;
  }
  if (blockIdx.x == 0) {
    int32_t* c = reinterpret_cast<int32_t*>(ctl);
    for (int i = threadIdx.x; i < 4 * kCtlInts; i += blockDim.x) c[i] = 0;
    unsigned long long* k = reinterpret_cast<unsigned long long*>(cnt);
    for (int i = threadIdx.x; i < (int)(kCntParts * sizeof(Counters) / 8); i += blockDim.x) k[i] = 0;
    for (int i = threadIdx.x; i < kBarLines * 32; i += blockDim.x) bar[i] = 0;
    if (threadIdx.x < 4) flags[threadIdx.x] = 0;
  }
}

// Fresh barrier lines before a further persistent launch of the same solve.
__global__ void k_delta_bar_reset(uint32_t* bar) {
  for (int i = threadIdx.x; i < kBarLines * 32; i += blockDim.x) bar[i] = 0;
}

// LDS of the phase kernels.
struct DShared {
  int tile, tend, tot, ok;
  int32_t vid[kDDrainTile];    // drain: vertex of each grid-queue entry of the tile
  int32_t start[kDDrainTile];  // its first arc
  int32_t du[kDDrainTile];     // its distance
  int32_t off[kDDrainTile];    // exclusive scan of the entries' arcs
  int32_t scan[3 * SMDP_WAVES_PER_WG];
  int32_t out[SMDP_WAVES_PER_WG * kOutCap];
  int32_t pair[SMDP_WAVES_PER_WG * 2 * kDIlp * SMDP_WAVE];
  int32_t wl[SMDP_WAVES_PER_WG * 3 * kWaveGroups * SMDP_WAVE];  // wave passes: per wave, counts, bases, distances
  int32_t flush[SMDP_WAVES_PER_WG + 1];
  unsigned long long red[SMDP_WAVES_PER_WG * 8];
};

// Per-workgroup state of one phase's expansion.
struct DTile {
  int32_t* s_pair;   // this wave's 2 * kDIlp * SMDP_WAVE ints
  int32_t* wl;       // this wave's 3 * kWaveGroups * SMDP_WAVE ints
  Ctl* cin;          // the phase's input slot (grid queue, claims)
  Ctl* cout;         // the phase's output slot (light queue, next bucket)
  int split;
  int lim;
  int label;
  bool settle;
  bool pull;     // settle: the vertices beyond the bucket read their in-arcs
  bool tail;     // the open last bucket: light phases scan whole rows
  bool census;   // sweep or light pass of automatic settles: count the vertices found newly below lim
  bool members;  // ... and, in a bucket's first sweep, the vertices found in the bucket
  unsigned long long edges, skipped, pg, verts;
  unsigned long long jin, jout;  // census sums of this thread
  int nbmin, nrem, nsel;
};

// One phase's state in this workgroup; out must be initialised on the
// phase's output queue.
__device__ __forceinline__ void tile_init(const DArgs& a, const DPhase& d, int p, DShared& sh, DTile& t) {
  t.s_pair = sh.pair + (threadIdx.x / SMDP_WAVE) * 2 * kDIlp * SMDP_WAVE;
  t.wl = sh.wl + (threadIdx.x / SMDP_WAVE) * 3 * kWaveGroups * SMDP_WAVE;
  t.cin = a.ctl + (p & 3);
  t.cout = a.ctl + ((p + 1) & 3);
  // a short light queue has fewer tiles than workgroups: its entries of a
  // few dozen arcs go to the grid queue, which every workgroup drains
  t.split = (d.mode == kLight && d.nin < a.small_q) ? min(a.split, kDSmallSplit) : a.split;
  t.tail = d.tail != 0;
  t.lim = t.tail ? INT32_MAX : (d.bucket + 1) << a.shift;  // first distance beyond the bucket
  t.label = p + 1;
  t.settle = d.mode == kSettle;
  t.pull = t.settle && d.pull;
  t.census = a.settle == kSettleAuto && !t.settle && !t.tail;
  t.members = t.census && d.mode == kSweepLight && !d.resweep;
  t.jin = t.jout = 0;
  t.nbmin = INT32_MAX;
  t.nrem = 0;
  t.nsel = 0;
}

// Relaxes the arcs [s_start + j - s_off] of the NE staged entries,
// edge-balanced: kDIlp arcs per thread per step, all loads of a step issued
// before its atomics.  Every thread of the workgroup calls it.
//
// PULL (a pull settle): the entries are vertices beyond the bucket and the
// arcs their in-arcs.  An in-arc from a source already below the bucket's
// end (settled) gives the entry the candidate dist[source] + w; no such
// candidate is below lim (those were relaxed by the light passes), so the
// sources read are never written in this pass and the result does not depend
// on the order of the candidates.  A vertex's in-arcs may be spread over
// several threads: each candidate below the distance the sweep read goes
// through the same non-returning atomicMin as a push.
template <int NE, bool PULL>
__device__ __forceinline__ void relax_staged(const DArgs& a, DTile& t, const int32_t* s_vid, const int32_t* s_start,
                                             const int32_t* s_du, const int32_t* s_off, int j0, int tot, Sink& out) {
  for (int jb = j0; jb < tot; jb += kDIlp * SMDP_WG) {
    int j[kDIlp], own[kDIlp], k[kDIlp], v[kDIlp], alt[kDIlp], seen[kDIlp];
    bool rel[kDIlp], want[kDIlp];
#pragma unroll
    for (int q = 0; q < kDIlp; ++q) j[q] = min(jb + (int)threadIdx.x + q * SMDP_WG, tot - 1);
    tile_owners<kDIlp, NE>(s_off, j, own);
#pragma unroll
    for (int q = 0; q < kDIlp; ++q) k[q] = s_start[own[q]] + (j[q] - s_off[own[q]]);
    if constexpr (PULL) {
      int wk[kDIlp];
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        const int2v e = stream_load(a.rcw + k[q]);
        v[q] = e.x;  // the arc's source
        wk[q] = e.y;
      }
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) seen[q] = a.dist[v[q]];
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        const bool ok = jb + (int)threadIdx.x + q * SMDP_WG < tot;
        rel[q] = ok && seen[q] < t.lim;
        t.skipped += ok && !rel[q];
        t.edges += rel[q];
        alt[q] = rel[q] ? seen[q] + wk[q] : INT32_MAX;
        want[q] = rel[q] && alt[q] < s_du[own[q]];
        v[q] = s_vid[own[q]];  // the candidate is for the entry itself
      }
      relax_packed_d<kDIlp>(a, v, alt, want, true, t.label, t.s_pair, out, t.nbmin, t.nrem);
    } else {
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        const int2v e = stream_load(a.cw + k[q]);
        v[q] = e.x;
        alt[q] = s_du[own[q]] + e.y;
      }
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        const bool ok = jb + (int)threadIdx.x + q * SMDP_WG < tot;
        rel[q] = ok && (t.settle ? alt[q] >= t.lim : alt[q] < t.lim);
        t.skipped += ok && !rel[q];
        t.edges += rel[q];
      }
      // only arcs this pass relaxes gather their distance
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) seen[q] = rel[q] ? a.dist[v[q]] : 0;
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) want[q] = rel[q] && alt[q] < seen[q];
      // a candidate to a vertex read at or beyond lim brings that vertex
      // below lim: its queue entry is flagged for the census
      if (t.census)
#pragma unroll
        for (int q = 0; q < kDIlp; ++q) v[q] |= want[q] && seen[q] >= t.lim ? kJoinFlag : 0;
      relax_packed_d<kDIlp>(a, v, alt, want, t.settle, t.label, t.s_pair, out, t.nbmin, t.nrem);
    }
  }
}

// A vertex at distance dv is expanded by a pass over `bucket`: it lies in the
// bucket (in the open last bucket, at or beyond it).
__device__ __forceinline__ bool selected(int dv, int bucket, bool tail, int shift) {
  const int b = dv >> shift;
  return b == bucket || (tail && b > bucket);
}

// Pull settle pass (each wave on its own; no workgroup barrier).  A wave
// reads kWaveGroups groups of 64 consecutive vertices at once (groups dealt
// grid-stride over the waves): a vertex of the bucket with out-arcs counts as
// settled, one beyond it as seen, and one with in-arcs at or beyond the
// bucket's end (unreached ones included) is pulled.  The pulled vertices'
// in-arcs are laid end to end (a wave scan of their counts, kept in the
// wave's LDS area with each vertex's first in-arc and distance) and read 64 *
// kDIlp at a time, each lane finding its arc's vertex by a binary search of
// the counts; a segmented minimum over the lanes of each vertex leaves one
// candidate per vertex per step, which goes through a non-returning
// atomicMin when it is below the distance the vertex had (a vertex's arcs can
// span steps).  A vertex with split or more in-arcs goes to the grid queue
// as chunks instead.
__device__ __forceinline__ void pull_pass(const DArgs& a, const DPhase& d, DTile& t, Sink& out) {
  constexpr int G = kWaveGroups;
  constexpr int GW = G * SMDP_WAVE;
  static_assert((GW & (GW - 1)) == 0, "binary search over a power of two");
  const int lane = lane_id();
  const int gw = (int)blockIdx.x * SMDP_WAVES_PER_WG + (int)threadIdx.x / SMDP_WAVE;
  const int nw = (int)gridDim.x * SMDP_WAVES_PER_WG;
  const int ngroups = (a.n + SMDP_WAVE - 1) / SMDP_WAVE;
  int32_t* s_pre = t.wl;         // inclusive in-arc counts of the wave's G * 64 vertices
  int32_t* s_base = s_pre + GW;  // first in-arc minus the vertex's first position in the wave's arc list
  int32_t* s_dv = s_base + GW;   // distance read
  for (int g0 = gw; g0 < ngroups; g0 += G * nw) {
    int v[G], dv[G], st[G], deg[G];
    bool hi[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int grp = g0 + g * nw;
      v[g] = grp < ngroups ? grp * SMDP_WAVE + lane : a.n;
      dv[g] = -1;  // not read
      hi[g] = false;
      if (v[g] < a.n) {
        const bool ho = d_has_arcs(a, v[g]);
        hi[g] = d_has_in(a, v[g]);
        if (ho || hi[g]) dv[g] = a.dist[v[g]];
        if (ho && dv[g] != SMDP_SSSP_INF) {
          const int b = dv[g] >> a.shift;
          if (b == d.bucket) {
            ++t.nsel;
          } else if (b > d.bucket) {
            t.nbmin = min(t.nbmin, dv[g]);
            ++t.nrem;
          }
        }
      }
    }
    // in-rows of the pulled vertices (loads of vertex 0's for the others)
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const bool pk = hi[g] && dv[g] >= t.lim;
      const int x = pk ? v[g] : 0;
      st[g] = rec_field(a, x, kRecIn);
      deg[g] = pk ? rec_field(a, x, kRecInEnd) - st[g] : 0;
      t.verts += pk;
    }
    // long in-rows go to the grid queue
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int nch = deg[g] >= t.split ? (deg[g] + a.chunk - 1) / a.chunk : 0;
      const int incl = wave_inclusive_scan(nch);
      const int wtot = __shfl(incl, SMDP_WAVE - 1);
      if (wtot > 0) {
        int b0 = 0;
        if (lane == 0) b0 = atomicAdd(&t.cin->hlen, wtot);
        b0 = __shfl(b0, 0);
        for (int c = 0; c < nch; ++c) a.hq[b0 + incl - nch + c] = Chunk{v[g], st[g] + c * a.chunk};
        t.pg += nch;
        if (nch) deg[g] = 0;
      }
    }
    int tot = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int incl = wave_inclusive_scan(deg[g]);
      s_pre[g * SMDP_WAVE + lane] = tot + incl;
      s_base[g * SMDP_WAVE + lane] = st[g] - (tot + incl - deg[g]);
      s_dv[g * SMDP_WAVE + lane] = dv[g];
      tot += __shfl(incl, SMDP_WAVE - 1);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int f0 = 0; f0 < tot; f0 += kDIlp * SMDP_WAVE) {
      int f[kDIlp], o[kDIlp], src[kDIlp], wk[kDIlp], du[kDIlp];
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) f[q] = min(f0 + q * SMDP_WAVE + lane, tot - 1);
      // o = the number of vertices whose arcs all come before arc f
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        int x = 0;
#pragma unroll
        for (int step = GW / 2; step > 0; step >>= 1)
          if (s_pre[x + step - 1] <= f[q]) x += step;
        o[q] = x;
      }
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        const int2v e = stream_load(a.rcw + s_base[o[q]] + f[q]);
        src[q] = e.x;
        wk[q] = e.y;
      }
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) du[q] = a.dist[src[q]];
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        const bool ok = f0 + q * SMDP_WAVE + lane < tot;
        const bool r = ok && du[q] < t.lim;
        t.edges += r;
        t.skipped += ok && !r;
        // the smallest candidate over this step's lanes of vertex o[q] (the
        // lanes' vertices do not decrease with the lane; clamped lanes repeat
        // the last arc's vertex with no candidate)
        int c = r ? du[q] + wk[q] : INT32_MAX;
#pragma unroll
        for (int dd = 1; dd < SMDP_WAVE; dd <<= 1) {
          const int oo = __shfl_up(o[q], dd);
          const int cc = __shfl_up(c, dd);
          if (lane >= dd && oo == o[q]) c = min(c, cc);
        }
        const int onext = __shfl_down(o[q], 1);
        if ((lane == SMDP_WAVE - 1 || onext != o[q]) && c < s_dv[o[q]]) {
          const int x = (g0 + (o[q] / SMDP_WAVE) * nw) * SMDP_WAVE + o[q] % SMDP_WAVE;
          __hip_atomic_fetch_min(&a.dist[x], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ++out.tries;
          ++out.wins;
          if (d_has_arcs(a, x)) {
            t.nbmin = min(t.nbmin, c);
            ++t.nrem;
          }
        }
      }
    }
    // the LDS area is written again by the next groups
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Entries per group of a wave pass over L entries: 64 for sweeps; a short
// light queue is spread over the grid's waves a few entries each (the next
// power of two of L / waves), so that its arcs are not left to a few waves.
__device__ __forceinline__ int wave_group(int L, bool queue) {
  if (!queue) return SMDP_WAVE;
  const int nw = (int)gridDim.x * SMDP_WAVES_PER_WG;
  const int per = (L + nw - 1) / nw;
  int e = 1;
  while (e < per && e < SMDP_WAVE) e <<= 1;
  return e;
}

// Pass 1 of a sweep, light or push-settle phase, each wave on its own (no
// workgroup barrier, so a compute unit's waves wait on their memory steps
// independently).  Entries are the vertices 0 .. n-1 (sweeps) or the light
// queue's entries, in groups of 64 dealt grid-stride over the waves, a wave
// reading kWaveGroups groups at once.  The selected entries' arc ranges (the
// pass's bands of their rows) are laid end to end by a wave scan, kept in the
// wave's LDS area with each entry's first arc and distance, and read 64 *
// kDIlp at a time, each lane finding its arc's entry by a binary search of
// the counts; candidates go through relax_packed_d.  An entry with split or
// more arcs goes to the grid queue as chunks instead.
__device__ __forceinline__ void wave_pass(const DArgs& a, const DPhase& d, int p, DTile& t, Sink& out) {
  constexpr int G = kWaveGroups;
  constexpr int GW = G * SMDP_WAVE;
  static_assert((GW & (GW - 1)) == 0, "binary search over a power of two");
  const int lane = lane_id();
  const int gw = (int)blockIdx.x * SMDP_WAVES_PER_WG + (int)threadIdx.x / SMDP_WAVE;
  const int nw = (int)gridDim.x * SMDP_WAVES_PER_WG;
  const int L = d.nin;
  const int bucket = d.bucket;
  const int32_t* qin = d.mode == kLight ? ((p & 1) ? a.q1 : a.q0) : nullptr;  // queued vertices have arcs
  const int E = wave_group(L, qin != nullptr);  // entries per group (lanes E .. 63 idle while reading entries)
  const int ngroups = (L + E - 1) / E;
  int32_t* s_pre = t.wl;         // inclusive arc counts of the wave's G * 64 entries
  int32_t* s_base = s_pre + GW;  // first arc minus the entry's first position in the wave's arc list
  int32_t* s_du = s_base + GW;   // the entry's distance
  for (int g0 = gw; g0 < ngroups; g0 += G * nw) {
    int v[G], dv[G], st[G], deg[G];
    bool sel[G], join[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int i = (g0 + g * nw) * E + lane;
      const bool in = lane < E && g0 + g * nw < ngroups && i < L;
      const int raw = in ? (qin ? qin[i] : i) : -1;
      join[g] = in && qin && raw < 0;
      v[g] = in ? raw & ~kJoinFlag : -1;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      dv[g] = SMDP_SSSP_INF;
      if (v[g] >= 0 && (qin || d_has_arcs(a, v[g]))) dv[g] = a.dist[v[g]];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      sel[g] = dv[g] != SMDP_SSSP_INF && selected(dv[g], bucket, t.tail, a.shift);
      if (t.settle && !sel[g] && dv[g] != SMDP_SSSP_INF && (dv[g] >> a.shift) > bucket) {
        t.nbmin = min(t.nbmin, dv[g]);
        ++t.nrem;
      }
    }
    // the pass's arc range of each selected entry (vertex 0's for the others)
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int x = sel[g] ? v[g] : 0;
      const int du = sel[g] ? dv[g] : 0;
      int e;
      if (t.tail) {
        st[g] = rec_field(a, x, 0);
        e = rec_field(a, x, kRecEnd);
      } else if (t.settle) {
        st[g] = settle_start(a, x, du, t.lim);
        e = rec_field(a, x, kRecEnd);
      } else {
        st[g] = rec_field(a, x, 0);
        e = light_end(a, x, du, t.lim);
      }
      deg[g] = sel[g] ? e - st[g] : 0;
      t.verts += sel[g];
      t.nsel += sel[g];
      // the census: the bucket's first sweep counts the vertices it finds in
      // the bucket, a light pass those flagged when they came below lim (the
      // degrees share the record's line; vertices without out-arcs are never
      // expanded or queued and stay out, which makes a pull look a little
      // dearer)
      if ((t.members || (t.census && join[g])) && sel[g]) {
        const int2v dg = rec_degrees(a, x);
        t.jout += (unsigned)dg.x;
        t.jin += (unsigned)dg.y;
      }
    }
    // long ranges go to the grid queue
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int nch = deg[g] >= t.split ? (deg[g] + a.chunk - 1) / a.chunk : 0;
      const int incl = wave_inclusive_scan(nch);
      const int wtot = __shfl(incl, SMDP_WAVE - 1);
      if (wtot > 0) {
        int b0 = 0;
        if (lane == 0) b0 = atomicAdd(&t.cin->hlen, wtot);
        b0 = __shfl(b0, 0);
        for (int c = 0; c < nch; ++c) a.hq[b0 + incl - nch + c] = Chunk{v[g], st[g] + c * a.chunk};
        t.pg += nch;
        if (nch) deg[g] = 0;
      }
    }
    int tot = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int incl = wave_inclusive_scan(deg[g]);
      s_pre[g * SMDP_WAVE + lane] = tot + incl;
      s_base[g * SMDP_WAVE + lane] = st[g] - (tot + incl - deg[g]);
      s_du[g * SMDP_WAVE + lane] = dv[g];
      tot += __shfl(incl, SMDP_WAVE - 1);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int f0 = 0; f0 < tot; f0 += kDIlp * SMDP_WAVE) {
      int f[kDIlp], o[kDIlp], vt[kDIlp], alt[kDIlp], seen[kDIlp];
      bool rel[kDIlp], want[kDIlp];
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) f[q] = min(f0 + q * SMDP_WAVE + lane, tot - 1);
      // o = the number of entries whose arcs all come before arc f
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        int x = 0;
#pragma unroll
        for (int step = GW / 2; step > 0; step >>= 1)
          if (s_pre[x + step - 1] <= f[q]) x += step;
        o[q] = x;
      }
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        const int2v e = stream_load(a.cw + s_base[o[q]] + f[q]);
        vt[q] = e.x;
        alt[q] = s_du[o[q]] + e.y;
      }
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) {
        const bool ok = f0 + q * SMDP_WAVE + lane < tot;
        rel[q] = ok && (t.settle ? alt[q] >= t.lim : alt[q] < t.lim);
        t.skipped += ok && !rel[q];
        t.edges += rel[q];
      }
      // only arcs this pass relaxes gather their distance
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) seen[q] = rel[q] ? a.dist[vt[q]] : 0;
#pragma unroll
      for (int q = 0; q < kDIlp; ++q) want[q] = rel[q] && alt[q] < seen[q];
      // as in relax_staged: entries of vertices brought below lim are flagged
      if (t.census)
#pragma unroll
        for (int q = 0; q < kDIlp; ++q) vt[q] |= want[q] && seen[q] >= t.lim ? kJoinFlag : 0;
      relax_packed_d<kDIlp>(a, vt, alt, want, t.settle, t.label, t.s_pair, out, t.nbmin, t.nrem);
    }
    // the LDS area is written again by the next groups
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Pass 1 of phase d (all threads of the workgroup, each wave on its own).
__device__ __forceinline__ void phase_tiles(const DArgs& a, const DPhase& d, int p, DShared& sh, DTile& t,
                                            Sink& out) {
  (void)sh;
  if (t.pull) pull_pass(a, d, t, out);
  else wave_pass(a, d, p, t, out);
}

// Pass 2 of phase d (all threads): the H chunks of the phase's grid queue,
// kDDrainTile entries per workgroup pass (fewer while the tiles would not
// reach every workgroup), expanded edge-balanced.  Returns the arcs relaxed.
__device__ __forceinline__ unsigned long long phase_chunks(const DArgs& a, const DPhase& d, int H, DShared& sh,
                                                           DTile& t, Sink& out) {
  int ept = kDDrainTile;
  while (ept > 1 && (H + ept / 2 - 1) / (ept / 2) <= (int)gridDim.x) ept >>= 1;
  const int ntiles = (H + ept - 1) / ept;
  const unsigned long long e0 = t.edges;
  uint32_t dead = 0;  // claim state of wave 0
  for (;;) {
    if (threadIdx.x < SMDP_WAVE) {
      int f = 0, e = 0;
      const bool got = claim_items_wave(t.cin->chunks, ntiles, 1, dead, &f, &e);
      if (threadIdx.x == 0) sh.tile = got ? f : -1;
    }
    __syncthreads();
    const int tile = sh.tile;
    if (tile < 0) break;
    if (threadIdx.x < SMDP_WAVE) {
      const int idx = tile * ept + (int)threadIdx.x;
      int len = 0, st = 0, du = 0, v = 0;
      if ((int)threadIdx.x < ept && idx < H) {
        const Chunk c = a.hq[idx];
        v = c.v;
        du = a.dist[c.v];
        // a light entry's distance may have fallen since it was cut: its
        // range only grows, and the vertex is queued again for the rest (a
        // pulling vertex is written only by its own chunks, so du stays at
        // or beyond lim)
        const int vend = t.pull ? rec_field(a, c.v, kRecInEnd)
                                : (t.settle || t.tail ? rec_field(a, c.v, kRecEnd) : light_end(a, c.v, du, t.lim));
        const int end = vend - c.first > a.chunk ? c.first + a.chunk : vend;
        len = end - c.first;
        st = c.first;
      }
      const int incl = wave_inclusive_scan(len);
      if ((int)threadIdx.x < kDDrainTile) {
        sh.off[threadIdx.x] = incl - len;
        sh.vid[threadIdx.x] = v;
        sh.start[threadIdx.x] = st;
        sh.du[threadIdx.x] = du;
      }
      if (threadIdx.x == kDDrainTile - 1) sh.tot = incl;
    }
    __syncthreads();
    if (t.pull) relax_staged<kDDrainTile, true>(a, t, sh.vid, sh.start, sh.du, sh.off, 0, sh.tot, out);
    else relax_staged<kDDrainTile, false>(a, t, sh.vid, sh.start, sh.du, sh.off, 0, sh.tot, out);
    __syncthreads();
  }
  return t.edges - e0;
}

// End of a pass: the workgroup's staged queue entries go out, a settle pass
// records its next-bucket minimum and a counting pass its census.  All
// threads.
__device__ __forceinline__ void end_pass(const DArgs& a, DShared& sh, DTile& t, Sink& out) {
  wg_flush(out, sh.flush);
  if (t.settle) store_next_bucket(t.cout, t.nbmin, t.nrem, t.nsel, a.shift, sh.scan);
  else if (t.census) store_joins(t.cout, t.jin, t.jout, sh.red);
}

// The workgroup's work counters, one atomic per counter.  All threads.
__device__ __forceinline__ void phase_counters(const DArgs& a, DShared& sh, DTile& t, Sink& out,
                                               unsigned long long e2) {
  const unsigned long long v[8] = {t.edges, t.verts, out.pushes, t.pg, out.tries, out.wins, t.skipped, e2};
  const int f[8] = {kCntEdges, kCntVerts, kCntPush0, kCntPush3, kCntTries, kCntWins, kCntSkipped, kCntEdges2};
  wg_add_counters<8>(a.cnt, v, f, sh.red);
}

__device__ __forceinline__ void phase_begin(const DArgs& a, const DPhase& d, int p) {
  if (blockIdx.x == 0) {
    ctl_clear(a.ctl, (p + 2) & 3);  // phase p + 1's output slot; no kernel of phase p uses it
    if (threadIdx.x == 0) {
      a.ph[p & 3] = d;
      if (d.mode != kDone) atomicMax(&cnt_of(a.cnt)->rounds, (unsigned long long)(p + 1));
      if (a.progress)
        __hip_atomic_store(a.progress, d.mode == kDone ? -(p + 1) : p + 1, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// Out-arcs of the reached vertices into the work counters (grid-stride, 8
// vertices per thread with their loads issued at once).
__device__ __forceinline__ void count_reached(const DArgs& a) {
  constexpr int K = 8;
  unsigned long long acc = 0;
  const int64_t stride = (int64_t)gridDim.x * SMDP_WG;
  for (int64_t v0 = (int64_t)blockIdx.x * SMDP_WG + threadIdx.x; v0 < a.n; v0 += K * stride) {
    int d[K], r0[K], r1[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t v = min(v0 + k * stride, (int64_t)a.n - 1);
      d[k] = a.dist[v];
      r0[k] = a.row[v];
      r1[k] = a.row[v + 1];
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (v0 + k * stride < a.n && d[k] != SMDP_SSSP_INF) acc += (unsigned long long)(r1[k] - r0[k]);
  }
  acc = wave_sum(acc);
  if (lane_id() == 0 && acc) atomicAdd(&cnt_of(a.cnt)->reached, acc);
}

// One phase's pass 1 (launched per phase).  The first phase past the end of
// the solve counts the out-arcs of the reached vertices (the solve's work in
// traversed edges), in a launch the host enqueued anyway.
__global__ __launch_bounds__(SMDP_WG) void k_delta_step(DArgs a, int p) {
  const DPhase d = phase_of(a, p);
  phase_begin(a, d, p);
  if (d.mode == kDone) {
    if (p > 0 && a.ph[(p - 1) & 3].mode != kDone) count_reached(a);
    return;
  }
  // more workgroups than a short light queue has groups
  if (d.mode == kLight &&
      (int)blockIdx.x * SMDP_WAVES_PER_WG >= (d.nin + wave_group(d.nin, true) - 1) / wave_group(d.nin, true))
    return;
  __shared__ DShared sh;
  DTile t;
  tile_init(a, d, p, sh, t);
  t.edges = t.skipped = t.pg = t.verts = 0;
  Sink out;
  out.init(sh.out + (threadIdx.x / SMDP_WAVE) * kOutCap, (p & 1) ? a.q0 : a.q1, &t.cout->qlen, a.n);
  phase_tiles(a, d, p, sh, t, out);
  log_phase(a, p, &d, wave_sum(t.edges));
  end_pass(a, sh, t, out);
  phase_counters(a, sh, t, out, 0);
}

// The same phase's pass 2: its grid chunks.
__global__ __launch_bounds__(SMDP_WG) void k_delta_drain(DArgs a, int p) {
  const DPhase d = a.ph[p & 3];
  if (d.mode == kDone) return;
  const int H = a.ctl[p & 3].hlen;
  int ept = kDDrainTile;
  while (ept > 1 && (H + ept / 2 - 1) / (ept / 2) <= (int)gridDim.x) ept >>= 1;
  if ((int)blockIdx.x >= (H + ept - 1) / ept) return;
  __shared__ DShared sh;
  DTile t;
  tile_init(a, d, p, sh, t);
  t.edges = t.skipped = t.pg = t.verts = 0;
  Sink out;
  out.init(sh.out + (threadIdx.x / SMDP_WAVE) * kOutCap, (p & 1) ? a.q0 : a.q1, &t.cout->qlen, a.n);
  const unsigned long long e2 = phase_chunks(a, d, H, sh, t, out);
  log_phase(a, p, nullptr, wave_sum(e2));
  end_pass(a, sh, t, out);
  phase_counters(a, sh, t, out, e2);
}

// Phases [p0, pend) in one launch whose workgroups are all resident, with a
// grid barrier after each pass (the chunk pass only when the phase pushed
// chunks).  Every workgroup computes the same phase descriptor after each
// barrier, so all take the same barriers and stop at the same phase.
// flags[1] = the first phase not run (the finished solve's kDone phase, or
// pend); flags[0] = 1 when a barrier gave up waiting.
__global__ __launch_bounds__(SMDP_WG) void k_delta_run(DArgs a, int p0, int pend, uint32_t* bar, int32_t* flags) {
  __shared__ DShared sh;
  DTile t;
  t.edges = t.skipped = t.pg = t.verts = 0;
  unsigned long long e2 = 0;
  Sink out;
  out.init(sh.out + (threadIdx.x / SMDP_WAVE) * kOutCap, a.q0, &a.ctl[0].qlen, a.n);
  uint32_t epoch = 0;
  int p = p0;
  for (; p < pend; ++p) {
    if (__hip_atomic_load(&flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    const DPhase d = phase_of(a, p);
    phase_begin(a, d, p);
    if (d.mode == kDone) break;
    tile_init(a, d, p, sh, t);
    out.q = (p & 1) ? a.q0 : a.q1;
    out.qlen = &t.cout->qlen;
    const unsigned long long eb = t.edges;
    phase_tiles(a, d, p, sh, t, out);
    end_pass(a, sh, t, out);
    if (!grid_barrier(bar, ++epoch, &sh.ok)) {
      if (threadIdx.x == 0) atomicExch(&flags[0], 1);
      break;
    }
    const int H = __hip_atomic_load(&t.cin->hlen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (H > 0) {
      e2 += phase_chunks(a, d, H, sh, t, out);
      end_pass(a, sh, t, out);
      if (!grid_barrier(bar, ++epoch, &sh.ok)) {
        if (threadIdx.x == 0) atomicExch(&flags[0], 1);
        break;
      }
    }
    log_phase(a, p, &d, wave_sum(t.edges - eb));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) flags[1] = p;
  phase_counters(a, sh, t, out, e2);
}

// Group of an arc for the band layout: its light band (w >> bshift, below
// kBands) or kBands for a heavy arc (w >= delta).
__device__ __forceinline__ int arc_group(int w, int delta, int bshift) { return w >= delta ? kBands : w >> bshift; }

// Rows of the renumbered copy reordered by weight band: band j < kBands holds
// the light arcs with j * width <= w < (j + 1) * width (width = 1 << bshift =
// delta / kBands, at least 1), band kBands the heavy arcs (w >= delta); each
// band keeps the arcs' original order.  Writes the reordered arcs twice (as
// separate columns and weights for the unordered rounds, as {target, weight}
// pairs for the bucketed passes) and the row's record: start, the first arc
// of each band j >= 1, end and out-degree.  One wave per row; lane j counts
// group j.
__global__ __launch_bounds__(SMDP_WG) void k_band_rows(const int32_t* __restrict__ row,
                                                       const int32_t* __restrict__ col_in,
                                                       const int32_t* __restrict__ w_in, int32_t* col_out,
                                                       int32_t* w_out, int2v* cw_out, int32_t* rec, int32_t n,
                                                       int32_t delta, int32_t bshift) {
  static_assert(kBands < SMDP_WAVE, "one lane per group");
  const int lane = lane_id();
  const int64_t nw = (int64_t)gridDim.x * SMDP_WAVES_PER_WG;
  for (int64_t u = (int64_t)blockIdx.x * SMDP_WAVES_PER_WG + threadIdx.x / SMDP_WAVE; u < n; u += nw) {
    const int s = row[u], e = row[u + 1];
    int cnt = 0;
    for (int b = s; b < e; b += SMDP_WAVE) {
      const int k = b + lane;
      const int g = k < e ? arc_group(w_in[k], delta, bshift) : -1;
#pragma unroll
      for (int j = 0; j <= kBands; ++j) {
        const int c = __popcll(ballot(g == j));
        if (lane == j) cnt += c;
      }
    }
    const int mine = lane <= kBands ? cnt : 0;
    int base = s + wave_inclusive_scan(mine) - mine;  // lane j: first arc of group j
    int32_t* r = rec + u * kRec;
    if (lane <= kBands) r[lane] = base;  // lane 0: the row's start
    if (lane == kRecEnd) r[lane] = e;
    if (lane == kRecDeg) r[lane] = e - s;
    for (int b = s; b < e; b += SMDP_WAVE) {
      const int k = b + lane;
      const bool in = k < e;
      const int wk = in ? w_in[k] : 0;
      const int g = in ? arc_group(wk, delta, bshift) : -1;
      int pos = 0;
#pragma unroll
      for (int j = 0; j <= kBands; ++j) {
        const uint64_t m = ballot(g == j);
        const int bj = __shfl(base, j);
        if (g == j) pos = bj + mbcnt(m);
        if (lane == j) base += __popcll(m);
      }
      if (in) {
        const int c = col_in[k];
        col_out[pos] = c;
        w_out[pos] = wk;
        cw_out[pos] = int2v{c, wk};
      }
    }
  }
}

// In-arcs of the copy: row u's arcs scattered to their targets' in-rows
// (cur = the in-row starts, advanced by one atomic per arc), one wave per
// row.  The order inside an in-row follows the atomics; a pull settle takes
// the minimum over it, which does not depend on that order.
__global__ __launch_bounds__(SMDP_WG) void k_in_arcs(const int32_t* __restrict__ row, const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ w, int32_t n, int32_t* cur,
                                                     int2v* rcw) {
  const int lane = lane_id();
  const int64_t nw = (int64_t)gridDim.x * SMDP_WAVES_PER_WG;
  for (int64_t u = (int64_t)blockIdx.x * SMDP_WAVES_PER_WG + threadIdx.x / SMDP_WAVE; u < n; u += nw) {
    const int e = row[u + 1];
    for (int k = row[u] + lane; k < e; k += SMDP_WAVE) {
      const int pos = atomicAdd(&cur[col[k]], 1);
      rcw[pos] = int2v{(int32_t)u, w[k]};
    }
  }
}

// In-row start, end and in-degree of every vertex into its record.
__global__ void k_rec_in(const int32_t* __restrict__ rrow, int32_t n, int32_t* rec) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const int s = rrow[v], e = rrow[v + 1];
    int32_t* r = rec + v * kRec;
    r[kRecIn] = s;
    r[kRecInEnd] = e;
    r[kRecDeg + 1] = e - s;
  }
}

// The copy's in-arcs, in-row fields of the records and in-arc bits for the
// pull settle, built once per resident graph (after the records) from the
// in-row offsets the renumbering computed.
static int ensure_pull(smdp_graph* g, Relabel* rl) {
  if (rl->rrow) return SMDP_OK;
  const int64_t n = g->n, m = std::max<int64_t>(1, g->m);
  if ((int64_t)rl->h_rrow.size() != n + 1 || !rl->w || !rl->rec) return SMDP_ERR_ARG;
  hipStream_t s = g->ctx->stream;
  int32_t *rrow = nullptr, *rcw = nullptr, *cur = nullptr;
  auto release = [&]() {
    for (int32_t* q : {rrow, rcw, cur})
      if (q) (void)hipFree(q);
  };
  if (hipMalloc((void**)&rrow, sizeof(int32_t) * (n + 1)) != hipSuccess ||
      hipMalloc((void**)&rcw, sizeof(int2v) * m) != hipSuccess ||
      hipMalloc((void**)&cur, sizeof(int32_t) * (n + 1)) != hipSuccess) {
    release();
    return SMDP_ERR_NOMEM;
  }
  bool ok = hipMemcpyAsync(rrow, rl->h_rrow.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s) ==
                hipSuccess &&
            hipMemcpyAsync(cur, rrow, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToDevice, s) == hipSuccess;
  if (ok && g->m > 0) {
    const int64_t blocks = std::min<int64_t>((n + SMDP_WAVES_PER_WG - 1) / SMDP_WAVES_PER_WG, 16384);
    hipLaunchKernelGGL(k_in_arcs, dim3((unsigned)std::max<int64_t>(1, blocks)), dim3(SMDP_WG), 0, s, rl->row,
                       rl->col, rl->w, g->n, cur, reinterpret_cast<int2v*>(rcw));
    ok = hipGetLastError() == hipSuccess;
  }
  if (ok && n > 0) {
    hipLaunchKernelGGL(k_rec_in, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, s, rrow,
                       g->n, rl->rec);
    ok = hipGetLastError() == hipSuccess;
  }
  if (!ok || hipStreamSynchronize(s) != hipSuccess) {
    release();
    return SMDP_ERR_DEVICE;
  }
  (void)hipFree(cur);
  cur = nullptr;
  const int rc = build_hasout(g, rrow, &rl->hasin);
  if (rc != SMDP_OK) {
    release();
    return rc;
  }
  rl->rrow = rrow;
  rl->rcw = rcw;
  std::vector<int32_t>().swap(rl->h_rrow);
  return SMDP_OK;
}

// log2 of the band width for bucket width delta (a power of two).
static int band_shift(int delta) {
  int shift = 0;
  while ((1 << shift) < delta) ++shift;
  int b = 0;
  while ((1 << b) < kBands) ++b;
  return shift > b ? shift - b : 0;
}

// Reorders the renumbered copy's rows into weight bands for bucket width
// delta (once per delta), writes their records and {target, weight} pairs, and
// allocates the phase descriptors.  The records' in-row fields, written once
// by ensure_pull, do not depend on delta.
static int ensure_buckets(smdp_graph* g, Relabel* rl, int delta) {
  hipStream_t s = g->ctx->stream;
  // 4 phase descriptors, the persistent launch's barrier lines and flags
  if (!rl->phase) SMDP_CHECK(alloc_once(&rl->phase, (int64_t)4 * kPhaseInts + kBarLines * 32 + 32));
  if (!rl->hctl && hipHostMalloc(&rl->hctl, sizeof(Ctl) + sizeof(DPhase), hipHostMallocDefault) != hipSuccess) {
    rl->hctl = nullptr;
    return SMDP_ERR_NOMEM;
  }
  if (rl->rec && rl->band_delta == delta) return SMDP_OK;
  const int64_t m = std::max<int64_t>(1, g->m);
  SMDP_CHECK(alloc_once(&rl->rec, std::max<int64_t>(1, (int64_t)g->n * kRec)));
  SMDP_CHECK(alloc_once(&rl->cw, 2 * m));
  int32_t *col2 = nullptr, *w2 = nullptr;
  if (hipMalloc((void**)&col2, sizeof(int32_t) * m) != hipSuccess) return SMDP_ERR_NOMEM;
  if (hipMalloc((void**)&w2, sizeof(int32_t) * m) != hipSuccess) {
    (void)hipFree(col2);
    return SMDP_ERR_NOMEM;
  }
  const int64_t blocks = std::min<int64_t>(((int64_t)g->n + SMDP_WAVES_PER_WG - 1) / SMDP_WAVES_PER_WG, 16384);
  hipLaunchKernelGGL(k_band_rows, dim3((unsigned)std::max<int64_t>(1, blocks)), dim3(SMDP_WG), 0, s, rl->row,
                     rl->col, rl->w, col2, w2, reinterpret_cast<int2v*>(rl->cw), rl->rec, g->n, delta,
                     band_shift(delta));
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    (void)hipFree(col2);
    (void)hipFree(w2);
    return SMDP_ERR_DEVICE;
  }
  (void)hipFree(rl->col);
  (void)hipFree(rl->w);
  rl->col = col2;
  rl->w = w2;
  rl->band_delta = delta;
  return SMDP_OK;
}

// Bucketed SSSP from src (an id of the renumbered copy); results stay in
// g->dist in the copy's numbering.  *ms is the device time from the first to
// the last launch, *launches the kernels launched.  With profiling on, fills
// the context's per-phase log.  *reached_counted: the done phase added the
// reached vertices' out-arcs to the counters, which are then already in the
// context's pinned buffer (counters_staged).  The copy must have no negative
// weight.
int run_delta(smdp_graph* g, Relabel* rl, int32_t src, const smdp_params& p, int grid, float* ms_out,
              int* launches_out, bool* reached_counted) {
  *reached_counted = false;
  smdp_ctx* ctx = g->ctx;
  if (rl->minw < 0 || (g->m > 0 && !g->w)) return SMDP_ERR_ARG;
  const int delta = p.delta > 0 ? p.delta : 64;
  SMDP_CHECK(ensure_buckets(g, rl, delta));
  SMDP_CHECK(build_hasout(g, rl->row, &rl->hasout));
  // the pull settle's in-arcs; an automatic choice that finds no room for
  // them settles by push
  int settle = p.settle;
  if (settle != kSettlePush) {
    const int rc = ensure_pull(g, rl);
    if (rc == SMDP_ERR_NOMEM && settle == kSettleAuto) settle = kSettlePush;
    else if (rc != SMDP_OK) return rc;
  }
  int shift = 0;
  while ((1 << shift) < delta) ++shift;

  DArgs a;
  a.row = rl->row;
  a.rec = rl->rec;
  a.cw = reinterpret_cast<const int2v*>(rl->cw);
  a.hasout = rl->hasout;
  a.dist = g->dist;
  a.stamp = g->stamp;
  a.q0 = g->q[0];
  a.q1 = g->q[1];
  a.hq = g->hq;
  a.ctl = g->ctl;
  a.ph = reinterpret_cast<DPhase*>(rl->phase);
  a.cnt = g->cnt;
  const bool pull = settle != kSettlePush;
  a.rcw = pull ? reinterpret_cast<const int2v*>(rl->rcw) : nullptr;
  a.hasin = pull ? rl->hasin : nullptr;
  a.m = g->m;
  a.settle = settle;
  a.rlog = ctx->profiling ? ctx->rlog : nullptr;
  a.progress = nullptr;
  a.small_q = grid * SMDP_WG;
  a.n = g->n;
  a.shift = shift;
  a.bshift = band_shift(delta);
  a.split = p.split_threshold;
  a.chunk = p.chunk;
  hipStream_t s = ctx->stream;
  if (a.rlog) SMDP_HIP(hipMemsetAsync(a.rlog, 0, sizeof(unsigned long long) * 2 * kRoundLogCap, s));
  prof_reset(ctx);

  int32_t* phase_buf = rl->phase;
  uint32_t* bar = reinterpret_cast<uint32_t*>(phase_buf + 4 * kPhaseInts);
  int32_t* flags = phase_buf + 4 * kPhaseInts + kBarLines * 32;
  // The persistent launch needs every workgroup resident at once: its grid
  // is capped at what the device holds (occupancy x CUs).
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_delta_run), SMDP_WG,
                                                   0) != hipSuccess)
    per_cu = 0;
  const int pgrid = std::min(grid, per_cu * ctx->cus);
  bool persistent = p.persistent && per_cu >= 1;

  int launches = 0;
  int ph = 0;
  for (int attempt = 0;; ++attempt) {
    SMDP_HIP(hipEventRecord(ctx->ev_begin, s));
    hipLaunchKernelGGL(k_delta_init, dim3((unsigned)std::min<int64_t>(((int64_t)g->n + 255) / 256, 4096)), dim3(256),
                       0, s, g->dist, g->stamp, g->n, src, g->ctl, g->cnt, bar, flags);
    ++launches;
    ph = 0;
    if (!persistent) break;
    // all phases in launches of up to kRunPhases phases each
    int32_t* hf = ctx->pinned;
    bool gave_up = false;
    a.small_q = pgrid * SMDP_WG;
    for (;;) {
      SMDP_CHECK(prof_mark(ctx, 0));
      hipLaunchKernelGGL(k_delta_run, dim3(pgrid), dim3(SMDP_WG), 0, s, a, ph, ph + kRunPhases, bar, flags);
      SMDP_CHECK(prof_mark(ctx, 0));
      ++launches;
      SMDP_HIP(hipGetLastError());
      SMDP_HIP(hipMemcpyAsync(hf, flags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      SMDP_HIP(hipStreamSynchronize(s));
      if (hf[0]) {
        gave_up = true;
        break;
      }
      const int next = hf[1];
      if (next < ph + kRunPhases) {
        ph = next;
        break;
      }
      ph = next;
      if (ph >= p.max_rounds) return SMDP_ERR_OVERFLOW;
      hipLaunchKernelGGL(k_delta_bar_reset, dim3(1), dim3(256), 0, s, bar);
      ++launches;
    }
    if (!gave_up) break;
    // A barrier gave up: the workgroups were not all resident (another
    // process on the device).  Solve again with one launch per pass.
    if (attempt > 0) return SMDP_ERR_DEVICE;
    persistent = false;
    prof_reset(ctx);
    a.small_q = grid * SMDP_WG;
  }
  if (!persistent) {
    // One launch pair per phase.  The step kernel reports each phase it starts
    // in host memory, so the host keeps a few phases queued ahead of the
    // device (never letting the queue drain while it checks) and stops
    // enqueueing at the phase that finds the solve done; a blocking check per
    // batch left the device idle ~35 us each time, and batches sized blindly
    // ran up to 25 phases past the end.
    a.progress = ctx->progress_dev;
    volatile int32_t* prog = ctx->progress;
    *prog = 0;
    int enq = 0;
    auto enqueue = [&](int k) -> int {
      for (int j = 0; j < k; ++j, ++enq) {
        SMDP_CHECK(prof_mark(ctx, 0));
        hipLaunchKernelGGL(k_delta_step, dim3(grid), dim3(SMDP_WG), 0, s, a, enq);
        SMDP_CHECK(prof_mark(ctx, 0));
        SMDP_CHECK(prof_mark(ctx, 1));
        hipLaunchKernelGGL(k_delta_drain, dim3(grid), dim3(SMDP_WG), 0, s, a, enq);
        SMDP_CHECK(prof_mark(ctx, 1));
        launches += 2;
      }
      SMDP_HIP(hipGetLastError());
      return SMDP_OK;
    };
    SMDP_CHECK(enqueue(std::min(kFeedFirst, p.max_rounds)));
    int seen = 0;
    double t_seen = now_ms();
    for (;;) {
      const int v = *prog;
      if (v < 0) break;  // the kDone phase started: nothing more to enqueue
      if (v >= enq - kFeedSlack) {
        if (enq >= p.max_rounds) {
          // the cap is enqueued: done only if its last phase found the solve done
          SMDP_HIP(hipStreamSynchronize(s));
          if (*prog < 0) break;
          return SMDP_ERR_OVERFLOW;
        }
        SMDP_CHECK(enqueue(std::min(kFeed, p.max_rounds - enq)));
        continue;
      }
      if (v != seen) {
        seen = v;
        t_seen = now_ms();
      } else if (now_ms() - t_seen > 100.0) {
        // no progress report for a while: a fault stops the stream
        const hipError_t e = hipStreamQuery(s);
        if (e != hipSuccess && e != hipErrorNotReady) return hip_fail(e, "bucketed SSSP phases", __LINE__);
        t_seen = now_ms();
      }
    }
    ph = enq;
    // the done phase counted the reached arcs; the counters travel to the
    // host with the wait for the end event
    SMDP_HIP(hipMemcpyAsync(ctx->pinned, g->cnt, kCntParts * sizeof(Counters), hipMemcpyDeviceToHost, s));
    ctx->counters_staged = true;
    *reached_counted = true;
  }
  SMDP_HIP(hipEventRecord(ctx->ev_end, s));
  SMDP_HIP(hipEventSynchronize(ctx->ev_end));
  float ms = 0;
  SMDP_HIP(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
  if (a.rlog) {
    // one log line per phase, timed over its expand and drain launches
    const int k = std::min(ph, kRoundLogCap);
    std::vector<unsigned long long> h(2 * (size_t)k);
    if (k) SMDP_HIP(hipMemcpy(h.data(), ctx->rlog, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
    ctx->log_qlen.assign(k, 0);
    ctx->log_arcs.assign(k, 0);
    ctx->log_ms.assign(k, 0.0);
    for (int i = 0; i < k; ++i) {
      ctx->log_qlen[i] = (int64_t)h[2 * i];
      ctx->log_arcs[i] = (int64_t)h[2 * i + 1];
      if (persistent) continue;  // phases share one launch: no time per phase
      for (size_t e = 0; e < 4; e += 2) {
        const size_t at = (size_t)i * 4 + e;
        if (at + 1 >= ctx->prof_ev.size()) break;
        float t = 0;
        SMDP_HIP(hipEventElapsedTime(&t, ctx->prof_ev[at], ctx->prof_ev[at + 1]));
        ctx->log_ms[i] += t;
      }
    }
  }
  *ms_out = ms;
  *launches_out = launches;
  return SMDP_OK;
}

}  // namespace smdp
