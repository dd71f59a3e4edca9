// Library-internal structures: context, resident graph, control blocks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/smdp.h"

namespace smdp {

// Per-round control block.  Rounds use slots r&3: round r reads slot r&3,
// appends to slot (r+1)&3 and clears slot (r+2)&3, which no kernel of round r
// touches.  Padded to a cache line so slots do not share lines.
// Work is claimed from kClaimParts counters, one per 128-byte line: one
// counter serialises about 85M claims/s on MI355X, eight about 500M/s.
constexpr int kClaimParts = 8;
constexpr int kClaimStride = 32;  // ints between two counters

struct Ctl {
  int32_t qlen;        // frontier entries of this round
  int32_t hlen;        // grid-queue entries pushed during this round
  int32_t tile_next;   // dynamic tile counter (phase 1)
  int32_t chunk_next;  // dynamic grid-queue counter (phase 2)
  int32_t ticket;      // workgroups finished (last-workgroup reductions)
  int32_t pad[27];
  int32_t tiles[kClaimParts * kClaimStride];   // partitioned tile counters (phase 1)
  int32_t chunks[kClaimParts * kClaimStride];  // partitioned grid-queue counters (phase 2)
  // bucketed SSSP: the next bucket seen by this round's settle pass, stored
  // as INT32_MAX - bucket (0 = none) in kClaimParts copies (atomicMax)
  int32_t nbk[kClaimParts * kClaimStride];
  // bucketed SSSP: vertices the settle pass saw beyond its bucket (an upper
  // bound: a vertex lowered twice counts twice), in kClaimParts copies
  int32_t rem[kClaimParts * kClaimStride];
  int32_t sel[kClaimParts * kClaimStride];  // bucketed SSSP: vertices the settle pass expanded
  // bucketed SSSP: in- and out-degrees of the vertices a sweep or light pass
  // found newly below the bucket's end (the census that chooses between a
  // push and a pull settle), in kClaimParts copies
  uint32_t jin[kClaimParts * kClaimStride];
  uint32_t jout[kClaimParts * kClaimStride];
};
static_assert(sizeof(Ctl) % 128 == 0, "Ctl is made of 128-byte lines");
constexpr int kCtlInts = (int)(sizeof(Ctl) / sizeof(int32_t));

// Items [0, total) in kClaimParts equal ranges, one counter each (ctr[p *
// kClaimStride]).  A workgroup takes up to `grab` items at a time from range
// (blockIdx.x + k) % kClaimParts, starting at k = 0 (workgroups of one XCD
// share a counter first) and moving on when a range is exhausted; a plain read
// skips exhausted ranges without an atomic.  Called by one thread; returns
// false when every range is exhausted, else the items [*first, *end).
__device__ __forceinline__ bool claim_items(int32_t* ctr, int total, int grab, int& k, int* first, int* end) {
  for (; k < kClaimParts; ++k) {
    const int p = ((int)blockIdx.x + k) % kClaimParts;
    const int lo = (int)((int64_t)total * p / kClaimParts);
    const int size = (int)((int64_t)total * (p + 1) / kClaimParts) - lo;
    int32_t* c = ctr + p * kClaimStride;
    if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= size) continue;
    const int t = atomicAdd(c, grab);
    if (t < size) {
      *first = lo + t;
      *end = lo + min(t + grab, size);
      return true;
    }
  }
  return false;
}

// Zeroes control slot `slot` (block 0 only; every thread of it calls this).
__device__ __forceinline__ void ctl_clear(Ctl* ctl, int slot) {
  if (blockIdx.x != 0) return;
  int32_t* c = reinterpret_cast<int32_t*>(ctl + slot);
  for (int i = threadIdx.x; i < kCtlInts; i += blockDim.x) c[i] = 0;
}


// Work counters, accumulated with one atomic per wave.
struct Counters {
  unsigned long long edges;
  unsigned long long verts;
  unsigned long long push[4];
  unsigned long long rounds;   // max round index + 1 that saw work
  unsigned long long reached;  // out-arcs of reached vertices
  unsigned long long done;     // PageRank: iteration at which the stop rule fired (0 = not yet)
  unsigned long long edges2;   // arcs scanned by phase-2 (grid queue) kernels
  unsigned long long atom_try; // atomicMin attempts after the plain read said "improves"
  unsigned long long atom_win; // atomicMin calls that lowered the value
  unsigned long long skipped;  // bucketed SSSP: arcs read whose relaxation belongs to another pass
  unsigned long long pad[3];
};
static_assert(sizeof(Counters) == 128, "one 128-B line per counter copy");

// The statistics counters are kept in kCntParts copies on separate lines: a
// workgroup adds to copy blockIdx % kCntParts and the host combines them
// (sum_counters).  With one shared line, the end-of-kernel adds of every
// workgroup queued on it (about 22 ns per workgroup in a colouring round).
constexpr int kCntParts = 32;
__device__ __forceinline__ Counters* cnt_of(Counters* c) { return c + (blockIdx.x & (kCntParts - 1)); }
// Copies the kCntParts copies from the device (synchronously on stream s)
// and combines them: sums, and the maximum of rounds and done.
int read_counters(smdp_ctx* ctx, const Counters* dev, Counters* out);

// Entry of the grid queue: vertex/row and the first edge of its chunk.
struct Chunk {
  int32_t v;
  int32_t first;
};

}  // namespace smdp

struct smdp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  char name[256] = {0};
  char arch[64] = {0};
  smdp_params params[SMDP_APP_COUNT];
  int profiling = 0;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  std::vector<hipEvent_t> prof_ev;  // pairs of events around profiled launches
  std::vector<int> prof_kind;       // kernel kind of each pair
  int32_t* pinned = nullptr;        // small host readback buffer (hipHostMalloc)
  // Host memory kernels write progress into while they run (coherent, mapped):
  // progress is the host view, progress_dev the device address of the same words.
  int32_t* progress = nullptr;
  int32_t* progress_dev = nullptr;
  // set when a run enqueued the copy of its counters into `pinned` before its
  // end event, so the statistics need no copy and wait of their own
  bool counters_staged = false;
  int32_t* stage[2] = {nullptr, nullptr};  // pinned staging of large uploads (hipHostMalloc, lazily)
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  unsigned long long* rlog = nullptr;  // per-round (queue length, arcs) when profiling
  std::vector<int64_t> log_qlen, log_arcs;
  std::vector<double> log_ms;
  // device memory of the one-solve graphs of drop-in calls, kept and grown
  // as needed so that such a call allocates nothing once warm
  void* arena = nullptr;
  size_t arena_bytes = 0;
};

namespace smdp {
constexpr int kRoundLogCap = 4096;
constexpr int kPinnedInts = 2048;  // smdp_ctx::pinned: readback of small results and of the work counters

// Fixed working memory of smdp_sssp_device_noalloc (api.hip holds the array):
// control slots, counters, a flag, then stamp, two queues and the has-arcs
// bits for n <= SMDP_NOALLOC_MAX_N vertices, and the grid queue for
// n + m / 64 + 64 chunks with m <= SMDP_NOALLOC_MAX_M.
constexpr int64_t kNoAllocScratchBytes = int64_t(12) << 20;
// Sets up the array's device address for `device` (once; the device must be
// current) / returns it (nullptr before set-up).
int scratch_setup(int device);
void* scratch_ptr(int device);
}

struct smdp_graph {
  smdp_ctx* ctx = nullptr;
  int32_t n = 0, m = 0;
  int32_t* row = nullptr;
  int32_t* col = nullptr;
  int32_t* w = nullptr;
  // frontier state (SSSP / BFS / coloring)
  int32_t* dist = nullptr;   // dist / level / colour
  int32_t* stamp = nullptr;  // round label of the last enqueue
  int32_t* q[2] = {nullptr, nullptr};
  smdp::Chunk* hq = nullptr;
  int64_t hq_cap = 0;
  smdp::Ctl* ctl = nullptr;
  smdp::Counters* cnt = nullptr;
  // SpMV / PageRank
  float* val = nullptr;
  float* x = nullptr;
  float* y = nullptr;
  int32_t* outdeg = nullptr;
  float* rank[2] = {nullptr, nullptr};
  float* contrib = nullptr;
  double* partials = nullptr;  // per-tile FP64 partials
  int64_t partials_cap = 0;
  float* chunk_part = nullptr;  // grid-split row partials (sum, comp)
  int64_t chunk_part_cap = 0;
  int rank_cur = 0;             // which rank buffer holds the result
  // coloring
  uint32_t* prio = nullptr;
  uint32_t prio_version = 0;    // bumped at every priority upload (per-arc priority copies follow it)
  int32_t* resume = nullptr;
  int32_t* rmax = nullptr;
  // multi-source SSSP
  int32_t* mdist = nullptr;
  int64_t mdist_cap = 0;
  int32_t msrc_k = 0;
  void* msplan = nullptr;   // batched multi-source state and its captured graph (batch.hip)
  int last_res = -1;
  int val_version = 0;      // bumped by smdp_graph_set_values
  int outdeg_version = 0;   // bumped by smdp_graph_set_outdeg
  void* rowbufs = nullptr;  // row-reduce family buffers (rowreduce.hip)
  void* colorbufs = nullptr; // coloring buffers (color.hip)
  void* relabel = nullptr;   // renumbered copy for the tuned frontier (frontier.hip)
  bool dist_perm = false;    // dist holds the renumbered copy's result
  bool transient = false;    // created by a drop-in call for one solve
  bool arena = false;        // row, col, w, ctl, cnt, dist, stamp, q and hq live in the context's arena
  int8_t forest = -1;        // every vertex has at most one in-arc: 1 yes, 0 no, -1 not checked yet
  uint32_t* hasout = nullptr; // tuned frontier: bit v set when row v has arcs
  int32_t* tree_lo = nullptr; // tree: level bookkeeping and slots (tree.hip)
  int32_t* tree_inner = nullptr; // tree: per vertex, its arcs to vertices that have arcs (tree.hip)
};

namespace smdp {

// Record a HIP failure (message kept per thread) and return SMDP_ERR_DEVICE.
int hip_fail(hipError_t e, const char* what, int line);

#define SMDP_HIP(x)                                              \
  do {                                                           \
    hipError_t _e = (x);                                         \
    if (_e != hipSuccess) return smdp::hip_fail(_e, #x, __LINE__); \
  } while (0)

#define SMDP_CHECK(x)          \
  do {                         \
    int _rc = (x);             \
    if (_rc != SMDP_OK) return _rc; \
  } while (0)

// A temporary device buffer of a set-up step, freed on every return from it.
template <typename T>
struct DeviceScratch {
  T* p = nullptr;
  DeviceScratch() = default;
  DeviceScratch(const DeviceScratch&) = delete;
  DeviceScratch& operator=(const DeviceScratch&) = delete;
  ~DeviceScratch() {
    if (p) (void)hipFree(p);
  }
};

// Allocation helper: (re)allocates *p to hold `count` elements when too small.
template <typename T>
int ensure(T** p, int64_t* cap, int64_t count) {
  if (*p && *cap >= count) return SMDP_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  size_t bytes = sizeof(T) * (size_t)(count > 0 ? count : 1);
  if (hipMalloc((void**)p, bytes) != hipSuccess) {
    *p = nullptr;
    return SMDP_ERR_NOMEM;
  }
  *cap = count;
  return SMDP_OK;
}

template <typename T>
int alloc_once(T** p, int64_t count) {
  if (*p) return SMDP_OK;
  size_t bytes = sizeof(T) * (size_t)(count > 0 ? count : 1);
  if (hipMalloc((void**)p, bytes) != hipSuccess) {
    *p = nullptr;
    return SMDP_ERR_NOMEM;
  }
  return SMDP_OK;
}

const smdp_params& params_of(smdp_ctx* ctx, int app);
int grid_size(smdp_ctx* ctx, const smdp_params& p);
double now_ms();

// Profiling helpers around the dominant kernel of a call.
void prof_reset(smdp_ctx* ctx);
// Records one event; call before and after each launch of kernel `kind`
// (0: phase-1 / dominant kernel, 1: phase-2 kernel).
int prof_mark(smdp_ctx* ctx, int kind);
// Rounds to enqueue in the next batch of a host loop that reads the next
// round's queue length q after round r: batches double (up to cap) while the
// queue grows; once it shrinks, the next batch is the number of rounds that
// its measured decay per round needs to empty it (q * rho^j < 1), plus one,
// so few empty rounds are enqueued after the last one with work.
struct RoundBatcher {
  int batch;
  int cap;
  int32_t q_prev = 0;
  int r_prev = 0;
  RoundBatcher(int first, int cap_) : batch(std::max(1, first)), cap(cap_) {}
  void advance(int32_t q, int r) {
    int next = std::min(batch * 2, cap);
    if (r_prev > 0 && q > 0 && q < q_prev) {
      const double rho = std::pow((double)q / q_prev, 1.0 / (r - r_prev));
      const double need = std::log((double)q) / -std::log(rho);
      if (std::isfinite(need) && need >= 0.0) next = std::max(1, std::min(cap, (int)std::ceil(need) + 1));
    }
    q_prev = q;
    r_prev = r;
    batch = next;
  }
};

// Progress feed of the bucketed phases (delta.hip), the tuned rounds
// (frontier.hip) and PageRank's iterations (rowreduce.hip): the device reports
// each step it starts in mapped host memory; the host enqueues kFeedFirst
// steps before reading it, then kFeed more whenever at most kFeedSlack are
// still queued ahead of the device.  Steps enqueued after the last return at
// once but still cost a launch each: on RMAT-22 SSSP (first, feed, slack) =
// (8, 4, 3) left 4 empty phase pairs and took 2.083 ms a solve, (4, 2, 1)
// 2.062 ms.
#ifndef SMDP_FEED_FIRST
#define SMDP_FEED_FIRST 4
#endif
#ifndef SMDP_FEED
#define SMDP_FEED 2
#endif
#ifndef SMDP_FEED_SLACK
#define SMDP_FEED_SLACK 1
#endif
constexpr int kFeedFirst = SMDP_FEED_FIRST;
constexpr int kFeed = SMDP_FEED;
constexpr int kFeedSlack = SMDP_FEED_SLACK;

// Runs body(lo, hi) over [0, count) split into contiguous ranges, one per
// host thread (at most 16), each of at least `grain` items; with one range
// the calling thread runs it (starting threads costs tens of microseconds).
template <typename F>
void parallel_ranges(int64_t count, F body, int64_t grain = 256) {
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>({(int64_t)std::thread::hardware_concurrency(), 16,
                                                             count / std::max<int64_t>(1, grain)}));
  if (nt == 1) {
    body(0, count);
    return;
  }
  std::vector<std::thread> th;
  for (int64_t t = 0; t < nt; ++t) th.emplace_back([=] { body(count * t / nt, count * (t + 1) / nt); });
  for (auto& x : th) x.join();
}

// Renumbered copy of a resident graph for the tuned frontier solution
// (frontier.hip builds it; delta.hip partitions its rows).  Vertices are
// renumbered by falling in-degree; rows keep their arcs in order until a
// bucketed SSSP partitions each row into kBands weight bands of its light
// arcs (w < delta), lightest band first, then the heavy arcs, and gives each
// row a record of its band starts.
constexpr int kBands = 8;  // weight bands of a row's light arcs (bucketed SSSP)

struct Relabel {
  int32_t* row = nullptr;
  int32_t* col = nullptr;
  int32_t* w = nullptr;
  int32_t* newid = nullptr;          // caller id -> copy id (device)
  uint32_t* hasout = nullptr;        // bit v set when row v of the copy has arcs
  std::vector<int32_t> h_newid;      // the same on the host (source mapping)
  int32_t minw = 0;                  // smallest arc weight (bucketing needs minw >= 0)
  // bucketed SSSP: per row one kRec-int record (kRec* fields, delta.hip):
  // the row's start, its kBands band starts, its end, its in-row and its
  // degrees; and the rows' arcs as {target, weight} pairs in band order
  int32_t* rec = nullptr;
  int32_t* cw = nullptr;
  int32_t band_delta = 0;            // the bucket width rec and cw were built for (0: not built)
  int32_t* phase = nullptr;          // bucketed SSSP: 4 phase descriptors (delta.hip)
  // bucketed SSSP, pull settle: the copy's in-row offsets, its in-arcs as
  // {source, weight} pairs, and bit v set when v has in-arcs; built once per
  // resident graph
  int32_t* rrow = nullptr;
  int32_t* rcw = nullptr;
  uint32_t* hasin = nullptr;
  std::vector<int32_t> h_rrow;       // rrow on the host, kept from the renumbering until uploaded
  void* hctl = nullptr;              // bucketed SSSP: pinned host copy of one control slot
};

// Renumbered frontier copy (frontier.hip): freed with the graph; the
// un-permuting gather writes the first `count` results in caller numbering.
void relabel_free(smdp_graph* g);
int relabel_unpermute(smdp_graph* g, int32_t* out_dev, int64_t count);
// Sets g->forest (1 when no vertex has two in-arcs), checked once per graph.
int ensure_forest(smdp_graph* g);
// Builds *out (once): bit v of word v / 32 set when row v of `row` has arcs.
int build_hasout(smdp_graph* g, const int32_t* row, uint32_t** out);
// A one-solve graph for a drop-in call: as smdp_graph_create, with its CSR
// and frontier buffers in the context's arena (api.hip).
int graph_create_transient(smdp_ctx* ctx, const smdp_csr* h, smdp_graph** out);
int prof_collect(smdp_ctx* ctx, smdp_stats* st);

void rowbufs_free(smdp_graph* g);
void colorbufs_free(smdp_graph* g);
void msplan_free(smdp_graph* g);
// The last multi-source run's [k][n] results (device) and their count.
int msplan_result(smdp_graph* g, const int32_t** out, int64_t* count);

// Copies counters into st and finishes timings.
int fill_stats(smdp_graph* g, smdp_stats* st, double ms_kernel, uint64_t launches);

// Called by every run before it writes a device buffer: until the run
// succeeds and names its result kind, smdp_graph_read reports no result
// (dist, stamp and the queues are shared by the apps, so a failed or
// different run leaves nothing that can be read as a previous result).
inline void begin_run(smdp_graph* g) {
  g->last_res = -1;
  g->dist_perm = false;
}

}  // namespace smdp
