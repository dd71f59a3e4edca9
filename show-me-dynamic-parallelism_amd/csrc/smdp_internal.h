// Library-internal structures: context, resident graph, control blocks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/smdp.h"

namespace smdp {

// Per-round control block.  Rounds use slots r&3: round r reads slot r&3,
// appends to slot (r+1)&3 and clears slot (r+2)&3, which no kernel of round r
// touches.  Padded to a cache line so slots do not share lines.
struct Ctl {
  int32_t qlen;        // frontier entries of this round
  int32_t hlen;        // grid-queue entries pushed during this round
  int32_t tile_next;   // dynamic tile counter (phase 1)
  int32_t chunk_next;  // dynamic grid-queue counter (phase 2)
  int32_t ticket;      // workgroups finished (last-workgroup reductions)
  int32_t pad[27];
};
static_assert(sizeof(Ctl) == 128, "Ctl must be one 128-byte line");

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
  unsigned long long pad[4];
};

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
  unsigned long long* rlog = nullptr;  // per-round (queue length, arcs) when profiling
  std::vector<int64_t> log_qlen, log_arcs;
  std::vector<double> log_ms;
};

namespace smdp {
constexpr int kRoundLogCap = 4096;
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
  int32_t* resume = nullptr;
  int32_t* rmax = nullptr;
  // multi-source SSSP
  int32_t* mdist = nullptr;
  int64_t mdist_cap = 0;
  int32_t msrc_k = 0;
  int32_t msrc_groups = 0;
  int last_res = -1;
  int val_version = 0;      // bumped by smdp_graph_set_values
  void* rowbufs = nullptr;  // row-reduce family buffers (rowreduce.hip)
  void* colorbufs = nullptr; // coloring buffers (color.hip)
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
int prof_collect(smdp_ctx* ctx, smdp_stats* st);

void rowbufs_free(smdp_graph* g);
void colorbufs_free(smdp_graph* g);

// Copies counters into st and finishes timings.
int fill_stats(smdp_graph* g, smdp_stats* st, double ms_kernel, uint64_t launches);

}  // namespace smdp
