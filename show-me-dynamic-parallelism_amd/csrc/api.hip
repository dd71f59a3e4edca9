// Context, resident graphs, parameters, errors and timing helpers.
//
// Replaces the per-app prepare_gpu()/clean_gpu() pairs (e.g.
// apps/sssp/src/sssp_wrapper.cu:54-145): device buffers live in a resident
// graph and are reused across calls instead of being allocated, copied and
// freed (and the device reset) on every call.
#include <hip/hip_runtime.h>

#include <thread>

#include <atomic>

#include <chrono>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>

#include "smdp_internal.h"
#include "stage_team.h"

namespace smdp {

static thread_local std::string g_last_error;

int hip_fail(hipError_t e, const char* what, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) at line %d: %s", (int)e, hipGetErrorString(e), line, what);
  g_last_error = buf;
  return SMDP_ERR_DEVICE;
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

static smdp_params default_params(int app) {
  smdp_params p;
  memset(&p, 0, sizeof p);
  p.heavy_threshold = 64;  // THRESHOLD, apps/sssp/src/sssp_kernel.cu:7
  p.split_threshold = 2048;
  p.chunk = 2048;
  p.wg_per_cu = 8;
  p.batch_rounds = 4;
  p.max_rounds = 1 << 22;
  p.delta = 64;
  p.group_share = 100;
  switch (app) {
    case SMDP_APP_SPMV:
    case SMDP_APP_PAGERANK:
      p.split_threshold = 4096;
      p.chunk = 4096;
      break;
    case SMDP_APP_MSSSP:
      p.split_threshold = 512;
      p.chunk = 256;
      break;
    case SMDP_APP_BFS:
      // rounds enqueued before the host first checks the queue: a BFS of up
      // to 8 levels then needs one check (4M-node Zipf tree, 6 levels: 0.232
      // -> 0.204 ms); SSSP was the same with 4, 8 or 16
      p.batch_rounds = 8;
      break;
    default: break;
  }
  return p;
}

const smdp_params& params_of(smdp_ctx* ctx, int app) { return ctx->params[app]; }

int grid_size(smdp_ctx* ctx, const smdp_params& p) {
  int g = ctx->cus * (p.wg_per_cu > 0 ? p.wg_per_cu : 4);
  return g > 0 ? g : 1;
}

void prof_reset(smdp_ctx* ctx) {
  for (auto e : ctx->prof_ev) (void)hipEventDestroy(e);
  ctx->prof_ev.clear();
  ctx->prof_kind.clear();
}

int prof_mark(smdp_ctx* ctx, int kind) {
  if (!ctx->profiling) return SMDP_OK;
  hipEvent_t e;
  SMDP_HIP(hipEventCreate(&e));
  if (ctx->prof_ev.size() % 2 == 0) ctx->prof_kind.push_back(kind);
  ctx->prof_ev.push_back(e);
  SMDP_HIP(hipEventRecord(e, ctx->stream));
  return SMDP_OK;
}

int prof_collect(smdp_ctx* ctx, smdp_stats* st) {
  if (!ctx->profiling || !st) return SMDP_OK;
  double ms[2] = {0, 0};
  uint64_t n[2] = {0, 0};
  for (size_t i = 0; i + 1 < ctx->prof_ev.size(); i += 2) {
    float t = 0;
    SMDP_HIP(hipEventElapsedTime(&t, ctx->prof_ev[i], ctx->prof_ev[i + 1]));
    int k = ctx->prof_kind[i / 2] ? 1 : 0;
    ms[k] += t;
    ++n[k];
  }
  st->ms_dominant = ms[0];
  st->dominant_launches = n[0];
  st->ms_phase2 = ms[1];
  st->phase2_launches = n[1];
  prof_reset(ctx);
  return SMDP_OK;
}

// The copies go through the context's pinned buffer (allocated with the
// context), so reading counters allocates nothing.
int read_counters(smdp_ctx* ctx, const Counters* dev, Counters* out) {
  static_assert(kCntParts * sizeof(Counters) <= kPinnedInts * sizeof(int32_t), "pinned buffer holds the counters");
  const Counters* parts = reinterpret_cast<const Counters*>(ctx->pinned);
  if (!ctx->counters_staged) {
    SMDP_HIP(hipMemcpyAsync(ctx->pinned, dev, kCntParts * sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream));
    SMDP_HIP(hipStreamSynchronize(ctx->stream));
  }
  ctx->counters_staged = false;
  Counters c;
  memset(&c, 0, sizeof c);
  for (int k = 0; k < kCntParts; ++k) {
    const Counters& p = parts[k];
    c.edges += p.edges;
    c.verts += p.verts;
    for (int i = 0; i < 4; ++i) c.push[i] += p.push[i];
    c.rounds = std::max(c.rounds, p.rounds);
    c.reached += p.reached;
    c.done = std::max(c.done, p.done);
    c.edges2 += p.edges2;
    c.atom_try += p.atom_try;
    c.atom_win += p.atom_win;
    c.skipped += p.skipped;
  }
  *out = c;
  return SMDP_OK;
}

int fill_stats(smdp_graph* g, smdp_stats* st, double ms_kernel, uint64_t launches) {
  if (!st) return SMDP_OK;
  Counters c;
  SMDP_CHECK(read_counters(g->ctx, g->cnt, &c));
  st->ms_kernel = ms_kernel;
  st->edges_relaxed = c.edges;
  st->vertices_expanded = c.verts;
  for (int i = 0; i < 4; ++i) st->q_push[i] = c.push[i];
  st->rounds = c.rounds;
  st->m_reached = c.reached;
  st->edges_phase2 = c.edges2;
  st->atomics_tried = c.atom_try;
  st->atomics_won = c.atom_win;
  st->arcs_skipped = c.skipped;
  st->kernel_launches = launches;
  return SMDP_OK;
}

// Host-to-device copy of `count` int32 through two pinned staging buffers:
// one thread team fills one buffer (its slice each) while the other is in
// DMA.  Copies from pageable memory ran at ~28 GB/s on this machine, from
// pinned memory at ~57 GB/s.  With check_n >= 0 every value must lie in
// [0, check_n) (the CSR column check, made in the same pass so host memory
// is read once); *bad is set otherwise.  Returns after the last copy.
constexpr int64_t kStageInts = int64_t(8) << 20;  // 32 MB per staging buffer
constexpr int64_t kStagedMin = int64_t(1) << 22;  // 4-byte entries: smaller arrays go up in one plain copy
static int staged_upload(smdp_ctx* ctx, int32_t* dst, const int32_t* src, int64_t count, int64_t check_n,
                         std::atomic<bool>* bad) {
  for (int i = 0; i < 2; ++i) {
    if (!ctx->stage[i] && hipHostMalloc((void**)&ctx->stage[i], sizeof(int32_t) * kStageInts, hipHostMallocDefault) !=
                              hipSuccess) {
      ctx->stage[i] = nullptr;
      return SMDP_ERR_NOMEM;
    }
    if (!ctx->stage_ev[i]) SMDP_HIP(hipEventCreateWithFlags(&ctx->stage_ev[i], hipEventDisableTiming));
  }
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)std::thread::hardware_concurrency(), 16));
  const bool ok = stage_team(
      nt, src, count, kStageInts, ctx->stage, check_n, bad,
      [&](int b) { return hipEventSynchronize(ctx->stage_ev[b]) == hipSuccess; },
      [&](int64_t off, int b, int64_t len) {
        return hipMemcpyAsync(dst + off, ctx->stage[b], sizeof(int32_t) * (size_t)len, hipMemcpyHostToDevice,
                              ctx->stream) == hipSuccess &&
               hipEventRecord(ctx->stage_ev[b], ctx->stream) == hipSuccess;
      });
  SMDP_HIP(hipStreamSynchronize(ctx->stream));
  return ok ? SMDP_OK : SMDP_ERR_DEVICE;
}

// One array of 4-byte entries (values, x, out-degrees, priorities).
template <typename T>
static int upload4(smdp_ctx* ctx, T** d, const T* s, int64_t cnt) {
  static_assert(sizeof(T) == sizeof(int32_t), "4-byte entries");
  if (!s) return SMDP_ERR_ARG;
  SMDP_CHECK(alloc_once(d, cnt));
  if (cnt >= kStagedMin) {
    std::atomic<bool> unused{false};
    return staged_upload(ctx, reinterpret_cast<int32_t*>(*d), reinterpret_cast<const int32_t*>(s), cnt, -1, &unused);
  }
  if (cnt > 0) SMDP_HIP(hipMemcpy(*d, s, sizeof(T) * cnt, hipMemcpyHostToDevice));
  return SMDP_OK;
}

// Working memory of the entry points that allocate nothing
// (smdp_sssp_device_noalloc): one fixed array per GPU, held by every GPU the
// library runs on until the process ends.  In host code `scratch` only names
// the symbol; its device address on each GPU comes from hipGetSymbolAddress,
// taken once per device (with that device current) when the first context on
// it is created, and is the only handle host code uses.
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {nullptr};
static std::mutex g_scratch_mu;

int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return SMDP_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return SMDP_OK;
  void* ptr = nullptr;
  SMDP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));  // the device is current (smdp_create)
  if (!ptr) return SMDP_ERR_DEVICE;
  g_scratch_ptr[device] = ptr;
  return SMDP_OK;
}

void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace smdp

using namespace smdp;

extern "C" {

const char* smdp_last_hip_error(void) { return g_last_error.c_str(); }

const char* smdp_strerror(int code) {
  switch (code) {
    case SMDP_OK: return "ok";
    case SMDP_ERR_ARG: return "bad argument";
    case SMDP_ERR_NOMEM: return "out of memory";
    case SMDP_ERR_DEVICE: return "device or HIP runtime error";
    case SMDP_ERR_SOLUTION: return "unsupported solution";
    case SMDP_ERR_OVERFLOW: return "queue or round capacity exceeded";
    case SMDP_ERR_IO: return "input file missing or malformed";
    case SMDP_ERR_FORMAT: return "wrong code for dataset";
    case SMDP_ERR_NOT_BUILT: return "library built without device code";
    default: return "unknown error";
  }
}

int smdp_create(int device, smdp_ctx** out) {
  if (!out) return SMDP_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SMDP_ERR_DEVICE;
  if (device < 0 || device >= count) return SMDP_ERR_ARG;
  SMDP_HIP(hipSetDevice(device));
  smdp_ctx* c = new smdp_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return SMDP_ERR_DEVICE;
  }
  c->cus = prop.multiProcessorCount;
  strncpy(c->name, prop.name, sizeof c->name - 1);
  strncpy(c->arch, prop.gcnArchName, sizeof c->arch - 1);
  // some drivers report no marketing name: ask again, else name the
  // architecture and compute units
  if (!c->name[0] && hipDeviceGetName(c->name, (int)sizeof c->name - 1, device) != hipSuccess) c->name[0] = 0;
  if (!c->name[0]) {
    char arch[64];
    strncpy(arch, prop.gcnArchName, sizeof arch - 1);
    arch[sizeof arch - 1] = 0;
    if (char* colon = strchr(arch, ':')) *colon = 0;
    snprintf(c->name, sizeof c->name, "AMD %s, %d CUs", arch, prop.multiProcessorCount);
  }
  for (int a = 0; a < SMDP_APP_COUNT; ++a) c->params[a] = default_params(a);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev_begin) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess ||
      hipHostMalloc((void**)&c->pinned, kPinnedInts * sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->progress, 64 * sizeof(int32_t), hipHostMallocMapped | hipHostMallocCoherent) !=
          hipSuccess ||
      hipHostGetDevicePointer((void**)&c->progress_dev, c->progress, 0) != hipSuccess) {
    smdp_destroy(c);
    return SMDP_ERR_DEVICE;
  }
  const int rc = scratch_setup(device);
  if (rc != SMDP_OK) {
    smdp_destroy(c);
    return rc;
  }
  *out = c;
  return SMDP_OK;
}

void smdp_destroy(smdp_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  prof_reset(c);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->progress) (void)hipHostFree(c->progress);
  for (int i = 0; i < 2; ++i) {
    if (c->stage[i]) (void)hipHostFree(c->stage[i]);
    if (c->stage_ev[i]) (void)hipEventDestroy(c->stage_ev[i]);
  }
  if (c->rlog) (void)hipFree(c->rlog);
  if (c->arena) (void)hipFree(c->arena);
  if (c->ev_begin) (void)hipEventDestroy(c->ev_begin);
  if (c->ev_end) (void)hipEventDestroy(c->ev_end);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int smdp_get_params(smdp_ctx* ctx, int app, smdp_params* out) {
  if (!ctx || !out || app < 0 || app >= SMDP_APP_COUNT) return SMDP_ERR_ARG;
  *out = ctx->params[app];
  return SMDP_OK;
}

int smdp_set_params(smdp_ctx* ctx, int app, const smdp_params* p) {
  if (!ctx || !p || app < 0 || app >= SMDP_APP_COUNT) return SMDP_ERR_ARG;
  smdp_params d = default_params(app);
  smdp_params q = *p;
  if (q.heavy_threshold <= 0) q.heavy_threshold = d.heavy_threshold;
  if (q.split_threshold <= 0) q.split_threshold = d.split_threshold;
  if (q.chunk <= 0) q.chunk = d.chunk;
  if (q.wg_per_cu <= 0) q.wg_per_cu = d.wg_per_cu;
  if (q.batch_rounds <= 0) q.batch_rounds = d.batch_rounds;
  if (q.max_rounds <= 0) q.max_rounds = d.max_rounds;
  if (q.delta <= 0) q.delta = d.delta;
  if (q.group_share <= 0) q.group_share = d.group_share;
  if (q.group_share < 10 || q.group_share > 1000) return SMDP_ERR_ARG;
  if (q.persistent != 0 && q.persistent != 1) return SMDP_ERR_ARG;
  if (q.settle != SMDP_SETTLE_AUTO && q.settle != SMDP_SETTLE_PUSH && q.settle != SMDP_SETTLE_PULL)
    return SMDP_ERR_ARG;
  if (q.wg_per_cu > 64 || q.chunk < 64) return SMDP_ERR_ARG;
  if ((q.delta & (q.delta - 1)) != 0 || q.delta > (1 << 20)) return SMDP_ERR_ARG;  // a power of two
  if (q.layout != SMDP_LAYOUT_AUTO && q.layout != SMDP_LAYOUT_CSR && q.layout != SMDP_LAYOUT_BINNED &&
      q.layout != SMDP_LAYOUT_HOT && q.layout != SMDP_LAYOUT_BUCKETS && q.layout != SMDP_LAYOUT_XCD8)
    return SMDP_ERR_ARG;
  if (q.layout == SMDP_LAYOUT_BINNED && app != SMDP_APP_PAGERANK) return SMDP_ERR_ARG;
  if (q.layout == SMDP_LAYOUT_HOT && app != SMDP_APP_SSSP && app != SMDP_APP_BFS) return SMDP_ERR_ARG;
  if (q.layout == SMDP_LAYOUT_BUCKETS && app != SMDP_APP_SSSP) return SMDP_ERR_ARG;
  ctx->params[app] = q;
  return SMDP_OK;
}

int smdp_device_info(smdp_ctx* ctx, char* name, int name_len, char* arch, int arch_len, int32_t* cus) {
  if (!ctx) return SMDP_ERR_ARG;
  if (name && name_len > 0) {
    strncpy(name, ctx->name, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (arch && arch_len > 0) {
    strncpy(arch, ctx->arch, arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  if (cus) *cus = ctx->cus;
  return SMDP_OK;
}

int smdp_set_profiling(smdp_ctx* ctx, int on) {
  if (!ctx) return SMDP_ERR_ARG;
  SMDP_HIP(hipSetDevice(ctx->device));
  if (on && !ctx->rlog) SMDP_CHECK(alloc_once(&ctx->rlog, 2 * (int64_t)kRoundLogCap));
  ctx->profiling = on ? 1 : 0;
  prof_reset(ctx);
  return SMDP_OK;
}

int smdp_round_log(smdp_ctx* ctx, int64_t* qlen, int64_t* arcs, double* ms, int32_t cap) {
  if (!ctx || cap < 0) return SMDP_ERR_ARG;
  const int32_t k = std::min<int32_t>(cap, (int32_t)ctx->log_qlen.size());
  for (int32_t i = 0; i < k; ++i) {
    if (qlen) qlen[i] = ctx->log_qlen[i];
    if (arcs) arcs[i] = ctx->log_arcs[i];
    if (ms) ms[i] = i < (int32_t)ctx->log_ms.size() ? ctx->log_ms[i] : 0.0;
  }
  return k;
}

// Validates and uploads h into g (buffers already set in g are used as they
// are); on failure destroys g.
static int graph_fill(smdp_ctx* ctx, const smdp_csr* h, smdp_graph* g, smdp_graph** out);

int smdp_graph_create(smdp_ctx* ctx, const smdp_csr* h, smdp_graph** out) {
  if (!ctx || !h || !out) return SMDP_ERR_ARG;
  *out = nullptr;
  SMDP_HIP(hipSetDevice(ctx->device));
  return graph_fill(ctx, h, new smdp_graph(), out);
}


static int graph_fill(smdp_ctx* ctx, const smdp_csr* h, smdp_graph* g, smdp_graph** out) {
  g->ctx = ctx;
  if (h->n < 0 || h->m < 0 || !h->row || (h->m > 0 && !h->col) || h->row[h->n] != h->m || h->row[0] != 0) {
    smdp_graph_destroy(g);
    return SMDP_ERR_ARG;
  }
  // Every kernel indexes row[u..u+1] and dist[col[e]] without bounds checks,
  // so the CSR is validated here once (on up to all host threads, 1M entries
  // each at least: a check reads an entry in well under a nanosecond, and a
  // small graph's drop-in call would otherwise spend its time starting them).
  // (columns of large graphs are checked while they are staged for upload, below)
  {
    constexpr int64_t kCheckGrain = int64_t(1) << 20;
    std::atomic<bool> bad{false};
    parallel_ranges(
        h->n,
        [&](int64_t b, int64_t e) {
          for (int64_t v = b; v < e; ++v)
            if (h->row[v + 1] < h->row[v]) bad = true;
        },
        kCheckGrain);
    if (h->m < kStagedMin)
      parallel_ranges(
          h->m,
          [&](int64_t b, int64_t e) {
            bool local = false;
            for (int64_t k = b; k < e; ++k) local |= (uint32_t)h->col[k] >= (uint32_t)h->n;
            if (local) bad = true;
          },
          kCheckGrain);
    if (bad) {
      smdp_graph_destroy(g);
      return SMDP_ERR_ARG;
    }
  }
  g->n = h->n;
  g->m = h->m;
  int rc = SMDP_OK;
  std::atomic<bool> bad_col{false};
  auto up = [&](int32_t** d, const int32_t* s, int64_t cnt, int64_t check_n) -> int {
    int r = alloc_once(d, cnt);
    if (r != SMDP_OK) return r;
    if (cnt >= kStagedMin) return staged_upload(ctx, *d, s, cnt, check_n, &bad_col);
    if (cnt > 0 && hipMemcpy(*d, s, sizeof(int32_t) * cnt, hipMemcpyHostToDevice) != hipSuccess)
      return SMDP_ERR_DEVICE;
    return SMDP_OK;
  };
  rc = up(&g->row, h->row, (int64_t)h->n + 1, -1);
  if (rc == SMDP_OK) rc = up(&g->col, h->col, h->m, h->n);
  if (rc == SMDP_OK && bad_col) rc = SMDP_ERR_ARG;  // a column out of range (checked while staged)
  if (rc == SMDP_OK && h->w) rc = up(&g->w, h->w, h->m, -1);
  if (rc == SMDP_OK) rc = alloc_once(&g->ctl, 4);
  if (rc == SMDP_OK) rc = alloc_once(&g->cnt, kCntParts);
  if (rc != SMDP_OK) {
    smdp_graph_destroy(g);
    return rc;
  }
  *out = g;
  return SMDP_OK;
}

void smdp_graph_destroy(smdp_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->ctx->device);
  if (g->arena) {
    // these live in the context's arena
    g->row = g->col = g->w = g->dist = g->stamp = g->q[0] = g->q[1] = nullptr;
    g->ctl = nullptr;
    g->cnt = nullptr;
    g->hq = nullptr;
  }
  void* bufs[] = {g->row,    g->col,      g->w,       g->dist,     g->stamp,  g->hasout, g->q[0],  g->q[1],
                  g->hq,     g->ctl,      g->cnt,     g->val,      g->x,      g->y,     g->outdeg,
                  g->rank[0], g->rank[1], g->contrib, g->partials, g->chunk_part, g->prio, g->resume,
                  g->rmax,   g->mdist,    g->tree_lo, g->tree_inner};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  rowbufs_free(g);
  colorbufs_free(g);
  msplan_free(g);
  relabel_free(g);
  delete g;
}

int smdp_graph_set_values(smdp_graph* g, const float* val) {
  if (!g) return SMDP_ERR_ARG;
  SMDP_HIP(hipSetDevice(g->ctx->device));
  ++g->val_version;  // layouts derived from the values must be rebuilt
  return upload4(g->ctx, &g->val, val, g->m);
}

int smdp_graph_set_x(smdp_graph* g, const float* x) {
  if (!g) return SMDP_ERR_ARG;
  SMDP_HIP(hipSetDevice(g->ctx->device));
  return upload4(g->ctx, &g->x, x, g->n);
}

int smdp_graph_set_outdeg(smdp_graph* g, const int32_t* outdeg) {
  if (!g || !outdeg) return SMDP_ERR_ARG;
  SMDP_HIP(hipSetDevice(g->ctx->device));
  ++g->outdeg_version;  // the dangling count derived from them must be redone
  return upload4(g->ctx, &g->outdeg, outdeg, g->n);
}

int smdp_graph_set_prio(smdp_graph* g, const uint32_t* prio) {
  if (!g || !prio) return SMDP_ERR_ARG;
  SMDP_HIP(hipSetDevice(g->ctx->device));
  ++g->prio_version;  // the per-arc priority copies of the colouring must be redone
  return upload4(g->ctx, &g->prio, prio, g->n);
}

int smdp_graph_read(smdp_graph* g, int what, void* out, int64_t count) {
  if (!g || !out || count < 0) return SMDP_ERR_ARG;
  SMDP_HIP(hipSetDevice(g->ctx->device));
  const void* src = nullptr;
  int64_t avail = 0;
  size_t esz = 4;
  switch (what) {
    case SMDP_RES_DIST:
    case SMDP_RES_LEVEL:
    case SMDP_RES_COLOR:
      src = g->dist;
      avail = g->n;
      break;
    case SMDP_RES_Y:
      src = g->y;
      avail = g->n;
      break;
    case SMDP_RES_RANK:
      src = g->rank[g->rank_cur];
      avail = g->n;
      break;
    case SMDP_RES_HEIGHT:
    case SMDP_RES_DESC:
      if (g->last_res != SMDP_RES_HEIGHT) return SMDP_ERR_ARG;  // both come from the last tree run
      src = what == SMDP_RES_HEIGHT ? g->dist : g->stamp;
      avail = g->n;
      what = SMDP_RES_HEIGHT;
      break;
    case SMDP_RES_MDIST:
      // [k][n] of the last multi-source run, source-major
      if (g->last_res != SMDP_RES_MDIST) return SMDP_ERR_ARG;
      {
        const int32_t* r = nullptr;
        SMDP_CHECK(msplan_result(g, &r, &avail));
        src = r;
      }
      break;
    default: return SMDP_ERR_ARG;
  }
  // Only the result of the last run is defined (dist/level/colour share a buffer).
  if (what != g->last_res) return SMDP_ERR_ARG;
  if (!src || count > avail) return SMDP_ERR_ARG;
  if ((what == SMDP_RES_DIST || what == SMDP_RES_LEVEL) && g->dist_perm) {
    // the solve ran on the renumbered copy: back to the caller's numbering
    SMDP_CHECK(relabel_unpermute(g, g->q[1], count));
    src = g->q[1];
  }
  SMDP_HIP(hipMemcpyAsync(out, src, esz * (size_t)count, hipMemcpyDeviceToHost, g->ctx->stream));
  SMDP_HIP(hipStreamSynchronize(g->ctx->stream));
  return SMDP_OK;
}

}  // extern "C"

namespace smdp {
int graph_create_transient(smdp_ctx* ctx, const smdp_csr* h, smdp_graph** out) {
  if (!ctx || !h || !out || h->n < 0 || h->m < 0) return SMDP_ERR_ARG;
  *out = nullptr;
  SMDP_HIP(hipSetDevice(ctx->device));
  const int64_t n = std::max<int64_t>(1, h->n), m = std::max<int64_t>(1, h->m);
  // the grid queue holds the most any solution can push (chunks are at
  // least 64 arcs): its capacity never has to grow
  const int64_t hq_cap = n + m / 64 + 64;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t sizes[] = {al(4 * (size_t)(n + 1)),           al(4 * (size_t)m),  al(h->w ? 4 * (size_t)m : 0),
                          al(4 * sizeof(Ctl)),                al(kCntParts * sizeof(Counters)),
                          al(4 * (size_t)n),                  al(4 * (size_t)n),  al(4 * (size_t)n),
                          al(4 * (size_t)n),                  al(sizeof(Chunk) * (size_t)hq_cap)};
  size_t total = 0;
  for (size_t b : sizes) total += b;
  if (total > ctx->arena_bytes) {
    if (ctx->arena) (void)hipFree(ctx->arena);
    ctx->arena = nullptr;
    ctx->arena_bytes = 0;
    const size_t want = total + total / 4;
    if (hipMalloc(&ctx->arena, want) != hipSuccess) {
      ctx->arena = nullptr;
      return SMDP_ERR_NOMEM;
    }
    ctx->arena_bytes = want;
  }
  char* base = static_cast<char*>(ctx->arena);
  size_t off = 0;
  auto take = [&](size_t b) {
    void* p = base + off;
    off += b;
    return p;
  };
  smdp_graph* g = new smdp_graph();
  g->arena = true;
  g->transient = true;
  g->row = static_cast<int32_t*>(take(sizes[0]));
  g->col = static_cast<int32_t*>(take(sizes[1]));
  void* wp = take(sizes[2]);
  if (h->w) g->w = static_cast<int32_t*>(wp);
  g->ctl = static_cast<Ctl*>(take(sizes[3]));
  g->cnt = static_cast<Counters*>(take(sizes[4]));
  g->dist = static_cast<int32_t*>(take(sizes[5]));
  g->stamp = static_cast<int32_t*>(take(sizes[6]));
  g->q[0] = static_cast<int32_t*>(take(sizes[7]));
  g->q[1] = static_cast<int32_t*>(take(sizes[8]));
  g->hq = static_cast<Chunk*>(take(sizes[9]));
  g->hq_cap = hq_cap;
  return graph_fill(ctx, h, g, out);
}
}  // namespace smdp
