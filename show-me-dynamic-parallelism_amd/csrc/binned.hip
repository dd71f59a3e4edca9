// Binned PageRank iteration: the in-arcs are regrouped once per graph so that
// every iteration streams HBM instead of gathering from it.
//
// Reference: pagerank_thread_pull_kernel and its consolidated variants
// (apps/pagerank/src/pagerank_kernel.cu:72-601) gather rank[u] / outdeg[u]
// for every in-arc of a row.  Each gather is a random 4-byte read; on MI355X
// those run at about 80 G/s, far below what the same bytes cost when streamed.
//
// Here one iteration is two passes over a layout built once per graph
// (deterministic propagation blocking).  Sources are cut into chunks of
// 2^kSrcChunkLog ids, rows into bins of consecutive rows holding at most
// kBinCap arcs, and a segment is the run of arcs of one (chunk, bin) pair.
//  * scatter (k_bin_scatter): a work item is a range of bins of one chunk.
//    The workgroup stages the chunk's contributions in LDS, reads the source
//    index of every arc of those segments (stored chunk-major, so one
//    contiguous run) and writes the contribution to the arc's slot in the
//    bin-major value array, each segment a run of its own.
//  * reduce (k_bin_reduce): a work item is one bin.  Its values are one
//    contiguous run; the workgroup copies them into LDS at each arc's in-CSR
//    position, sums every row from LDS in arc order with the compensated sum
//    of device_common.h and runs the PageRank epilogue.  Rows of at least
//    kBinWide arcs have a bin of their own, cut into pieces; the last
//    workgroup merges each wide row's piece sums in order.
// Per arc and iteration the passes move 2 + 4 bytes (scatter) and 4 + 2 bytes
// (reduce).  Results do not depend on the schedule: every light row is summed
// in in-CSR arc order by one thread, every heavy row by one wave in a fixed
// lane order, and every partial is combined in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "device_common.h"
#include "rowreduce.h"
#include "smdp_internal.h"

namespace smdp {

constexpr int kScatNT = 1024;                 // threads per scatter workgroup (one per CU: ~146 KB of LDS)
constexpr int kBinNT = 512;                   // threads per reduce workgroup (two per CU: ~76 KB of LDS)
constexpr int kW = kBinNT / SMDP_WAVE;
constexpr int kSW = kScatNT / SMDP_WAVE;
constexpr int kSrcChunkLog = 15;
constexpr int kSrcChunk = 1 << kSrcChunkLog;  // sources per chunk: 128 KB of contributions in LDS
constexpr int kBinCap = 16380;                // arcs per bin: 64 KB of values in LDS; with a start
                                              // off a 4-arc boundary, 4096 groups of 4 still cover it
constexpr int kRowsPerThread = 3;
constexpr int kBinRows = kRowsPerThread * kBinNT;  // rows per bin
constexpr int kBinWide = kBinCap;             // rows with at least this many arcs get their own bin
constexpr int kRowWave = 64;                  // rows of a bin with at least this many arcs: one wave each
constexpr int kRowGroup = 8;                  // rows with [8, 64) arcs: 8 lanes each; shorter rows: one lane
constexpr int kScatterArcs = 262144;          // arcs per scatter work item (at most)
constexpr int kScatterBins = 2048;            // bins per scatter work item: their segment offsets sit in LDS
constexpr int kScatterBlock = 4 * SMDP_WAVE;  // arcs a wave loads with one 4-arc load per lane
constexpr int kScatterBlocks = kScatterArcs / kScatterBlock + 1;
constexpr int kWidePiece = 16384;             // arcs per piece of a wide row
constexpr int kQuads = 8;                     // 4-arc groups each thread loads at once
constexpr int kPad = 8;                       // entries past m in vals and dst: 4-arc loads may touch them
constexpr int kFront = 4;                     // entries before vals and dst: a bin without arcs that starts
                                              // at arc 0 loads the group of 4 just below it (never used)
constexpr int kSrcPad = kScatterBlock + 8;    // entries past m in src: a scatter block may read that far
constexpr int64_t kMaxSegments = int64_t(1) << 27;  // bound on nchunks x nbins (host plan memory)

static_assert(kBinCap <= 65535 && kSrcChunk <= 65536, "positions and offsets are stored in 16 bits");
static_assert(kBinWide <= kBinCap, "a row below kBinWide always fits in a bin");
static_assert(kSrcChunk % (4 * kScatNT) == 0, "contributions are loaded as float4, kSrcChunk/4/kScatNT each");

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));

// Workgroup g takes the scatter items of s_group[g] .. s_group[g+1]-1: a run
// of items of about equal arcs, mostly of one chunk, whose contributions are
// loaded into LDS once per chunk.
__global__ __launch_bounds__(kScatNT) void k_bin_scatter(RArgs a, BArgs b) {
  if (a.st->done) return;
  __shared__ float s_c[kSrcChunk];
  __shared__ int s_cm[kScatterBins + 1];          // segment starts relative to the item's first arc, plus its end
  __shared__ int s_bm[kScatterBins];              // segment starts in vals
  __shared__ uint16_t s_blk[kScatterBlocks + 1];  // segment holding the first arc of each block
  const int B = b.nbins;
  const int lane = lane_id();
  const int wave = threadIdx.x / SMDP_WAVE;
  int loaded = -1;
  for (int it = b.s_group[blockIdx.x]; it < b.s_group[blockIdx.x + 1]; ++it) {
    const int32_t* item = b.s_item + 5 * it;
    const int ch = item[0], b0 = item[1], b1 = item[2], first = item[3], last = item[4];
    const int nb = b1 - b0;
    // arcs are addressed from the 4-aligned position below the item's first
    const int base = first & ~3;
    const int32_t* cm = b.cm + (int64_t)ch * (B + 1) + b0;
    const int32_t* bm = b.bm + (int64_t)ch * B + b0;
    __syncthreads();  // the previous item's reads of LDS are done
    // segments clipped to the item's arcs [first, last): the first and the
    // last may be cut
    for (int i = threadIdx.x; i <= nb; i += kScatNT) s_cm[i] = min(max(cm[i], first), last) - base;
    for (int i = threadIdx.x; i < nb; i += kScatNT) s_bm[i] = bm[i] + max(first - cm[i], 0);
    if (ch != loaded) {
      // the chunk's contributions: kSrcChunk/4/kScatNT float4 loads per thread, all in flight
      const int lo = ch << kSrcChunkLog;
      const int cnt = min(kSrcChunk, a.n - lo);
      constexpr int Q = kSrcChunk / 4 / kScatNT;
      f32x4 c4[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int i = 4 * ((int)threadIdx.x + q * kScatNT);
        if (i + 3 < cnt) {
          c4[q] = *reinterpret_cast<const f32x4*>(a.x + lo + i);
        } else {
          for (int e = 0; e < 4; ++e) c4[q][e] = i + e < cnt ? a.x[lo + i + e] : 0.f;
        }
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) *reinterpret_cast<f32x4*>(s_c + 4 * ((int)threadIdx.x + q * kScatNT)) = c4[q];
      loaded = ch;
    }
    __syncthreads();
    const int lo_j = first - base;  // relative positions below lo_j belong to the item before
    const int tot = s_cm[nb];       // relative end
    const int nblk = (tot + kScatterBlock - 1) / kScatterBlock;
    for (int i = threadIdx.x; i < nb; i += kScatNT) {
      const int s0 = max(s_cm[i], lo_j), s1 = s_cm[i + 1];
      for (int k = (s0 + kScatterBlock - 1) / kScatterBlock; k * kScatterBlock < s1; ++k) s_blk[k] = (uint16_t)i;
    }
    if (threadIdx.x == 0) s_blk[0] = 0;  // block 0 may start below lo_j: its arcs there are skipped
    __syncthreads();
    // wave w takes blocks w, w + kSW, ...: kQuads of them at once, loads first.
    // Lane l takes arcs l, l + 64, l + 128, l + 192 of a block, so each load
    // reads 128 contiguous bytes of src and each store writes 64 consecutive
    // slots of vals when they lie in one segment (lanes taking 4 consecutive
    // arcs wrote at a 16-byte stride: 152 -> 106 us per RMAT-22 iteration).
    for (int k0 = wave; k0 < nblk; k0 += kQuads * kSW) {
      uint16_t sv[kQuads][4];
#pragma unroll
      for (int u = 0; u < kQuads; ++u) {
        const int k = min(k0 + u * kSW, nblk - 1);
#pragma unroll
        for (int e = 0; e < 4; ++e)
          sv[u][e] = __builtin_nontemporal_load(b.src + base + k * kScatterBlock + e * SMDP_WAVE + lane);
      }
#pragma unroll
      for (int u = 0; u < kQuads; ++u) {
        const int k = k0 + u * kSW;
        if (k >= nblk) continue;
        int c = s_blk[k];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int j = k * kScatterBlock + e * SMDP_WAVE + lane;
          if (j >= lo_j && j < tot) {
            while (s_cm[c + 1] <= j) ++c;
            b.vals[s_bm[c] + (j - s_cm[c])] = s_c[sv[u][e]];
          }
        }
      }
    }
  }
}

// Sum of v over the workgroup in a fixed order.
__device__ __forceinline__ double2 bin_sum2(double2 v, double2* lds) {
  v.x = wave_sum_dpp(v.x);
  v.y = wave_sum_dpp(v.y);
  if (lane_id() == 0) lds[threadIdx.x / SMDP_WAVE] = v;
  __syncthreads();
  double2 t = make_double2(0.0, 0.0);
#pragma unroll
  for (int w = 0; w < kW; ++w) {
    t.x += lds[w].x;
    t.y += lds[w].y;
  }
  __syncthreads();
  return t;
}

// Reduce items are claimed from a counter, one at a time (about 4300 claims per
// RMAT-22 iteration, well within one counter's rate; eight partitioned
// counters were measured 10% slower here).  A workgroup claims
// its next item when it starts one and reads that item's descriptor while it
// sums the rows of the current one, so neither latency is on its path.
__global__ __launch_bounds__(kBinNT, 2) void k_bin_reduce(RArgs a, BArgs b, int k) {
  if (a.st->done) return;
  const float base = iteration_base(a, k);
  const int lane = lane_id();
  const int wave = threadIdx.x / SMDP_WAVE;
  __shared__ float s_v[kBinCap];
  __shared__ uint16_t s_row[kBinRows + 2];  // row offsets in the bin
  __shared__ uint16_t s_perm[kBinRows];     // the bin's rows by falling degree
  __shared__ float s_res[kBinRows];         // row sums, by row
  __shared__ int s_last;
  __shared__ int s_next;
  __shared__ float2 s_red[kW];
  __shared__ double2 s_red2[kW];
  unsigned long long edges = 0, rows = 0, pushes = 0;
  const int nitems = b.nitems;
  if (threadIdx.x == 0) s_next = atomicAdd(&a.st->work_next, 1);
  __syncthreads();
  int it = s_next;
  __syncthreads();  // s_next is written again below
  int4 d0 = make_int4(0, 0, 0, 0), d1 = make_int4(0, 0, 0, 0);
  if (it < nitems) {
    d0 = b.r_item[2 * it];
    d1 = b.r_item[2 * it + 1];
  }
  while (it < nitems) {
    int nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(&a.st->work_next, 1);  // the next item, claimed now
    const int beg = d0.x, end = d0.y, r0 = d0.z, nr = d0.w;
    const int nwave_rows = d1.x, ngroup_rows = d1.y, part = d1.z;
    // the run [beg, end) is read as aligned groups of 4 arcs
    const int a0 = beg & ~3;
    const int nq = (end - a0 + 3) >> 2;
    if (nr < 0) {
      // ---- a piece of a wide row: one compensated sum over a contiguous run
      if (threadIdx.x == 0) s_next = nxt;
      __syncthreads();
      it = s_next;
      if (it < nitems) {
        d0 = b.r_item[2 * it];
        d1 = b.r_item[2 * it + 1];
      }
      CSum acc;
      for (int qb = 0; qb < nq; qb += kQuads * kBinNT) {
        f32x4 v[kQuads];
#pragma unroll
        for (int t = 0; t < kQuads; ++t) {
          const int qi = min(qb + t * kBinNT + (int)threadIdx.x, nq - 1);
          v[t] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(b.vals + a0) + qi);
        }
#pragma unroll
        for (int t = 0; t < kQuads; ++t) {
          const int qi = qb + t * kBinNT + (int)threadIdx.x;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int j = a0 + 4 * qi + e;
            if (qi < nq && j >= beg && j < end) acc.add(v[t][e]);
          }
        }
      }
      if (threadIdx.x == 0) edges += end - beg;
      const CSum h = wg_csum<kBinNT>(acc, s_red);
      if (threadIdx.x == 0) b.wpart[part] = make_float2(h.s, h.c);
      continue;
    }
    // ---- a bin of rows.  Every load the bin needs is issued before any of
    // them is used: row offsets, the degree order, old ranks and
    // out-degrees of this thread's rows, and the whole run of values and
    // positions (kQuads groups of 4 arcs per thread cover a full bin).
    constexpr int kRowLoads = (kBinRows + 1 + kBinNT - 1) / kBinNT;
    int roff[kRowLoads];
    uint16_t perm[kRowsPerThread];
    float rold[kRowsPerThread];
    int od[kRowsPerThread];
    // Loads go to clamped positions (a bin has at least one row; a bin without
    // arcs loads the group of 4 below its start, kFront entries of padding at
    // arc 0): a load under a per-lane condition made the compiler wait for it
    // before the next one, which held back the bin's value loads by three
    // round trips.  Values loaded past nr or outside [beg, end) are never used.
#pragma unroll
    for (int q = 0; q < kRowLoads; ++q) roff[q] = a.row[r0 + min((int)threadIdx.x + q * kBinNT, nr)];
#pragma unroll
    for (int q = 0; q < kRowsPerThread; ++q) {
      const int r = r0 + min((int)threadIdx.x + q * kBinNT, nr - 1);
      perm[q] = b.rperm[r];
      rold[q] = a.rold[r];
      od[q] = a.outdeg[r];
    }
    static_assert(kQuads * kBinNT * 4 >= kBinCap + 3, "one round of quads covers a bin");
    f32x4 v[kQuads];
    u16x4 dd[kQuads];
#pragma unroll
    for (int t = 0; t < kQuads; ++t) {
      const int qi = min(t * kBinNT + (int)threadIdx.x, nq - 1);
      v[t] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(b.vals + a0) + qi);
      dd[t] = __builtin_nontemporal_load(reinterpret_cast<const u16x4*>(b.dst + a0) + qi);
    }
#pragma unroll
    for (int q = 0; q < kRowLoads; ++q) {
      const int i = (int)threadIdx.x + q * kBinNT;
      if (i <= nr) s_row[i] = (uint16_t)(roff[q] - beg);
    }
#pragma unroll
    for (int q = 0; q < kRowsPerThread; ++q) {
      const int r = (int)threadIdx.x + q * kBinNT;
      if (r < nr) s_perm[r] = perm[q];
    }
#pragma unroll
    for (int t = 0; t < kQuads; ++t) {
      const int qi = t * kBinNT + (int)threadIdx.x;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = a0 + 4 * qi + e;
        if (qi < nq && j >= beg && j < end) s_v[dd[t][e]] = v[t][e];
      }
    }
    if (threadIdx.x == 0) s_next = nxt;
    __syncthreads();
    it = s_next;
    if (it < nitems) {
      d0 = b.r_item[2 * it];
      d1 = b.r_item[2 * it + 1];
    }
    // Row sums.  Rows in falling degree order: the first nwave_rows take a
    // wave each, the next ngroup_rows 8 lanes each, the rest one lane each,
    // so the lanes of one task have rows of about the same length.
    const int ngroup_tasks = (ngroup_rows + SMDP_WAVE / kRowGroup - 1) / (SMDP_WAVE / kRowGroup);
    const int nlane_rows = nr - nwave_rows - ngroup_rows;
    const int ntasks = nwave_rows + ngroup_tasks + (nlane_rows + SMDP_WAVE - 1) / SMDP_WAVE;
    pushes += threadIdx.x == 0 ? nwave_rows + ngroup_rows : 0;
    for (int task = wave; task < ntasks; task += kW) {
      if (task < nwave_rows) {
        const int r = s_perm[task];
        const int s = s_row[r], e = s_row[r + 1];
        CSum acc;
        int q = s + lane;
        for (; q + 3 * SMDP_WAVE < e; q += 4 * SMDP_WAVE) {
          const float x0 = s_v[q], x1 = s_v[q + SMDP_WAVE], x2 = s_v[q + 2 * SMDP_WAVE], x3 = s_v[q + 3 * SMDP_WAVE];
          acc.add(x0);
          acc.add(x1);
          acc.add(x2);
          acc.add(x3);
        }
        for (; q < e; q += SMDP_WAVE) acc.add(s_v[q]);
        acc = wave_csum_dpp(acc);
        if (lane == 0) s_res[r] = acc.value();
      } else if (task < nwave_rows + ngroup_tasks) {
        const int i = nwave_rows + (task - nwave_rows) * (SMDP_WAVE / kRowGroup) + lane / kRowGroup;
        const bool valid = i < nwave_rows + ngroup_rows;
        const int r = valid ? s_perm[i] : 0;
        const int s = valid ? s_row[r] : 0, e = valid ? s_row[r + 1] : 0;
        CSum acc;
        for (int q = s + (lane & (kRowGroup - 1)); q < e; q += kRowGroup) acc.add(s_v[q]);
        acc = group8_csum(acc);
        if (valid && (lane & (kRowGroup - 1)) == 4) s_res[r] = acc.value();
      } else {
        const int i = nwave_rows + ngroup_rows + (task - nwave_rows - ngroup_tasks) * SMDP_WAVE + lane;
        if (i < nr) {
          const int r = s_perm[i];
          const int s = s_row[r], e = s_row[r + 1];
          CSum acc;
          for (int q = s; q < e; ++q) acc.add(s_v[q]);
          s_res[r] = acc.value();
        }
      }
    }
    __syncthreads();
    double res = 0.0, dang = 0.0;
#pragma unroll
    for (int q = 0; q < kRowsPerThread; ++q) {
      const int r = (int)threadIdx.x + q * kBinNT;
      if (r >= nr) continue;
      edges += s_row[r + 1] - s_row[r];
      ++rows;
      pr_epilogue(a, r0 + r, s_res[r], base, rold[q], od[q], res, dang);
    }
    // bin_sum2 ends with a barrier: LDS is free for the next item afterwards
    const double2 t = bin_sum2(make_double2(res, dang), s_red2);
    if (threadIdx.x == 0) b.bpart[part] = t;
  }
  count_work(a.cnt, edges, rows, pushes);
  if (arrive_last(a.st, &s_last)) {
    double res = 0.0, dang = 0.0;
    for (int i = threadIdx.x; i < b.nwide; i += kBinNT) {
      CSum h;
      for (int p = b.wfirst[i]; p < b.wfirst[i + 1]; ++p) {
        CSum o;
        o.s = b.wpart[p].x;
        o.c = b.wpart[p].y;
        h.merge(o);
      }
      epilogue<true>(a, b.wrow[i], h.value(), base, res, dang);
    }
    if (threadIdx.x == 0 && b.nwide) atomicAdd(&cnt_of(a.cnt)->verts, (unsigned long long)b.nwide);
    finish_pass<true, kBinNT>(a, k, make_double2(res, dang), s_red2);
  }
}

// ---------------------------------------------------------------- host side

namespace {
struct BinnedPlan {
  int32_t* cm = nullptr;
  int32_t* bm = nullptr;
  uint16_t* src = nullptr;
  uint16_t* dst = nullptr;
  float* vals = nullptr;
  uint16_t* rperm = nullptr;
  int32_t* s_item = nullptr;
  int32_t* s_group = nullptr;
  int4* r_item = nullptr;
  int32_t* wfirst = nullptr;
  int32_t* wrow = nullptr;
  float2* wpart = nullptr;
  double2* bpart = nullptr;
  BArgs args{};
};

}  // namespace

void binned_free(void* plan) {
  BinnedPlan* p = static_cast<BinnedPlan*>(plan);
  if (!p) return;
  void* ptrs[] = {p->cm,      p->bm,     p->src,     p->dst,    p->vals,   p->rperm, p->s_item,
                  p->s_group, p->r_item, p->wfirst,  p->wrow,   p->wpart,  p->bpart};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete p;
}

template <typename T>
static int upload(T** dptr, const std::vector<T>& h) {
  SMDP_CHECK(alloc_once(dptr, std::max<int64_t>(1, (int64_t)h.size())));
  if (!h.empty()) SMDP_HIP(hipMemcpy(*dptr, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return SMDP_OK;
}

// Builds the layout once per graph (the graph's arcs never change), for the
// context's CU count.  Every step is deterministic: threads own disjoint bin
// ranges.
int binned_ensure(smdp_graph* g, void** slot, BArgs* out) {
  if (*slot) {
    *out = static_cast<BinnedPlan*>(*slot)->args;
    return SMDP_OK;
  }
  const int64_t n = g->n, m = g->m;
  if (n <= 0) return SMDP_ERR_ARG;
  std::vector<int32_t> row((size_t)n + 1), col((size_t)std::max<int64_t>(1, m));
  SMDP_HIP(hipMemcpy(row.data(), g->row, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
  if (m) SMDP_HIP(hipMemcpy(col.data(), g->col, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  const int64_t K = (n + kSrcChunk - 1) >> kSrcChunkLog;
  const int64_t cus = std::max(1, g->ctx->cus);

  // bins: consecutive rows, at most kBinCap arcs and kBinRows rows; a wide
  // row is a bin alone
  std::vector<int32_t> brow(1, 0);
  std::vector<uint8_t> wide;
  int64_t cur = 0;
  for (int64_t r = 0; r < n; ++r) {
    const int64_t d = row[r + 1] - row[r];
    if (d >= kBinWide) {
      if (brow.back() < r) {
        brow.push_back((int32_t)r);
        wide.push_back(0);
      }
      brow.push_back((int32_t)(r + 1));
      wide.push_back(1);
      cur = 0;
      continue;
    }
    if (brow.back() < r && (cur + d > kBinCap || r - brow.back() >= kBinRows)) {
      brow.push_back((int32_t)r);
      wide.push_back(0);
      cur = 0;
    }
    cur += d;
  }
  if (brow.back() < n) {
    brow.push_back((int32_t)n);
    wide.push_back(0);
  }
  const int64_t B = (int64_t)wide.size();
  if (K * (B + 1) > kMaxSegments) return SMDP_ERR_ARG;

  // arcs per (chunk, bin)
  std::vector<int32_t> cnt((size_t)(K * B), 0);
  parallel_ranges(B, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; ++b)
      for (int32_t e = row[brow[b]]; e < row[brow[b + 1]]; ++e) cnt[(size_t)((col[e] >> kSrcChunkLog) * B + b)]++;
  });
  // segment starts: chunk-major in src, bin-major in vals/dst (a bin's run
  // starts at its first in-CSR arc)
  std::vector<int32_t> cm((size_t)(K * (B + 1))), bm((size_t)std::max<int64_t>(1, K * B));
  int64_t pos = 0;
  for (int64_t c = 0; c < K; ++c) {
    for (int64_t b = 0; b < B; ++b) {
      cm[(size_t)(c * (B + 1) + b)] = (int32_t)pos;
      pos += cnt[(size_t)(c * B + b)];
    }
    cm[(size_t)(c * (B + 1) + B)] = (int32_t)pos;
  }
  for (int64_t b = 0; b < B; ++b) {
    int64_t at = row[brow[b]];
    for (int64_t c = 0; c < K; ++c) {
      bm[(size_t)(c * B + b)] = (int32_t)at;
      at += cnt[(size_t)(c * B + b)];
    }
  }
  std::vector<uint16_t> src((size_t)(m + kSrcPad), 0), dst((size_t)(kFront + m + kPad), 0);
  parallel_ranges(B, [&](int64_t b0, int64_t b1) {
    std::vector<int32_t> cs((size_t)K), bs((size_t)K);
    for (int64_t b = b0; b < b1; ++b) {
      for (int64_t c = 0; c < K; ++c) {
        cs[(size_t)c] = cm[(size_t)(c * (B + 1) + b)];
        bs[(size_t)c] = bm[(size_t)(c * B + b)];
      }
      const int32_t e0 = row[brow[b]];
      for (int32_t e = e0; e < row[brow[b + 1]]; ++e) {
        const int32_t u = col[e];
        const int32_t c = u >> kSrcChunkLog;
        src[cs[(size_t)c]++] = (uint16_t)(u & (kSrcChunk - 1));
        dst[(size_t)kFront + bs[(size_t)c]++] = wide[b] ? 0 : (uint16_t)(e - e0);
      }
    }
  });

  // rows of each bin by falling degree (ties by row), and the class sizes
  std::vector<uint16_t> rperm((size_t)(n + kPad), 0);
  std::vector<int32_t> bcls((size_t)(2 * B), 0);
  parallel_ranges(B, [&](int64_t b0, int64_t b1) {
    std::vector<int32_t> ord;
    for (int64_t b = b0; b < b1; ++b) {
      if (wide[b]) continue;
      const int32_t r0 = brow[b], nr = brow[b + 1] - r0;
      ord.resize((size_t)nr);
      for (int32_t i = 0; i < nr; ++i) ord[(size_t)i] = i;
      auto deg = [&](int32_t i) { return row[r0 + i + 1] - row[r0 + i]; };
      std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) { return deg(x) > deg(y); });
      int32_t nw = 0, ng = 0;
      for (int32_t i = 0; i < nr; ++i) {
        rperm[(size_t)(r0 + i)] = (uint16_t)ord[(size_t)i];
        const int32_t d = deg(i);
        nw += d >= kRowWave;
        ng += d >= kRowGroup && d < kRowWave;
      }
      bcls[(size_t)(2 * b)] = nw;
      bcls[(size_t)(2 * b + 1)] = ng;
    }
  });

  // scatter: one workgroup per CU, each an equal share of the chunk-major
  // arcs, cut into items at chunk ends and where an item would exceed
  // kScatterArcs arcs or kScatterBins bins (a segment may be cut)
  std::vector<int32_t> s_item, s_group(1, 0);
  {
    int64_t c = 0, b0 = 0;
    for (int64_t grp = 0; grp < cus; ++grp) {
      int64_t p = m * grp / cus;
      const int64_t pend = m * (grp + 1) / cus;
      while (p < pend) {
        while (cm[(size_t)(c * (B + 1) + B)] <= p) {  // the chunk holding p
          ++c;
          b0 = 0;
        }
        const int32_t* cmc = &cm[(size_t)(c * (B + 1))];
        if (cmc[b0] > p) b0 = 0;
        while (cmc[b0 + 1] <= p) ++b0;  // the segment holding p
        const int64_t limit = std::min<int64_t>({pend, (int64_t)cmc[B], p + kScatterArcs});
        int64_t b1 = b0 + 1;
        while (b1 < B && cmc[b1] < limit && b1 - b0 < kScatterBins) ++b1;
        const int64_t last = std::min<int64_t>(limit, cmc[b1]);
        s_item.insert(s_item.end(), {(int32_t)c, (int32_t)b0, (int32_t)b1, (int32_t)p, (int32_t)last});
        p = last;
      }
      s_group.push_back((int32_t)(s_item.size() / 5));
    }
  }
  // reduce: items in row order (bins, and the pieces of each wide row),
  // claimed by the workgroups one at a time
  std::vector<int4> r_item;
  std::vector<int32_t> wfirst(1, 0), wrow;
  int32_t nparts = 0, npieces = 0;
  for (int64_t b = 0; b < B; ++b) {
    const int32_t e0 = row[brow[b]], e1 = row[brow[b + 1]];
    if (wide[b]) {
      for (int64_t p = e0; p < e1; p += kWidePiece) {
        const int32_t pe = (int32_t)std::min<int64_t>(e1, p + kWidePiece);
        r_item.push_back(make_int4((int32_t)p, pe, brow[b], -1));
        r_item.push_back(make_int4(0, 0, npieces++, 0));
      }
      wfirst.push_back(npieces);
      wrow.push_back(brow[b]);
    } else {
      const int32_t nr = brow[b + 1] - brow[b];
      r_item.push_back(make_int4(e0, e1, brow[b], nr));
      r_item.push_back(make_int4(bcls[(size_t)(2 * b)], bcls[(size_t)(2 * b + 1)], nparts++, 0));
    }
  }

  BinnedPlan* p = new BinnedPlan();
  int rc = SMDP_OK;
  auto step = [&](int r) {
    if (rc == SMDP_OK) rc = r;
  };
  step(upload(&p->cm, cm));
  step(upload(&p->bm, bm));
  step(upload(&p->src, src));
  step(upload(&p->dst, dst));
  step(upload(&p->rperm, rperm));
  step(upload(&p->s_item, s_item));
  step(upload(&p->s_group, s_group));
  step(upload(&p->r_item, r_item));
  step(upload(&p->wfirst, wfirst));
  step(upload(&p->wrow, wrow));
  if (rc == SMDP_OK) step(alloc_once(&p->vals, kFront + m + kPad));
  if (rc == SMDP_OK) step(alloc_once(&p->wpart, std::max<int64_t>(1, npieces)));
  if (rc == SMDP_OK) step(alloc_once(&p->bpart, std::max<int64_t>(1, nparts)));
  if (rc != SMDP_OK) {
    binned_free(p);
    return rc;
  }
  BArgs& a = p->args;
  a.cm = p->cm;
  a.bm = p->bm;
  a.src = p->src;
  a.dst = p->dst + kFront;
  a.vals = p->vals + kFront;
  a.rperm = p->rperm;
  a.s_item = p->s_item;
  a.s_group = p->s_group;
  a.r_item = p->r_item;
  a.wfirst = p->wfirst;
  a.wrow = p->wrow;
  a.wpart = p->wpart;
  a.bpart = p->bpart;
  a.nchunks = (int32_t)K;
  a.nbins = (int32_t)B;
  a.nsgroups = (int32_t)(s_group.size() - 1);
  a.nitems = (int32_t)(r_item.size() / 2);
  a.nwide = (int32_t)wrow.size();
  a.nparts = nparts;
  *slot = p;
  *out = a;
  return SMDP_OK;
}

// One iteration: the scatter holds ~146 KB of LDS per workgroup (one per CU)
// and runs the plan's work runs; the reduce holds ~76 KB (two per CU).
void binned_pass(smdp_ctx* ctx, hipStream_t s, const RArgs& a, const BArgs& b, int k, int* launches) {
  prof_mark(ctx, 1);
  hipLaunchKernelGGL(k_bin_scatter, dim3(b.nsgroups), dim3(kScatNT), 0, s, a, b);
  prof_mark(ctx, 1);
  prof_mark(ctx, 0);
  hipLaunchKernelGGL(k_bin_reduce, dim3(2 * ctx->cus), dim3(kBinNT), 0, s, a, b, k);
  prof_mark(ctx, 0);
  *launches += 2;
}

}  // namespace smdp
