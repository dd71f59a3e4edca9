// Device pieces shared by the frontier kernels (frontier.hip: unordered
// rounds for SSSP and BFS; delta.hip: bucketed SSSP).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"
#include "smdp_internal.h"

namespace smdp {

constexpr int kOutCap = 512;  // per-wave LDS staging of next-queue entries

// Next-queue output of one wave: LDS staging + the global queue it flushes to.
struct Sink {
  WaveOut<kOutCap> wo;
  int32_t* q;
  int32_t* qlen;
  unsigned long long pushes;
  unsigned long long tries;  // atomicMin attempts (plain read said "improves")
  unsigned long long wins;   // atomicMin that lowered the value
  __device__ __forceinline__ void init(int32_t* lds_region, int32_t* queue, int32_t* counter,
                                       int32_t queue_cap = INT32_MAX) {
    wo.init(lds_region, queue_cap);
    q = queue;
    qlen = counter;
    pushes = tries = wins = 0;
  }
  // wo.push returns the wave's count; lane 0 keeps it, so a wave sum counts it once
  __device__ __forceinline__ void push(bool pred, int32_t v) {
    const int c = wo.push(pred, v, q, qlen);
    if (lane_id() == 0) pushes += c;
  }
  __device__ __forceinline__ void flush() { wo.flush(q, qlen); }
};

// Flushes the wave's queue staging and adds its work counters (one atomic per
// counter per wave, on the workgroup's counter copy).  Wave-uniform.
__device__ __forceinline__ void flush_counters(Counters* cnt, unsigned long long edges, unsigned long long verts,
                                               unsigned long long p0, unsigned long long p1,
                                               unsigned long long p2, unsigned long long p3,
                                               Sink& out) {
  out.flush();
  unsigned long long tries = wave_sum(out.tries), wins = wave_sum(out.wins);
  if (lane_id() == 0) {
    if (tries) atomicAdd(&cnt_of(cnt)->atom_try, tries);
    if (wins) atomicAdd(&cnt_of(cnt)->atom_win, wins);
  }
  edges = wave_sum(edges);
  verts = wave_sum(verts);
  p0 = wave_sum(p0);
  p1 = wave_sum(p1);
  p2 = wave_sum(p2);
  p3 = wave_sum(p3);
  if (lane_id() == 0) {
    if (edges) atomicAdd(&cnt_of(cnt)->edges, edges);
    if (verts) atomicAdd(&cnt_of(cnt)->verts, verts);
    if (p0) atomicAdd(&cnt_of(cnt)->push[0], p0);
    if (p1) atomicAdd(&cnt_of(cnt)->push[1], p1);
    if (p2) atomicAdd(&cnt_of(cnt)->push[2], p2);
    if (p3) atomicAdd(&cnt_of(cnt)->push[3], p3);
  }
}

// End of a pass for the whole workgroup: the waves' staged queue entries go
// out with one atomic for the workgroup (one per wave queued the workgroups'
// final flushes of a large pass on the counter one after another).  All
// threads; s_n needs SMDP_WAVES_PER_WG + 1 ints.
__device__ __forceinline__ void wg_flush(Sink& out, int* s_n) {
  constexpr int W = SMDP_WAVES_PER_WG;
  const int wave = threadIdx.x / SMDP_WAVE;
  if (lane_id() == 0) s_n[wave] = out.wo.n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int k = 0; k < W; ++k) {
      const int c = s_n[k];
      s_n[k] = tot;
      tot += c;
    }
    s_n[W] = tot ? atomicAdd(out.qlen, tot) : 0;
  }
  __syncthreads();
  const int base = s_n[W] + s_n[wave];
  for (int j = lane_id(); j < out.wo.n; j += SMDP_WAVE)
    if (base + j < out.wo.qcap) out.q[base + j] = out.wo.buf[j];
  out.wo.n = 0;
  __syncthreads();
}

// claim_items for a whole wave: lanes 0 .. kClaimParts-1 read the range
// counters at once (lane j: range (blockIdx.x + j) % kClaimParts), and the
// first range with items left is claimed with one atomic, so an exhausted
// grid costs one round of reads instead of one read per range in turn.
// `dead` (bit j: range of lane j found exhausted) starts at 0 and is kept by
// the caller between claims.  Every lane of the wave calls it with the same
// arguments and gets the same result.
__device__ __forceinline__ bool claim_items_wave(int32_t* ctr, int total, int grab, uint32_t& dead, int* first,
                                                 int* end) {
  const int lane = lane_id();
  const int p = ((int)blockIdx.x + lane) % kClaimParts;
  const int lo = (int)((int64_t)total * p / kClaimParts);
  const int size = (int)((int64_t)total * (p + 1) / kClaimParts) - lo;
  int32_t* c = ctr + p * kClaimStride;
  for (;;) {
    bool avail = false;
    if (lane < kClaimParts && !((dead >> lane) & 1u))
      avail = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < size;
    const uint64_t m = ballot(avail);
    dead |= (uint32_t)~m & ((1u << kClaimParts) - 1u);
    if (m == 0) return false;
    const int leader = first_lane(m);
    int t = 0;
    if (lane == leader) t = atomicAdd(c, grab);
    t = __shfl(t, leader);
    const int llo = __shfl(lo, leader), lsize = __shfl(size, leader);
    if (t < lsize) {
      *first = llo + t;
      *end = llo + min(t + grab, lsize);
      return true;
    }
    dead |= 1u << leader;  // another workgroup took the last items first
  }
}

// Word indices of the Counters fields, for wg_add_counters.
enum : int {
  kCntEdges = 0, kCntVerts = 1, kCntPush0 = 2, kCntPush3 = 5, kCntRounds = 6, kCntEdges2 = 9,
  kCntTries = 10, kCntWins = 11, kCntSkipped = 12
};

// Adds K per-thread counts to the workgroup's counter copy: one atomic per
// non-zero counter per workgroup.  All threads; s_red needs
// SMDP_WAVES_PER_WG * K words.
template <int K>
__device__ __forceinline__ void wg_add_counters(Counters* cnt, const unsigned long long (&v)[K], const int (&field)[K],
                                                unsigned long long* s_red) {
  const int wave = threadIdx.x / SMDP_WAVE;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const unsigned long long x = wave_sum(v[k]);
    if (lane_id() == 0) s_red[wave * K + k] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if ((int)threadIdx.x == k) {
      unsigned long long sum = 0;
      for (int w = 0; w < SMDP_WAVES_PER_WG; ++w) sum += s_red[w * K + k];
      if (sum) atomicAdd(reinterpret_cast<unsigned long long*>(cnt_of(cnt)) + field[k], sum);
    }
  }
  __syncthreads();
}

}  // namespace smdp
