// Wave64 building blocks shared by every kernel family (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SMDP_WAVE 64
#define SMDP_WG 256
#define SMDP_WAVES_PER_WG (SMDP_WG / SMDP_WAVE)

namespace smdp {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Streaming load: CSR arrays are read once per pass, so they are loaded with
// the non-temporal hint and do not push the randomly gathered vectors (dist,
// x, PageRank contributions) out of L2.
template <typename T>
__device__ __forceinline__ T stream_load(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Number of set bits of `mask` below this lane (v_mbcnt_lo/hi).
__device__ __forceinline__ int mbcnt(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p ? 1 : 0); }

__device__ __forceinline__ int first_lane(uint64_t mask) { return __ffsll((unsigned long long)mask) - 1; }

// Wave-aggregated append: every active lane with `pred` gets a distinct slot
// in q[], at the cost of one atomic per wave.  Must be reached by all active
// lanes of the wave together.  Returns the number of items the wave appended.
__device__ __forceinline__ int wave_append(int32_t* __restrict__ q, int32_t* counter, int32_t item,
                                           bool pred) {
  uint64_t m = ballot(pred);
  if (m == 0) return 0;
  int leader = first_lane(m);
  int cnt = __popcll(m);
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, cnt);
  base = __shfl(base, leader);
  if (pred) q[base + mbcnt(m)] = item;
  return cnt;
}

// Per-wave output staging in LDS.  Items are appended with ballot + mbcnt
// (no atomics) and written to the global queue in coalesced runs with one
// global atomic per flush.  `n` lives in a register and must stay the same in
// every lane, so push() and flush() must be reached by all 64 lanes together
// (wave-uniform control flow); callers keep per-lane work predicated.
template <int CAP>
struct WaveOut {
  int32_t* buf;  // this wave's CAP-entry LDS region
  int n;
  int32_t qcap;  // entries the global queue holds: later ones are counted, not stored
  bool coherent;  // entries stored device-coherent (visible to other XCDs without an L2 write-back)

  __device__ __forceinline__ void init(int32_t* region, int32_t queue_cap = INT32_MAX, bool coh = false) {
    buf = region;
    n = 0;
    qcap = queue_cap;
    coherent = coh;
  }

  __device__ __forceinline__ void flush(int32_t* __restrict__ q, int32_t* counter) {
    if (n == 0) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint64_t act = ballot(true);
    const int leader = first_lane(act);
    const int me = mbcnt(act);
    const int stride = __popcll(act);
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, n);
    base = __shfl(base, leader);
    if (coherent) {
      for (int j = me; j < n; j += stride)
        if (base + j < qcap) __hip_atomic_store(q + base + j, buf[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      for (int j = me; j < n; j += stride)
        if (base + j < qcap) q[base + j] = buf[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    n = 0;
  }

  // Returns the number of items the wave appended.
  __device__ __forceinline__ int push(bool pred, int32_t item, int32_t* __restrict__ q, int32_t* counter) {
    const uint64_t m = ballot(pred);
    if (m == 0) return 0;
    if (pred) buf[n + mbcnt(m)] = item;
    const int c = __popcll(m);
    n += c;
    if (n > CAP - SMDP_WAVE) flush(q, counter);
    return c;
  }
};

// Owners of K edge slots in a tile: for each j, the last index i of the
// SMDP_WG-entry non-decreasing array off[] with off[i] <= j (off[0] == 0).
// A fixed 8-step search, interleaved over the K slots so their LDS reads
// overlap instead of forming one dependent chain each.
template <int K, int N = SMDP_WG>
__device__ __forceinline__ void tile_owners(const int* off, const int (&j)[K], int (&own)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) own[k] = 0;
#pragma unroll
  for (int step = N / 2; step > 0; step >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (off[own[k] + step] <= j[k]) own[k] += step;
  }
}

// Inclusive prefix sum across the active wave (all 64 lanes must be active).
__device__ __forceinline__ int wave_inclusive_scan(int x) {
  const int l = lane_id();
#pragma unroll
  for (int off = 1; off < SMDP_WAVE; off <<= 1) {
    int y = __shfl_up(x, off);
    if (l >= off) x += y;
  }
  return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int off = SMDP_WAVE / 2; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int off = SMDP_WAVE / 2; off > 0; off >>= 1) x = max(x, __shfl_xor(x, off));
  return x;
}

// Exclusive scan over a 256-thread workgroup; returns the exclusive prefix and
// writes the total to *total.  `lds` needs SMDP_WAVES_PER_WG ints.  Contains
// barriers: every thread of the workgroup must call it.
__device__ __forceinline__ int wg_exclusive_scan(int x, int* lds, int* total) {
  const int wave = threadIdx.x / SMDP_WAVE;
  int inc = wave_inclusive_scan(x);
  if (lane_id() == SMDP_WAVE - 1) lds[wave] = inc;
  __syncthreads();
  int before = 0, sum = 0;
#pragma unroll
  for (int k = 0; k < SMDP_WAVES_PER_WG; ++k) {
    int v = lds[k];
    before += (k < wave) ? v : 0;
    sum += v;
  }
  __syncthreads();
  *total = sum;
  return before + inc - x;
}

// Exclusive scan over a workgroup of NT threads (NT a multiple of 64); `lds`
// needs NT/64 ints.  Contains barriers: every thread must call it.
template <int NT>
__device__ __forceinline__ int wg_exclusive_scan_n(int x, int* lds, int* total) {
  constexpr int W = NT / SMDP_WAVE;
  const int wave = threadIdx.x / SMDP_WAVE;
  int inc = wave_inclusive_scan(x);
  if (lane_id() == SMDP_WAVE - 1) lds[wave] = inc;
  __syncthreads();
  int before = 0, sum = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    int v = lds[k];
    before += (k < wave) ? v : 0;
    sum += v;
  }
  __syncthreads();
  *total = sum;
  return before + inc - x;
}

// ---- compensated FP32 arithmetic --------------------------------------------
// These error-free transformations need every operation rounded on its own:
// contraction into fma (HIP's default -ffp-contract=fast) would break them.
// TwoSum (Knuth): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
#pragma clang fp contract(off)
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// A running sum with a second word that collects the rounding errors
// (Ogita-Rump-Oishi Sum2 / Dot2).  The error word absorbs the exact error of
// every addition and, via fma, of every product.
struct CSum {
  float s, c;
  __device__ __forceinline__ CSum() : s(0.f), c(0.f) {}
  __device__ __forceinline__ void add(float x) {
#pragma clang fp contract(off)
    float t, e;
    two_sum(s, x, t, e);
    s = t;
    c += e;
  }
  // Adds the exact product a*b: p = fl(a*b), fma(a,b,-p) is its rounding error.
  __device__ __forceinline__ void add_prod(float a, float b) {
#pragma clang fp contract(off)
    float p = a * b;
    float pe = __builtin_fmaf(a, b, -p);
    add(p);
    c += pe;
  }
  __device__ __forceinline__ void merge(const CSum& o) {
#pragma clang fp contract(off)
    add(o.s);
    c += o.c;
  }
  __device__ __forceinline__ float value() const {
#pragma clang fp contract(off)
    return s + c;
  }
};

__device__ __forceinline__ CSum wave_csum(CSum x) {
#pragma unroll
  for (int off = SMDP_WAVE / 2; off > 0; off >>= 1) {
    CSum o;
    o.s = __shfl_xor(x.s, off);
    o.c = __shfl_xor(x.c, off);
    x.merge(o);
  }
  return x;
}

// ---- DPP reductions of compensated sums -----------------------------------
// The shuffles move (sum, error) pairs with DPP or readlane, so no LDS is
// touched.  CSum::merge is not commutative in the last bit of the error
// term, so lanes of a group may end with slightly different values; every
// result is read from one fixed lane, which makes it a fixed function of the
// inputs.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ CSum dpp_merge(CSum x) {
  CSum o;
  o.s = dpp_f<CTRL>(x.s);
  o.c = dpp_f<CTRL>(x.c);
  x.merge(o);
  return x;
}
constexpr int kDppXor1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppShr4 = 0x114;  // row_shr:4
constexpr int kDppShr8 = 0x118;  // row_shr:8

// Sum of each aligned group of 8 lanes, read from lane 4 of the group.
__device__ __forceinline__ CSum group8_csum(CSum x) {
  x = dpp_merge<kDppXor1>(x);
  x = dpp_merge<kDppXor2>(x);
  return dpp_merge<kDppShr4>(x);
}

// Sum over the wave; valid in every lane.
__device__ __forceinline__ CSum wave_csum_dpp(CSum x) {
  x = group8_csum(x);
  x = dpp_merge<kDppShr8>(x);  // lane 15 of each row of 16 holds the row's sum
  CSum t;
#pragma unroll
  for (int r = 0; r < SMDP_WAVE / 16; ++r) {
    CSum o;
    o.s = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x.s), 16 * r + 15));
    o.c = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x.c), 16 * r + 15));
    t.merge(o);
  }
  return t;
}

// FP64 sum over the wave with DPP steps; every lane returns the same value,
// formed from the row sums in lanes 15, 31, 47, 63 in that order.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ double wave_sum_dpp(double x) {
  x += dpp_d<kDppXor1>(x);
  x += dpp_d<kDppXor2>(x);
  x += dpp_d<kDppShr4>(x);
  x += dpp_d<kDppShr8>(x);
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  double t = 0.0;
#pragma unroll
  for (int r = 0; r < SMDP_WAVE / 16; ++r) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 16 * r + 15);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 16 * r + 15);
    t += __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
  }
  return t;
}

}  // namespace smdp
