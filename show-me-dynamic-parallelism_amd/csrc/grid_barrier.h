// Grid barrier for kernels whose workgroups are all resident at once.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smdp {

constexpr int kBarLines = 10;  // 8 arrival lines, a total, a release; 32 words each

// Workgroup b arrives on line b % 8 (about one eighth of the workgroups per
// line: arrivals on one address run one at a time); the last arrival of a
// line arrives on line 8, and the last of those publishes the epoch on line
// 9, which everyone polls.  Counts only grow (epoch e is complete at e
// arrivals per member), and the lines are zeroed before each launch.  Gives
// up after about a second of polling: returns false in every workgroup that
// gave up, so a launch whose workgroups are not all resident ends with an
// error instead of waiting for ever.
constexpr int kBarSpinLimit = 1 << 21;
__device__ __forceinline__ bool grid_barrier(uint32_t* bar, uint32_t epoch, int* s_ok) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t nwg = gridDim.x;
    const uint32_t grp = blockIdx.x & 7;
    const uint32_t members = (nwg - grp + 7) / 8;  // workgroups g < nwg with g % 8 == grp
    const uint32_t ngroups = nwg < 8 ? nwg : 8;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this workgroup's writes (made before the __syncthreads)
    if (atomicAdd(&bar[32 * grp], 1u) + 1 == epoch * members &&
        atomicAdd(&bar[32 * 8], 1u) + 1 == epoch * ngroups)
      __hip_atomic_store(&bar[32 * 9], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    // relaxed polls (an acquire load invalidates the caches at every poll,
    // which slowed every workgroup of the XCD), one acquire fence after
    int ok = 1;
    for (int spins = 0; __hip_atomic_load(&bar[32 * 9], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch;) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > kBarSpinLimit) {
        ok = 0;
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // everyone's writes, for the whole workgroup
    *s_ok = ok;
  }
  __syncthreads();
  return *s_ok != 0;
}

}  // namespace smdp
