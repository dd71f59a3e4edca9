// Counter-based random numbers for the seeded generators, and a restatement of
// glibc's rand() for the weights the reference's readers draw.
#pragma once
#include <stdint.h>
#include <math.h>

namespace smdp {

// SplitMix64 finaliser (Steele, Lea, Flood 2014; constants as published by
// Vigna).  Each draw is splitmix64(seed*2^40 + stream*2^32 + i), so a value
// depends only on (seed, stream, i) and never on thread count or order.
static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

static inline uint64_t rng_at(uint64_t seed, uint32_t stream, uint64_t i) {
  return splitmix64((seed << 40) + ((uint64_t)stream << 32) + i);
}

// Integer uniform in [lo, hi] by multiply-shift on the top 32 bits.
static inline int32_t uniform_int(uint64_t r, int32_t lo, int32_t hi) {
  uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
  return (int32_t)((int64_t)lo + (int64_t)(((r >> 32) * span) >> 32));
}

// Double uniform in [0, 1) with 53 random bits.
static inline double uniform01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Draw stream ids.
enum : uint32_t {
  STREAM_RMAT = 0,
  STREAM_WEIGHT = 1,
  STREAM_PERM = 2,
  STREAM_SOURCE = 3,
  STREAM_UNIFORM_DST = 4,
  STREAM_SMALL_PARAMS = 5,
  STREAM_PRIO = 7,
  STREAM_ZIPF = 8,
  STREAM_SMALL_BASE = 16,  // graph i of a small-graph batch uses 16 + i
};

// glibc's default rand(): the additive lagged-Fibonacci generator of
// random_r.c in its TYPE_3 configuration (degree 31, separation 3).  The seed
// fills the table with the Park-Miller LCG x' = 16807 x mod (2^31 - 1)
// (Schrage's factorisation 127773 * 16807 + 2836), the first 310 outputs are
// discarded, and each output is (r[i-31] + r[i-3]) mod 2^32 shifted right by
// one.  The reference seeds it with srand(7) before drawing edge weights
// (common/global.cpp:87, :164), so this reproduces its weights bit for bit
// without touching the process-wide rand() state.
class GlibcRand {
 public:
  explicit GlibcRand(uint32_t seed) { reseed(seed); }
  void reseed(uint32_t seed) {
    int32_t word = (int32_t)(seed == 0 ? 1 : seed);
    tab_[0] = word;
    for (int i = 1; i < 31; ++i) {
      int32_t hi = word / 127773, lo = word % 127773;
      word = 16807 * lo - 2836 * hi;
      if (word < 0) word += 2147483647;
      tab_[i] = word;
    }
    f_ = 3;
    r_ = 0;
    for (int i = 0; i < 310; ++i) next();
  }
  int32_t next() {
    tab_[f_] = (int32_t)((uint32_t)tab_[f_] + (uint32_t)tab_[r_]);
    int32_t out = (int32_t)(((uint32_t)tab_[f_] >> 1) & 0x7fffffff);
    if (++f_ == 31) f_ = 0;
    if (++r_ == 31) r_ = 0;
    return out;
  }

 private:
  int32_t tab_[31];
  int f_, r_;
};

}  // namespace smdp
