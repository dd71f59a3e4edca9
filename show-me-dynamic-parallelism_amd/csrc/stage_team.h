// Host-only: the thread team behind the library's pinned-memory uploads.
//
// `count` 4-byte entries of `src` go to the device in chunks of `chunk`
// entries through two staging buffers.  nt threads (the caller is thread 0)
// fill a buffer together, each its own slice, while the copy issued from the
// other buffer is in flight; all device calls are made by thread 0 alone,
// between two meetings of the whole team.  The copy engine is reached only
// through the two callables, so the team runs the same way against a
// simulated engine:
//   wait(b)            -> bool   the copy issued from buffer b has finished
//   issue(off, b, len) -> bool   start copying stage[b][0, len) to entry off
// With check_n >= 0 every value must lie in [0, check_n) (*bad is set
// otherwise), checked in the pass that fills the buffer.  Returns false when a
// callable failed; the team still meets to the end and the caller waits for
// the last copy.
#pragma once
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

namespace smdp {

template <typename Wait, typename Issue>
bool stage_team(int nt, const int32_t* src, int64_t count, int64_t chunk, int32_t* const stage[2], int64_t check_n,
                std::atomic<bool>* bad, Wait wait, Issue issue) {
  nt = std::max(1, nt);
  const int64_t nchunks = count > 0 ? (count + chunk - 1) / chunk : 0;
  std::atomic<int> arrived{0};
  std::atomic<int64_t> phase{0};
  std::atomic<bool> ok{true};
  // all nt threads meet; thread 0 runs `step` while the others wait
  auto meet = [&](int t, int64_t& ph, auto step) {
    if (t == 0) {
      while (arrived.load(std::memory_order_acquire) != nt - 1) std::this_thread::yield();
      arrived.store(0, std::memory_order_relaxed);
      step();
      phase.store(++ph, std::memory_order_release);
    } else {
      arrived.fetch_add(1, std::memory_order_acq_rel);
      while (phase.load(std::memory_order_acquire) == ph) std::this_thread::yield();
      ++ph;
    }
  };
  auto worker = [&](int t) {
    int64_t ph = 0;
    for (int64_t c = 0; c < nchunks; ++c) {
      const int64_t off = c * chunk, len = std::min(chunk, count - off);
      const int b = (int)(c & 1);
      int32_t* buf = stage[b];
      // the buffer is free once the copy issued from it two chunks ago is done
      meet(t, ph, [&] {
        if (c >= 2 && ok.load() && !wait(b)) ok = false;
      });
      const int64_t lo = len * t / nt, hi = len * (t + 1) / nt;
      if (hi > lo) {
        memcpy(buf + lo, src + off + lo, sizeof(int32_t) * (size_t)(hi - lo));
        if (check_n >= 0) {
          bool out_of_range = false;
          for (int64_t k = lo; k < hi; ++k) out_of_range |= (uint32_t)buf[k] >= (uint64_t)check_n;
          if (out_of_range) bad->store(true);
        }
      }
      meet(t, ph, [&] {
        if (ok.load() && !issue(off, b, len)) ok = false;
      });
    }
  };
  std::vector<std::thread> team;
  for (int t = 1; t < nt; ++t) team.emplace_back(worker, t);
  worker(0);
  for (auto& x : team) x.join();
  return ok.load();
}

}  // namespace smdp
