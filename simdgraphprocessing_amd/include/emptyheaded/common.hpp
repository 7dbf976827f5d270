// common.hpp -- the pieces of src/common.hpp the hot path's callers use:
// the set type tags (src/common.hpp:171-181), the wall-clock timers
// (src/common.hpp:87-99), the per-thread padding constant, and par_for_range
// (src/common.hpp:213-309) for host-side loops in user code.
#pragma once

#include <atomic>
#include <chrono>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "eh.h"

#define PADDING 300

namespace common {

enum type : uint8_t {
  BITSET = EH_BITSET,
  PSHORT = EH_PSHORT,
  UINTEGER = EH_UINTEGER,
  BITPACKED = 3,
  VARIANT = 4,
  HYBRID = EH_HYBRID,
  BITSET_NEW = EH_BITSET_NEW,
  NEW_TYPE = EH_NEW_TYPE
};

enum graph_type { DIRECTED, UNDIRECTED };

inline double startClock() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

inline double stopClock(double t_in) { return startClock() - t_in; }

inline double stopClock(const std::string &in, double t_in) {
  const double t = startClock() - t_in;
  std::cout << "Time[" + in + "]: " << t << " s" << std::endl;
  return t;
}

// The reference reserves per-thread intersection scratch here; device kernels
// stage their sets in LDS, so nothing is reserved.
inline void alloc_scratch_space(size_t, size_t) {}

// Dynamic chunks of `block` indices over [from, to) on num_threads host threads:
// setup(tid) once per thread, body(tid, i) per index, finish(tid) once per thread.
inline void par_for_range(size_t num_threads, size_t from, size_t to, size_t block,
                          const std::function<void(size_t)> &setup,
                          const std::function<void(size_t, size_t)> &body,
                          const std::function<void(size_t)> &finish) {
  if (num_threads == 0) num_threads = 1;
  std::atomic<size_t> next(from);
  std::vector<std::thread> pool;
  for (size_t t = 0; t < num_threads; t++) setup(t);
  for (size_t t = 0; t < num_threads; t++)
    pool.emplace_back([&, t] {
      for (;;) {
        const size_t s = next.fetch_add(block, std::memory_order_relaxed);
        if (s >= to) break;
        const size_t e = s + block < to ? s + block : to;
        for (size_t i = s; i < e; i++) body(t, i);
      }
    });
  for (auto &th : pool) th.join();
  for (size_t t = 0; t < num_threads; t++) finish(t);
}

}  // namespace common
