// common.hpp -- the pieces of src/common.hpp the hot path's callers use:
// the set type tags (src/common.hpp:171-181), the wall-clock timers
// (src/common.hpp:87-99), the per-thread padding constant, the std names the
// reference makes global (src/common.hpp:67-68), and par_for_range
// (src/common.hpp:213-309), which runs an app's loop in rounds (deferred.hpp).
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

#include "deferred.hpp"
#include "eh.h"

#define PADDING 300

// the reference's apps use cout, endl, string ... unqualified (src/common.hpp:67-68)
using namespace std;
using namespace std::placeholders;

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
//
// The loop runs in rounds (deferred.hpp): while set_intersect calls in the bodies miss the
// session's cache, their row sets are answered on the device and setup + every body run
// again; finish runs once, after the round in which no call missed.  As in the reference's
// apps, per-thread state is (re)initialised in setup and combined in finish.  A loop whose
// bodies make no such call runs once.
inline void par_for_range(size_t num_threads, size_t from, size_t to, size_t block,
                          const std::function<void(size_t)> &setup,
                          const std::function<void(size_t, size_t)> &body,
                          const std::function<void(size_t)> &finish) {
  if (num_threads == 0) num_threads = 1;
  auto round = [&](emptyheaded::Session *ses) {
    std::atomic<size_t> next(from);
    std::vector<std::thread> pool;
    for (size_t t = 0; t < num_threads; t++) setup(t);
    for (size_t t = 0; t < num_threads; t++)
      pool.emplace_back([&, t] {
        emptyheaded::tl_session = ses;
        emptyheaded::tl_tid = t;
        for (;;) {
          const size_t s = next.fetch_add(block, std::memory_order_relaxed);
          if (s >= to) break;
          const size_t e = s + block < to ? s + block : to;
          for (size_t i = s; i < e; i++) body(t, i);
        }
        emptyheaded::tl_session = nullptr;
      });
    for (auto &th : pool) th.join();
  };
  // a loop nested in a round's body runs its calls at once
  if (!emptyheaded::deferral_enabled() || emptyheaded::tl_session) {
    round(nullptr);
  } else {
    emptyheaded::Session ses(num_threads);
    for (;;) {
      if (!ses.begin_round()) {
        round(nullptr);  // bounded: a last round with every call run at once
        break;
      }
      round(&ses);
      if (!ses.missed()) break;
      ses.flush();
    }
  }
  for (size_t t = 0; t < num_threads; t++) finish(t);
}

}  // namespace common
