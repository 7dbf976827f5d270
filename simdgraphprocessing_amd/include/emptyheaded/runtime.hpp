// runtime.hpp -- the process-wide device context behind the C++ surface.
//
// The reference runs everything on the host (src/emptyheaded.hpp:33-82).  Here
// one eh_ctx (device 0, or the id in EH_DEVICE) is created on first use and
// shared by every SparseMatrix, Set operation and app in the process.  Calls on
// one context must not overlap, so every C-ABI call made through this header
// holds the runtime's mutex.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

#include "eh.h"

namespace emptyheaded {

class Runtime {
 public:
  static Runtime &get() {
    static Runtime rt;
    return rt;
  }
  eh_ctx *ctx() const { return ctx_; }
  std::mutex &mutex() { return mu_; }

 private:
  Runtime() {
    const char *dev = std::getenv("EH_DEVICE");
    const eh_status s = eh_ctx_create(dev ? std::atoi(dev) : 0, &ctx_);
    if (s != EH_OK) {
      std::fprintf(stderr, "ERROR: no MI355X context: %s (%s)\n", eh_status_str(s), eh_last_error());
      std::exit(1);
    }
  }
  ~Runtime() { eh_ctx_destroy(ctx_); }
  Runtime(const Runtime &) = delete;
  Runtime &operator=(const Runtime &) = delete;
  eh_ctx *ctx_ = nullptr;
  std::mutex mu_;
};

class Error : public std::runtime_error {
 public:
  Error(eh_status s, const std::string &what)
      : std::runtime_error(what + ": " + eh_status_str(s) + " (" + eh_last_error() + ")"), status(s) {}
  eh_status status;
};

inline void check(eh_status s, const char *what) {
  if (s != EH_OK) throw Error(s, what);
}

}  // namespace emptyheaded
