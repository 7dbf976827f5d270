// deferred.hpp -- the rounds in which common::par_for_range runs an app's loop, so that the
// ops::set_intersect calls of an unchanged reference-style body (one call per pair, nested
// per level: apps/undirected_triangle_counting.cpp:45-52, apps/n_clique.cpp:30-50) become a
// few batched device calls instead of one device round trip each.
//
// Every set taken from a matrix row carries its origin (the matrix and the row, Set.hpp), and
// the result of intersecting such sets carries the union of their rows.  Inside a round, a
// set_intersect whose operands both come from one matrix looks its row set up in the
// session's cache: a hit returns the result (count and, when the caller's WRITE_VECTOR asks
// for it, the elements in the reference's result layout), a miss records the row set and
// returns an empty set.  After the round one eh_intersect_row_sets call per matrix answers
// every recorded row set, and the next round runs the bodies again from their setup.  The
// first round with no miss is the last: every call in it returned exactly the result the
// reference's single pass computes, so the thread-local sums it leaves are the app's answer.
// An app whose results feed deeper calls (n_clique's level sets) needs one round per level
// plus the last one; the triangle app needs two.
//
// Sets of any other origin, block layouts and row sets of more than EH_ROW_SET_MAX rows go
// to the device one call at a time, as outside par_for_range.  EH_DEFERRED=0 turns the rounds
// off (every call then runs at once).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "eh.h"
#include "runtime.hpp"

namespace emptyheaded {

// the rows of one matrix whose intersection a set is (sorted, distinct)
struct RowSetKey {
  const void *origin = nullptr;  // the eh_matrix
  uint32_t n = 0;
  uint32_t rows[EH_ROW_SET_MAX];

  bool operator==(const RowSetKey &o) const {
    return origin == o.origin && n == o.n && std::equal(rows, rows + n, o.rows);
  }
};

struct RowSetKeyHash {
  size_t operator()(const RowSetKey &k) const {
    uint64_t h = 1469598103934665603ull ^ reinterpret_cast<uintptr_t>(k.origin);
    for (uint32_t q = 0; q < k.n; q++) h = (h ^ k.rows[q]) * 1099511628211ull;
    return (size_t)(h ^ (h >> 29));
  }
};

// The row set of A & B from the row sets of A and B; false when it would exceed the maximum.
inline bool merge_row_sets(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, RowSetKey &out) {
  uint32_t i = 0, j = 0, n = 0;
  while (i < na || j < nb) {
    uint32_t x;
    if (j >= nb || (i < na && a[i] < b[j])) x = a[i++];
    else if (i >= na || b[j] < a[i]) x = b[j++];
    else { x = a[i++]; j++; }
    if (n == EH_ROW_SET_MAX) return false;
    out.rows[n++] = x;
  }
  out.n = n;
  return true;
}

struct RowSetResult {
  uint64_t count = 0;
  uint64_t off = 0;        // first element in Session::elems
  bool has_elems = false;
};

class Session {
 public:
  explicit Session(size_t threads) : pending_(threads ? threads : 1) {}

  // the cache is only written between rounds, so lookups need no lock
  const RowSetResult *find(const RowSetKey &k, bool need_elems) const {
    const auto it = cache_.find(k);
    if (it == cache_.end() || (need_elems && !it->second.has_elems)) return nullptr;
    return &it->second;
  }
  const uint32_t *elements(const RowSetResult &r) const { return elems_.data() + r.off; }

  void record(size_t tid, const RowSetKey &k, bool need_elems) {
    pending_[tid].push_back({k, need_elems});
    missed_.store(true, std::memory_order_relaxed);
  }

  bool begin_round() {
    missed_.store(false, std::memory_order_relaxed);
    return ++rounds_ <= kMaxRounds;
  }
  bool missed() const { return missed_.load(std::memory_order_relaxed); }
  size_t rounds() const { return rounds_; }

  // Answer every recorded row set on the device (one eh_intersect_row_sets call per matrix
  // and kind: counts only, or counts and elements).
  void flush() {
    std::unordered_map<RowSetKey, bool, RowSetKeyHash> want;  // row set -> elements needed
    for (auto &v : pending_) {
      for (const auto &p : v) {
        bool &w = want[p.first];
        w = w || p.second;
      }
      v.clear();
    }
    // group by matrix and kind
    std::unordered_map<const void *, std::vector<RowSetKey>> groups[2];
    for (const auto &kv : want) {
      const auto it = cache_.find(kv.first);
      if (it != cache_.end() && (it->second.has_elems || !kv.second)) continue;
      groups[kv.second ? 1 : 0][kv.first.origin].push_back(kv.first);
    }
    Runtime &rt = Runtime::get();
    std::lock_guard<std::mutex> lock(rt.mutex());
    for (int mat = 0; mat < 2; mat++)
      for (auto &g : groups[mat]) {
        const std::vector<RowSetKey> &keys = g.second;
        for (size_t b = 0; b < keys.size(); b += kChunk) {
          const size_t e = std::min(keys.size(), b + kChunk), nq = e - b;
          std::vector<uint32_t> ids;
          std::vector<uint64_t> roff(nq + 1, 0), counts(nq), eoff(nq + 1, 0);
          for (size_t q = 0; q < nq; q++) {
            ids.insert(ids.end(), keys[b + q].rows, keys[b + q].rows + keys[b + q].n);
            roff[q + 1] = ids.size();
          }
          eh_matrix *m = static_cast<eh_matrix *>(const_cast<void *>(g.first));
          check(eh_intersect_row_sets(m, ids.data(), roff.data(), nq, counts.data(), nullptr, 0, eoff.data()),
                "ops::set_intersect (batched)");
          const uint64_t base = elems_.size();
          if (mat) {
            elems_.resize(base + eoff[nq]);
            check(eh_intersect_row_sets(m, ids.data(), roff.data(), nq, counts.data(), elems_.data() + base,
                                        eoff[nq], eoff.data()),
                  "ops::set_intersect (batched)");
          }
          for (size_t q = 0; q < nq; q++) {
            RowSetResult &r = cache_[keys[b + q]];
            r.count = counts[q];
            r.has_elems = mat == 1;
            r.off = mat ? base + eoff[q] : 0;
          }
        }
      }
  }

  static constexpr size_t kMaxRounds = 64;  // past this, a last round runs every call at once

 private:
  static constexpr size_t kChunk = 1u << 20;
  std::unordered_map<RowSetKey, RowSetResult, RowSetKeyHash> cache_;
  std::vector<uint32_t> elems_;
  std::vector<std::vector<std::pair<RowSetKey, bool>>> pending_;
  std::atomic<bool> missed_{false};
  size_t rounds_ = 0;
};

// the session of the par_for_range round this thread runs (null outside one) and its thread index
inline thread_local Session *tl_session = nullptr;
inline thread_local size_t tl_tid = 0;

inline bool deferral_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("EH_DEFERRED");
    return !(e && e[0] == '0');
  }();
  return on;
}

}  // namespace emptyheaded
