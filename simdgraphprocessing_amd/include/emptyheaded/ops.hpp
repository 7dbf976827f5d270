// ops.hpp -- ops::set_intersect(Set<X>* C, const Set<Y>* A, const Set<Z>* B)
// (src/set/ops/intersection.hpp:829,880,956,1019,1023,1059,1062,1321,1325,2067),
// ops::set_union (union.hpp:6-127) and ops::set_difference (difference.hpp:5-84).
//
// One overload covers every (X, Y, Z): the pair is dispatched on A->type and
// B->type at run time by the device (eh_intersect).  As in the reference:
//   * an empty side gives an empty C;
//   * C's type follows the reference's result layouts (uint whenever a side is
//     uint, bitset for bitset x bitset, pshort for pshort x pshort/bitset,
//     bitset_new for bitset_new x bitset_new, new_type for new_type x new_type);
//   * C->data is caller-owned and must hold the result: at most
//     4*min(|A|,|B|) bytes for a uint result, 8 + 8*min(words) for a bitset
//     result, 6*min(|A|,|B|) for a pshort result, 8 + 20*min(|A|,|B|) for a
//     block result;
//   * WRITE_VECTOR, when defined to 0 before the include, counts without
//     writing C->data (apps/undirected_triangle_counting.cpp:1).
// Inside a par_for_range round (deferred.hpp) a call whose operands are rows of one matrix, or
// results of such calls, is answered from the round's cache of batched device results; every
// other call is one device round trip.  Device failures raise emptyheaded::Error.
#pragma once

#include <algorithm>
#include <mutex>

#include "Set.hpp"
#include "runtime.hpp"

#ifndef WRITE_VECTOR
#define WRITE_VECTOR 1
#endif

namespace ops {

namespace detail {

template <class S>
inline eh_row_view view_of(const S *s) {
  eh_row_view v;
  v.data = s->data;
  v.cardinality = s->cardinality;
  v.number_of_bytes = s->type == common::UINTEGER ? 4 * s->cardinality : s->number_of_bytes;
  v.type = (int32_t)s->type;
  v.reserved = 0;
  return v;
}

template <class S>
inline uint64_t words_of(const S *s) {
  return s->type == common::BITSET && s->number_of_bytes >= 8 ? (s->number_of_bytes - 8) / 8 : 0;
}

// the dispatcher's result layout for the operand types (src/set/ops/intersection.hpp:2066-2118)
inline common::type result_type(common::type ta, common::type tb) {
  if (ta == common::BITSET && tb == common::BITSET) return common::BITSET;
  if ((ta == common::PSHORT || tb == common::PSHORT) && ta != common::UINTEGER && tb != common::UINTEGER)
    return common::PSHORT;
  return common::UINTEGER;
}

// C->data from the sorted result elements, in layout rt; returns the byte count.  BITSET:
// [u64 first word][words] over the words both operands span (bitset.hpp:65-96,
// intersection.hpp:880-955); PSHORT: [prefix, count-1, low16...] per non-empty 16-bit
// prefix (pshort.hpp:46-78).
template <class Y, class Z>
inline size_t encode_result(uint8_t *out, const uint32_t *e, uint64_t n, common::type rt, const Set<Y> *A,
                            const Set<Z> *B) {
  if (rt == common::UINTEGER) {
    if (n) std::memcpy(out, e, 4 * n);
    return 4 * n;
  }
  if (rt == common::BITSET) {
    uint64_t a0 = 0, b0 = 0;
    std::memcpy(&a0, A->data, 8);
    std::memcpy(&b0, B->data, 8);
    const uint64_t ea = a0 + words_of(A), eb = b0 + words_of(B);
    const uint64_t start = std::max(a0, b0), end = std::min(ea, eb);
    const uint64_t total = start < end ? end - start : 0;
    std::memcpy(out, &start, 8);
    std::memset(out + 8, 0, 8 * total);
    for (uint64_t k = 0; k < n; k++) {
      uint8_t *w = out + 8 + 8 * ((e[k] >> 6) - start);
      uint64_t v;
      std::memcpy(&v, w, 8);
      v |= 1ull << (e[k] & 63u);
      std::memcpy(w, &v, 8);
    }
    return 8 + 8 * total;
  }
  uint16_t *o = reinterpret_cast<uint16_t *>(out);
  size_t w = 0;
  for (uint64_t k = 0; k < n;) {
    const uint32_t hi = e[k] >> 16;
    uint64_t j = k;
    while (j < n && (e[j] >> 16) == hi) j++;
    const uint16_t pre = (uint16_t)hi, cnt = (uint16_t)(j - k - 1);
    std::memcpy(o + w, &pre, 2);
    std::memcpy(o + w + 1, &cnt, 2);
    w += 2;
    for (; k < j; k++) {
      const uint16_t lo = (uint16_t)(e[k] & 0xFFFFu);
      std::memcpy(o + w++, &lo, 2);
    }
  }
  return 2 * w;
}

inline bool deferrable(common::type t) {
  return t == common::UINTEGER || t == common::BITSET || t == common::PSHORT;
}

}  // namespace detail

template <class X, class Y, class Z>
inline Set<X> *set_intersect(Set<X> *C, const Set<Y> *A, const Set<Z> *B) {
  emptyheaded::RowSetKey key;
  const bool rows_known = A->origin.origin && A->origin.origin == B->origin.origin &&
                          emptyheaded::merge_row_sets(A->origin.rows, A->origin.n, B->origin.rows, B->origin.n, key);
  if (rows_known) key.origin = A->origin.origin;
  emptyheaded::Session *ses = emptyheaded::tl_session;
  if (ses && rows_known && detail::deferrable(A->type) && detail::deferrable(B->type)) {
    const common::type rt = detail::result_type(A->type, B->type);
    C->density = 0.0;
    C->type = rt;
    C->origin = key;
    C->cardinality = 0;
    C->number_of_bytes = 0;
    // the reference returns an empty result as soon as either side is empty
    if (A->cardinality == 0 || B->cardinality == 0) return C;
    const bool want = WRITE_VECTOR != 0;
    const emptyheaded::RowSetResult *r = ses->find(key, want);
    if (!r) {
      ses->record(emptyheaded::tl_tid, key, want);
      return C;
    }
    C->cardinality = r->count;
    if (want) C->number_of_bytes = detail::encode_result(C->data, ses->elements(*r), r->count, rt, A, B);
    return C;
  }
  const eh_row_view a = detail::view_of(A), b = detail::view_of(B);
  const uint64_t mc = std::min<uint64_t>(A->cardinality, B->cardinality);
  const uint64_t cap =
      std::max<uint64_t>({4 * mc, 8 + 8 * std::min(detail::words_of(A), detail::words_of(B)), 6 * mc, 8 + 20 * mc});
  eh_row_view out;
  emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
  {
    std::lock_guard<std::mutex> lock(rt.mutex());
    emptyheaded::check(eh_intersect(rt.ctx(), &a, &b, WRITE_VECTOR ? 1 : 0, WRITE_VECTOR ? C->data : nullptr,
                                    WRITE_VECTOR ? cap : 0, &out),
                       "ops::set_intersect");
  }
  C->cardinality = out.cardinality;
  C->number_of_bytes = WRITE_VECTOR ? out.number_of_bytes : 0;
  C->density = 0.0;
  C->type = (common::type)out.type;
  if (rows_known) C->origin = key;
  else C->clear_origin();
  return C;
}

// ops::set_union(Set<bitset>* A, Set<X>* B) (src/set/ops/union.hpp:6-127): B's
// elements OR-ed into A's words in place (B bitset, uinteger or pshort; elements
// outside A's word span are dropped).  Like the reference, A's cardinality is left as
// it was.
template <class X>
inline void set_union(Set<bitset> *A, const Set<X> *B) {
  const eh_row_view a = detail::view_of(A), b = detail::view_of(B);
  eh_row_view out;
  emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
  std::lock_guard<std::mutex> lock(rt.mutex());
  emptyheaded::check(eh_union(rt.ctx(), &a, &b, A->data, A->number_of_bytes, &out), "ops::set_union");
  A->clear_origin();  // A's bytes are no longer those of its rows
}

// ops::set_difference(Set<uinteger>* C, const Set<bitset>* A, const Set<bitset>* B)
// (src/set/ops/difference.hpp:5-84): C = the elements of A not in B, sorted; C->data
// must hold 4 bytes per element of A.  C is left as it was unless both sides have bytes.
inline Set<uinteger> *set_difference(Set<uinteger> *C, const Set<bitset> *A, const Set<bitset> *B) {
  const eh_row_view a = detail::view_of(A), b = detail::view_of(B);
  eh_row_view out;
  emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
  {
    std::lock_guard<std::mutex> lock(rt.mutex());
    const uint64_t cap = A->number_of_bytes >= 8 ? 256 * ((A->number_of_bytes - 8) / 8) : 0;
    emptyheaded::check(eh_difference(rt.ctx(), &a, &b, C->data, cap, &out), "ops::set_difference");
  }
  if (A->number_of_bytes == 0 || B->number_of_bytes == 0) return C;
  C->cardinality = out.cardinality;
  C->number_of_bytes = out.number_of_bytes;
  C->density = 0.0;
  C->type = common::UINTEGER;
  C->clear_origin();
  return C;
}

}  // namespace ops
