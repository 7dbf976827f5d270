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
// Device failures raise emptyheaded::Error.
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

}  // namespace detail

template <class X, class Y, class Z>
inline Set<X> *set_intersect(Set<X> *C, const Set<Y> *A, const Set<Z> *B) {
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
  return C;
}

}  // namespace ops
