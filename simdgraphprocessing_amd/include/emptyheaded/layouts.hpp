// layouts.hpp -- the layout tag classes of src/set/layouts/ (uinteger.hpp,
// pshort.hpp, bitset.hpp, hybrid.hpp, bitset_new.hpp, new_type.hpp).  In the reference they hold the
// encoders and the per-layout loops; here the encoders run on the device, so a
// tag carries its type value, the matrix layout policy it selects, and the
// host iteration over a set's bytes that Set<T>::foreach needs.
#pragma once

#include <cstdint>
#include <cstring>

#include "common.hpp"

namespace emptyheaded {

// Iterate the blocks of a bitset_new image: [u32 ids][2 u64 words per block]
// (bitset_new.hpp:159-191).
template <typename F>
inline void foreach_blocks(F f, const uint8_t *data, size_t number_of_bytes) {
  const size_t nb = number_of_bytes / 20;
  for (size_t b = 0; b < nb; b++) {
    uint32_t id;
    std::memcpy(&id, data + 4 * b, 4);
    for (size_t h = 0; h < 2; h++) {
      uint64_t word;
      std::memcpy(&word, data + 4 * nb + 16 * b + 8 * h, 8);
      while (word) {
        f((uint32_t)(128u * id + 64u * h + (unsigned)__builtin_ctzll(word)));
        word &= word - 1;
      }
    }
  }
}

// Iterate the elements of a set held in one of the reference byte formats:
// uint u32[card]; bitset [u64 first word][u64 words]; pshort u16 partitions
// [prefix, count-1, low16...] (bitset.hpp:65-96, pshort.hpp:46-78); bitset_new
// blocks; new_type [u64 uint bytes][uint part][bitset_new part], uint part first
// (new_type.hpp:135-152).
template <typename F>
inline void foreach_bytes(F f, const uint8_t *data, size_t cardinality, size_t number_of_bytes, common::type t) {
  // a uint set is as long as its cardinality; the other layouts are walked by their bytes, as
  // the reference's bitset and pshort loops are (bitset.hpp:148-173), so a bitset filled in
  // place (bitset::set, set_union) iterates its bits whatever its cardinality field holds
  if (data == nullptr || (t == common::UINTEGER && cardinality == 0)) return;
  if (t == common::BITSET_NEW) {
    foreach_blocks(f, data, number_of_bytes);
  } else if (t == common::NEW_TYPE) {
    if (number_of_bytes < 8) return;
    uint64_t ub;
    std::memcpy(&ub, data, 8);
    for (size_t k = 0; k < ub / 4; k++) {
      uint32_t v;
      std::memcpy(&v, data + 8 + 4 * k, 4);
      f(v);
    }
    foreach_blocks(f, data + 8 + ub, number_of_bytes - 8 - ub);
  } else if (t == common::BITSET) {
    if (number_of_bytes < 16) return;
    uint64_t w0;
    std::memcpy(&w0, data, 8);
    const size_t nw = (number_of_bytes - 8) / 8;
    for (size_t w = 0; w < nw; w++) {
      uint64_t word;
      std::memcpy(&word, data + 8 + 8 * w, 8);
      while (word) {
        const unsigned b = (unsigned)__builtin_ctzll(word);
        f((uint32_t)((w0 + w) * 64 + b));
        word &= word - 1;
      }
    }
  } else if (t == common::PSHORT) {
    const size_t ns = number_of_bytes / 2;
    size_t i = 0;
    while (i + 1 < ns) {
      uint16_t pre, cnt1;
      std::memcpy(&pre, data + 2 * i, 2);
      std::memcpy(&cnt1, data + 2 * (i + 1), 2);
      const size_t cnt = (size_t)cnt1 + 1;
      for (size_t k = 0; k < cnt; k++) {
        uint16_t lo;
        std::memcpy(&lo, data + 2 * (i + 2 + k), 2);
        f(((uint32_t)pre << 16) | lo);
      }
      i += 2 + cnt;
    }
  } else {
    for (size_t k = 0; k < cardinality; k++) {
      uint32_t v;
      std::memcpy(&v, data + 4 * k, 4);
      f(v);
    }
  }
}

}  // namespace emptyheaded

#define EH_LAYOUT_TAG(NAME, TYPE, POLICY)                                                        \
  class NAME {                                                                                   \
   public:                                                                                       \
    static common::type get_type() { return TYPE; }                                              \
    static constexpr eh_layout_policy policy = POLICY;                                           \
    template <typename F>                                                                        \
    static void foreach(F f, const uint8_t *data, size_t card, size_t bytes, common::type t) {   \
      emptyheaded::foreach_bytes(f, data, card, bytes, t);                                       \
    }                                                                                            \
    /* the bit helpers of bitset.hpp:50-63 */                                                    \
    static size_t word_index(uint32_t bit_index) { return bit_index >> 6; }                      \
    static bool is_set(uint32_t index, const uint64_t *in_array, uint64_t start_index) {         \
      return (in_array[word_index(index) - start_index] >> (index & 63u)) & 1ull;                \
    }                                                                                            \
    static void set(uint32_t index, uint64_t *in_array, uint64_t start_index) {                  \
      in_array[word_index(index) - start_index] |= 1ull << (index & 63u);                        \
    }                                                                                            \
  };

// src/set/layouts/uinteger.hpp, pshort.hpp, bitset.hpp, hybrid.hpp
EH_LAYOUT_TAG(uinteger, common::UINTEGER, EH_LAYOUT_UINT)
EH_LAYOUT_TAG(pshort, common::PSHORT, EH_LAYOUT_PSHORT)
EH_LAYOUT_TAG(bitset, common::BITSET, EH_LAYOUT_BITSET)
EH_LAYOUT_TAG(hybrid, common::HYBRID, EH_LAYOUT_HYBRID)
// selectors this engine adds: compute_type's pshort band (hybrid.hpp:55-63), the
// byte-cost rule and the learned per-row selector
EH_LAYOUT_TAG(hybrid3, common::HYBRID, EH_LAYOUT_HYBRID3)
EH_LAYOUT_TAG(cost_hybrid, common::HYBRID, EH_LAYOUT_COST)
EH_LAYOUT_TAG(learned_hybrid, common::HYBRID, EH_LAYOUT_LEARNED)
// the block-bitset layouts (--layout=bitset_new / new_type, src/emptyheaded.hpp:60-65)
// and hybrid built with NEW_BITSET (src/set/layouts/hybrid.hpp:83-87)
EH_LAYOUT_TAG(bitset_new, common::BITSET_NEW, EH_LAYOUT_BITSET_NEW)
EH_LAYOUT_TAG(new_type, common::NEW_TYPE, EH_LAYOUT_NEW_TYPE)
EH_LAYOUT_TAG(hybrid_block, common::HYBRID, EH_LAYOUT_HYBRID_BLOCK)

#undef EH_LAYOUT_TAG
