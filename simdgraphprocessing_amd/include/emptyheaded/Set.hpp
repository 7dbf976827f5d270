// Set.hpp -- Set<T>, the non-owning view of one sorted set (src/set/Set.hpp:14-132):
// {data, cardinality, number_of_bytes, density, type}.  The bytes are in the
// reference's formats for `type`; foreach walks them on the host.
//
// Beyond the reference's fields a set records where it came from: the matrix and the rows
// whose intersection it is (one row for SparseMatrix::get_row; the union of the operands'
// rows for an ops::set_intersect result), so that par_for_range's rounds can answer the
// intersection on the device in a batch (deferred.hpp).  Sets built from other data have no
// origin.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>

#include "deferred.hpp"
#include "layouts.hpp"

template <class T>
class Set {
 public:
  uint8_t *data;
  size_t cardinality;
  size_t number_of_bytes;
  double density;
  common::type type;
  emptyheaded::RowSetKey origin;  // origin.origin == nullptr: not the rows of a matrix

  Set(uint8_t *data_in, size_t cardinality_in, size_t number_of_bytes_in, double density_in, common::type type_in)
      : data(data_in), cardinality(cardinality_in), number_of_bytes(number_of_bytes_in), density(density_in),
        type(type_in) {}

  Set(uint8_t *data_in, size_t cardinality_in, size_t number_of_bytes_in, common::type type_in)
      : data(data_in), cardinality(cardinality_in), number_of_bytes(number_of_bytes_in), density(0.0),
        type(type_in) {}

  // an owned zeroed buffer (the reference leaks it too; Set.hpp:48-56)
  explicit Set(size_t number_of_bytes_in)
      : data(new uint8_t[number_of_bytes_in]()), cardinality(0), number_of_bytes(number_of_bytes_in), density(0.0),
        type(T::get_type()) {}

  // an output buffer for ops::set_intersect
  explicit Set(uint8_t *data_in)
      : data(data_in), cardinality(0), number_of_bytes(0), density(0.0), type(T::get_type()) {}

  Set(uint8_t *data_in, size_t number_of_bytes_in)
      : data(data_in), cardinality(0), number_of_bytes(number_of_bytes_in), density(0.0), type(T::get_type()) {}

  Set() : data(nullptr), cardinality(0), number_of_bytes(0), density(0.0), type(T::get_type()) {}

  template <class U>
  Set(const Set<U> &in)
      : data(in.data), cardinality(in.cardinality), number_of_bytes(in.number_of_bytes), density(in.density),
        type(in.type), origin(in.origin) {}

  template <class U>
  Set(const Set<U> *in)
      : data(in->data), cardinality(in->cardinality), number_of_bytes(in->number_of_bytes), density(in->density),
        type(in->type), origin(in->origin) {}

  // this set is row `row` of matrix `m` (SparseMatrix::get_row)
  void set_origin_row(const void *m, uint32_t row) {
    origin.origin = m;
    origin.n = 1;
    origin.rows[0] = row;
  }
  void clear_origin() { origin.origin = nullptr; origin.n = 0; }

  template <typename F>
  void foreach(F f) const {
    emptyheaded::foreach_bytes(f, data, cardinality, number_of_bytes, type);
  }

  // hybrid::foreach_until ignores the early stop (hybrid.hpp:211-217); this one stops
  template <typename F>
  void foreach_until(F f) const {
    bool done = false;
    emptyheaded::foreach_bytes([&](uint32_t x) { if (!done) done = f(x); }, data, cardinality, number_of_bytes,
                               type);
  }

  // Set<T>::from_array (src/set/Set.hpp:161-165) for the layouts built on the host:
  // uinteger (a copy) and bitset ([u64 first word][words], bitset.hpp:65-96)
  static Set<T> from_array(uint8_t *set_data, const uint32_t *array_data, size_t data_size) {
    const double density = data_size ? (double)((array_data[data_size - 1] - array_data[0]) / data_size) : 0.0;
    size_t bytes = 0;
    if (T::get_type() == common::UINTEGER) {
      std::memcpy(set_data, array_data, 4 * data_size);
      bytes = 4 * data_size;
    } else if (T::get_type() == common::BITSET && data_size) {
      const uint64_t w0 = array_data[0] >> 6, nw = (array_data[data_size - 1] >> 6) - w0 + 1;
      std::memcpy(set_data, &w0, 8);
      std::memset(set_data + 8, 0, 8 * nw);
      for (size_t k = 0; k < data_size; k++) {
        uint64_t w;
        uint8_t *p = set_data + 8 + 8 * ((array_data[k] >> 6) - w0);
        std::memcpy(&w, p, 8);
        w |= 1ull << (array_data[k] & 63u);
        std::memcpy(p, &w, 8);
      }
      bytes = 8 + 8 * nw;
    }
    return Set<T>(set_data, data_size, bytes, density, T::get_type());
  }

  // T::par_foreach (e.g. uinteger.hpp:101-118): f(tid, x) for every element; the host
  // walk here is one thread, so tid is 0 and 1 thread is reported
  template <typename F>
  size_t par_foreach(size_t num_threads, F f) const {
    (void)num_threads;
    foreach([&](uint32_t x) { f((size_t)0, x); });
    return 1;
  }

  void copy_from(const Set<T> &src) {
    std::memcpy(data, src.data, src.number_of_bytes);
    cardinality = src.cardinality;
    number_of_bytes = src.number_of_bytes;
    density = src.density;
    type = src.type;
    origin = src.origin;
  }
};
