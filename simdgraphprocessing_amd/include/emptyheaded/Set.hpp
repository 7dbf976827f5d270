// Set.hpp -- Set<T>, the non-owning view of one sorted set (src/set/Set.hpp:14-132):
// {data, cardinality, number_of_bytes, density, type}.  The bytes are in the
// reference's formats for `type`; foreach walks them on the host.
#pragma once

#include <cstdint>
#include <cstring>

#include "layouts.hpp"

template <class T>
class Set {
 public:
  uint8_t *data;
  size_t cardinality;
  size_t number_of_bytes;
  double density;
  common::type type;

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
        type(in.type) {}

  template <class U>
  Set(const Set<U> *in)
      : data(in->data), cardinality(in->cardinality), number_of_bytes(in->number_of_bytes), density(in->density),
        type(in->type) {}

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

  void copy_from(const Set<T> &src) {
    std::memcpy(data, src.data, src.number_of_bytes);
    cardinality = src.cardinality;
    number_of_bytes = src.number_of_bytes;
    density = src.density;
    type = src.type;
  }
};
