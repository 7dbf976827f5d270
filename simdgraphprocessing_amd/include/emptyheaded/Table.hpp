// Table.hpp -- Table<T> and ParallelTable<T> (src/Table.hpp:4-88): a column store of result
// tuples per thread, and the per-thread tables of an app.
//
// Same members and calls as the reference.  Differences: write_table appends one row per
// element of the set (tuple[0 .. num_columns-2] and id_map[element]) and stops at
// column_size rows, where the reference's loop writes every element to one cell
// (src/Table.hpp:40-42); ParallelTable::allocate resets a thread's existing table instead of
// leaking it (par_for_range may run its setup once per round, common.hpp); the columns are
// allocated untouched, so the n_clique app's large estimate costs address space only.
#pragma once

#include <cstddef>
#include <fstream>
#include <iostream>
#include <string>

#include "Set.hpp"

#ifndef PADDING
#define PADDING 300
#endif

template <class T>
class Table {
 public:
  size_t num_columns;
  size_t column_size;
  size_t cardinality;

  T **data;
  T *tuple;
  size_t *padding;

  // The n_clique app reads data[query_depth * tid] (apps/n_clique.cpp:81), past the columns
  // for every thread but the first; the slots past num_columns are kept null so that read
  // stays inside the array (get_decoded_row does not use the pointer).
  static constexpr size_t kSpareSlots = 16 * 1024;

  Table(const size_t num_columns_in, const size_t column_size_in)
      : num_columns(num_columns_in), column_size(column_size_in), cardinality(0),
        data(new T *[num_columns_in + kSpareSlots]()), tuple(new T[num_columns_in ? num_columns_in : 1]()),
        padding(new size_t[PADDING]) {
    for (size_t i = 0; i < num_columns; i++) data[i] = new T[column_size ? column_size : 1];
  }
  Table(const Table &) = delete;
  Table &operator=(const Table &) = delete;
  ~Table() {
    for (size_t i = 0; i < num_columns; i++) delete[] data[i];
    delete[] data;
    delete[] tuple;
    delete[] padding;
  }

  template <class R>
  void write_table(const Set<R> &c, const T *id_map) {
    if (num_columns == 0) return;
    c.foreach([&](uint32_t x) {
      if (cardinality >= column_size) return;
      for (size_t i = 0; i + 1 < num_columns; i++) data[i][cardinality] = tuple[i];
      data[num_columns - 1][cardinality] = id_map[x];
      cardinality++;
    });
  }

  void print_data(std::ofstream &file) const {
    for (size_t i = 0; i < cardinality; i++) {
      for (size_t j = 0; j < num_columns; j++) file << data[j][i] << "\t";
      file << std::endl;
    }
  }
};

template <class T>
class ParallelTable {
 public:
  Table<T> **table;
  size_t num_threads;
  size_t num_columns;
  size_t column_size;
  size_t cardinality;

  ParallelTable(size_t num_threads_in, size_t num_columns_in, size_t column_size_in)
      : table(new Table<T> *[num_threads_in ? num_threads_in : 1]()), num_threads(num_threads_in),
        num_columns(num_columns_in), column_size(column_size_in), cardinality(0) {}
  ParallelTable(const ParallelTable &) = delete;
  ParallelTable &operator=(const ParallelTable &) = delete;
  ~ParallelTable() {
    for (size_t t = 0; t < num_threads; t++) delete table[t];
    delete[] table;
  }

  void allocate(size_t tid) {
    if (table[tid]) table[tid]->cardinality = 0;
    else table[tid] = new Table<T>(num_columns, column_size);
  }

  void print_data(const std::string &filename) const {
    std::cout << "Printing data to table: " << filename << std::endl;
    std::ofstream file(filename);
    for (size_t t = 0; t < num_threads; t++)
      if (table[t]) table[t]->print_data(file);
  }
};
