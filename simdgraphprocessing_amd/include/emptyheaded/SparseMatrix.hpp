// SparseMatrix.hpp -- SparseMatrix<T,R> (src/SparseMatrix.hpp:14-35,90-93,156-183)
// over a device matrix (eh_matrix).
//
// build() evaluates the caller's node/edge predicates on the host exactly as
// pack_data does (entry (i, j) is kept iff node_selection(j, 0) &&
// edge_selection(i, j, 0), src/SparseMatrix.hpp:249-266) and hands the device
// the cheapest equivalent filter: all entries, the apps' orientation nbr < node,
// or an explicit byte mask.  The device then builds the per-row layouts for T
// (uinteger / pshort / bitset / hybrid and the selectors this engine adds).
//
// Rows live in HBM.  get_row(i) returns a Set<T> over a host copy of row i in the
// reference's byte format, fetched on first use and kept for the matrix's life;
// row_arrays[i] is that copy (null until fetched).  A directed graph's matrix also
// holds the in-neighbourhoods (get_column, src/SparseMatrix.hpp:162-165) as device
// rows n..2n-1.  Unlike the reference, the destructor frees everything it owns.
#pragma once

#include <atomic>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include "MutableGraph.hpp"
#include "Set.hpp"
#include "common.hpp"
#include "runtime.hpp"

template <class T, class R>
class SparseMatrix {
 public:
  size_t matrix_size;           // rows
  size_t cardinality;           // kept entries
  size_t row_total_bytes_used;  // bytes the reference's row images would take
  size_t col_total_bytes_used;
  size_t max_nbrhood_size;
  bool symmetric;
  uint32_t *row_lengths;
  uint8_t **row_arrays;
  uint64_t *id_map;
  eh_matrix *handle;
  double build_seconds;  // device time of the layout build

  typedef std::function<bool(uint32_t, uint32_t)> NodeSel;
  typedef std::function<bool(uint32_t, uint32_t, uint32_t)> EdgeSel;

  static SparseMatrix *build(MutableGraph *g, const NodeSel &node_selection, const EdgeSel &edge_selection,
                             size_t num_threads) {
    std::vector<uint64_t> off;
    std::vector<uint32_t> nbr;
    g->to_csr(off, nbr);
    const size_t n = g->num_nodes, rows = off.size() - 1;  // rows = 2n for a directed graph
    std::vector<uint8_t> mask(nbr.size());
    std::vector<uint32_t> lengths(rows);
    std::atomic<bool> all_kept(true), lower(g->symmetric);
    common::par_for_range(
        num_threads ? num_threads : 1, 0, rows, 256, [](size_t) {},
        [&](size_t, size_t r) {
          const uint32_t i = (uint32_t)(r < n ? r : r - n);
          uint32_t kept = 0;
          bool a = true, l = true;
          for (uint64_t k = off[r]; k < off[r + 1]; k++) {
            const uint32_t j = nbr[k];
            const bool keep = node_selection(j, 0) && edge_selection(i, j, 0);
            mask[k] = keep ? 1 : 0;
            kept += keep;
            a = a && keep;
            l = l && (keep == (j < i));
          }
          lengths[r] = kept;
          if (!a) all_kept.store(false, std::memory_order_relaxed);
          if (!l) lower.store(false, std::memory_order_relaxed);
        },
        [](size_t) {});
    const eh_edge_filter filter = all_kept ? EH_FILTER_ALL : (lower ? EH_FILTER_LOWER : EH_FILTER_MASK);
    emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
    eh_matrix *m = nullptr;
    eh_matrix_info info;
    {
      std::lock_guard<std::mutex> lock(rt.mutex());
      emptyheaded::check(eh_matrix_build_host(rt.ctx(), (uint32_t)rows, off.data(), nbr.data(), filter,
                                              filter == EH_FILTER_MASK ? mask.data() : nullptr, T::policy, &m),
                         "SparseMatrix::build");
      emptyheaded::check(eh_matrix_info_get(m, &info), "SparseMatrix::build");
    }
    auto *sm = new SparseMatrix(m, info, g);
    sm->filter_ = filter;
    std::copy(lengths.begin(), lengths.begin() + n, sm->row_lengths);
    if (!g->symmetric) {
      std::cout << "Number of rows + columns: " << rows << std::endl;
    }
    std::cout << "Number of nodes: " << sm->matrix_size << std::endl;
    std::cout << "Number of edges: " << sm->cardinality << std::endl;
    std::cout << "ROW DATA SIZE (Bytes): " << sm->row_total_bytes_used << std::endl;
    std::cout << "Time[parallel section]: " << sm->build_seconds << " s" << std::endl;
    return sm;
  }

  ~SparseMatrix() {
    eh_matrix_free(handle);
    for (size_t i = 0; i < device_rows_; i++) delete[] rows_[i].bytes;
    delete[] row_arrays;
    delete[] row_lengths;
  }

  Set<T> get_row(uint32_t row) const {
    const RowCopy &rc = fetch(row);
    Set<T> s(rc.bytes, rc.cardinality, rc.number_of_bytes, rc.type);
    s.set_origin_row(handle, row);
    return s;
  }

  // in-neighbours of a directed graph; the row itself for an undirected one
  Set<T> get_column(uint32_t column) const {
    const uint32_t row = symmetric ? column : (uint32_t)(matrix_size + column);
    const RowCopy &rc = fetch(row);
    Set<T> s(rc.bytes, rc.cardinality, rc.number_of_bytes, rc.type);
    s.set_origin_row(handle, row);
    return s;
  }

  // the first row of maximum length (src/SparseMatrix.hpp:134-145)
  uint32_t get_max_row_id() const {
    size_t best = 0;
    uint32_t id = 0;
    for (size_t i = 0; i < matrix_size; i++)
      if (row_lengths[i] > best) {
        best = row_lengths[i];
        id = (uint32_t)i;
      }
    std::cout << "MAXIMUM NODE: " << id_map[id] << " INTERNAL ID: " << id << std::endl;
    return id;
  }

  // the row of an external id, by linear search; 0 when absent (src/SparseMatrix.hpp:147-154)
  uint32_t get_internal_id(uint64_t external_id) const {
    for (size_t i = 0; i < matrix_size; i++)
      if (id_map[i] == external_id) return (uint32_t)i;
    return 0;
  }

  // n_path's search (apps/n_path.cpp:41-85) over the rows [0, matrix_size) on the device
  eh_bfs_info bfs_path(uint32_t start, uint32_t max_depth) const {
    eh_bfs_info info;
    emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
    std::lock_guard<std::mutex> lock(rt.mutex());
    // a directed matrix holds its in-neighbourhoods as rows matrix_size .. 2 matrix_size - 1;
    // rounds may run bottom-up only over rows that kept every entry
    const eh_bfs_mode mode =
        filter_ != EH_FILTER_ALL ? EH_BFS_TOP_DOWN : symmetric ? EH_BFS_SYMMETRIC : EH_BFS_IN_ROWS;
    emptyheaded::check(eh_bfs_path(handle, (uint32_t)matrix_size, start, max_depth, mode, nullptr, 0, &info),
                       "SparseMatrix::bfs_path");
    return info;
  }

  // |row(a[p]) & row(b[p])| for many pairs at once on the device (columns of a
  // directed graph are rows matrix_size + i)
  std::vector<uint64_t> intersect_rows(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) const {
    std::vector<uint64_t> out(a.size());
    if (a.empty()) return out;
    emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
    std::lock_guard<std::mutex> lock(rt.mutex());
    emptyheaded::check(eh_intersect_rows(handle, a.data(), b.data(), a.size(), out.data()), "intersect_rows");
    return out;
  }

  // with the reference's COMPRESSION 0 the decode buffer is unused (SparseMatrix.hpp:170-183)
  Set<R> get_decoded_row(uint32_t row, uint32_t *decode_buffer) const {
    (void)decode_buffer;
    const RowCopy &rc = fetch(row);
    Set<R> s(rc.bytes, rc.cardinality, rc.number_of_bytes, rc.type);
    s.set_origin_row(handle, row);
    return s;
  }

  // undirected_triangle_counting's loop (apps/undirected_triangle_counting.cpp:39-61) on the device
  size_t triangle_count(eh_run_stats *stats = nullptr, int schedule = EH_SCHED_TUNED) const {
    eh_run_opts o = {};
    o.schedule = schedule;
    uint64_t c = 0;
    emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
    std::lock_guard<std::mutex> lock(rt.mutex());
    emptyheaded::check(eh_triangle_count(handle, &o, &c, stats), "triangle_count");
    return (size_t)c;
  }

  // n_clique's recursion to query_depth k (apps/n_clique.cpp:30-101) on the device
  long clique_count(uint32_t k, eh_run_stats *stats = nullptr, int schedule = EH_SCHED_TUNED) const {
    eh_run_opts o = {};
    o.schedule = schedule;
    uint64_t c = 0;
    emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
    std::lock_guard<std::mutex> lock(rt.mutex());
    emptyheaded::check(eh_clique_count(handle, k, &o, &c, stats), "clique_count");
    return (long)c;
  }

 private:
  struct RowCopy {
    uint8_t *bytes = nullptr;
    size_t cardinality = 0;
    size_t number_of_bytes = 0;
    common::type type = common::UINTEGER;
    std::atomic<bool> ready{false};
  };

  SparseMatrix(eh_matrix *m, const eh_matrix_info &info, MutableGraph *g)
      : matrix_size(g->num_nodes), cardinality(info.cardinality), row_total_bytes_used(info.ref_row_bytes),
        col_total_bytes_used(0), max_nbrhood_size(g->max_nbrhood_size), symmetric(g->symmetric),
        row_lengths(new uint32_t[g->num_nodes]), row_arrays(new uint8_t *[g->num_nodes]()),
        id_map(g->id_map->data()), handle(m), build_seconds(info.build_ms * 1e-3), device_rows_(info.n),
        rows_(new RowCopy[info.n]) {}

  const RowCopy &fetch(uint32_t row) const {
    RowCopy &rc = rows_[row];
    if (rc.ready.load(std::memory_order_acquire)) return rc;
    emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
    std::lock_guard<std::mutex> lock(rt.mutex());
    if (rc.ready.load(std::memory_order_relaxed)) return rc;
    eh_row_view v;
    eh_status s = eh_matrix_row(handle, row, nullptr, 0, &v);
    if (s != EH_OK && s != EH_ERR_CAPACITY) emptyheaded::check(s, "SparseMatrix::get_row");
    uint8_t *buf = new uint8_t[v.number_of_bytes + 64]();
    emptyheaded::check(eh_matrix_row(handle, row, buf, v.number_of_bytes + 64, &v), "SparseMatrix::get_row");
    rc.bytes = buf;
    rc.cardinality = v.cardinality;
    rc.number_of_bytes = v.number_of_bytes;
    rc.type = (common::type)v.type;
    if (row < matrix_size) row_arrays[row] = buf;
    rc.ready.store(true, std::memory_order_release);
    return rc;
  }

  size_t device_rows_;
  std::unique_ptr<RowCopy[]> rows_;
  eh_edge_filter filter_ = EH_FILTER_ALL;  // the entries the build kept
};
