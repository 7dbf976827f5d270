// symbiosity -- the app of apps/symbiosity.cpp:1-73 on the device: the number of
// vertices i whose out- and in-neighbourhoods overlap by more than N = 0.5 of their
// union, |row(i) & column(i)| / (|row(i)| + |column(i)| - |both|) > N.  All n
// intersections run as one eh_intersect_rows call (columns of a directed graph are
// device rows n + i; an undirected graph's column is its row).
#define WRITE_VECTOR 0

#include "emptyheaded/emptyheaded.hpp"

template <class T, class R>
class symbiosity : public application<T, R> {
 public:
  size_t count;
  double N;

  explicit symbiosity(Parser input_data) : application<T, R>(input_data), count(0), N(0.5) {}
  ~symbiosity() override { delete this->graph; }

  SparseMatrix<T, R> *materialize_graph() {
    auto node_selection = [](uint32_t node, uint32_t attribute) -> bool {
      (void)node;
      (void)attribute;
      return true;
    };
    auto edge_selection = [](uint32_t node, uint32_t nbr, uint32_t attribute) -> bool {
      (void)attribute;
      (void)node;
      (void)nbr;
      return true;
    };
    return SparseMatrix<T, R>::build(this->input_graph, node_selection, edge_selection, this->num_threads);
  }

  void run() override {
    this->graph = materialize_graph();
    const double start_time = common::startClock();
    SparseMatrix<T, R> *g = this->graph;
    const size_t n = g->matrix_size;
    std::vector<uint32_t> rows(n), cols(n);
    for (size_t i = 0; i < n; i++) {
      rows[i] = (uint32_t)i;
      cols[i] = (uint32_t)(g->symmetric ? i : n + i);
    }
    const std::vector<uint64_t> both = g->intersect_rows(rows, cols);
    count = 0;
    for (size_t i = 0; i < n; i++) {
      const double a = (double)g->get_row((uint32_t)i).cardinality;
      const double b = (double)g->get_column((uint32_t)i).cardinality;
      const double ratio = (double)both[i] / (a + b - (double)both[i]);
      if (ratio > N) count++;
    }
    common::stopClock("symbiosity", start_time);
    std::cout << "Count: " << count << std::endl << std::endl;
  }
};

template <class T, class R>
application<T, R> *init_app(Parser input_data) {
  return new symbiosity<T, R>(input_data);
}
