// similar_nodes -- the app of apps/similar_nodes.cpp:1-73 on the device: A is the
// longest row (get_max_row_id), and the count is the number of rows i with
// |A & row(i)| > N (N = 10).  All n intersections run as one eh_intersect_rows call.
#define WRITE_VECTOR 0

#include "emptyheaded/emptyheaded.hpp"

template <class T, class R>
class similar_nodes : public application<T, R> {
 public:
  size_t count;
  size_t N;

  explicit similar_nodes(Parser input_data) : application<T, R>(input_data), count(0), N(10) {}
  ~similar_nodes() override { delete this->graph; }

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
    const uint32_t node = this->graph->get_max_row_id();
    const size_t n = this->graph->matrix_size;
    std::vector<uint32_t> a(n, node), b(n);
    for (size_t i = 0; i < n; i++) b[i] = (uint32_t)i;
    const std::vector<uint64_t> common = this->graph->intersect_rows(a, b);
    count = 0;
    for (uint64_t c : common) count += c > N;
    common::stopClock("similar_nodes", start_time);
    std::cout << "Count: " << count << std::endl << std::endl;
  }
};

template <class T, class R>
application<T, R> *init_app(Parser input_data) {
  return new similar_nodes<T, R>(input_data);
}
