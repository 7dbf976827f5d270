// undirected_triangle_counting -- the app of apps/undirected_triangle_counting.cpp:1-71
// on the device: the matrix is built with the orientation nbr < node, and the
// sum over kept (i, j) of |row(i) & row(j)| runs as one counting pass.
#define WRITE_VECTOR 0

#include "emptyheaded/emptyheaded.hpp"

template <class T, class R>
class undirected_triangle_counting : public application<T, R> {
 public:
  size_t num_triangles;

  explicit undirected_triangle_counting(Parser input_data) : application<T, R>(input_data), num_triangles(0) {}
  ~undirected_triangle_counting() override { delete this->graph; }

  SparseMatrix<T, R> *materialize_graph() {
    auto node_selection = [](uint32_t node, uint32_t attribute) -> bool {
      (void)node;
      (void)attribute;
      return true;
    };
    auto edge_selection = [](uint32_t node, uint32_t nbr, uint32_t attribute) -> bool {
      (void)attribute;
      return nbr < node;
    };
    return SparseMatrix<T, R>::build(this->input_graph, node_selection, edge_selection, this->num_threads);
  }

  void run() override {
    this->graph = materialize_graph();
    const double start_time = common::startClock();
    num_triangles = this->graph->triangle_count(nullptr, this->schedule);
    common::stopClock("UNDIRECTED TRIANGLE COUNTING", start_time);
    std::cout << "Number of triangles: " << num_triangles << std::endl;
  }
};

template <class T, class R>
application<T, R> *init_app(Parser input_data) {
  return new undirected_triangle_counting<T, R>(input_data);
}
