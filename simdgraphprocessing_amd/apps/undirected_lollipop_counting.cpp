// undirected_lollipop_counting -- the app of apps/undirected_lollipop_counting.cpp:1-83
// on the device.  The reference adds |N(x)| for every x, y in N(x) and z in
// N(x) & N(y) with z < y; each pair {y, z} of adjacent neighbours of x is met once,
// so the count is sum over x of |N(x)| * t(x) with t(x) = (1/2) sum over y in N(x)
// of |N(x) & N(y)|, the triangles through x.  The 2m intersections |N(x) & N(y)|
// run as one eh_intersect_rows call.
#define WRITE_VECTOR 1

#include "emptyheaded/emptyheaded.hpp"

template <class T, class R>
class undirected_lollipop_counting : public application<T, R> {
 public:
  size_t num_lollipops;

  explicit undirected_lollipop_counting(Parser input_data) : application<T, R>(input_data), num_lollipops(0) {}
  ~undirected_lollipop_counting() override { delete this->graph; }

  SparseMatrix<T, R> *materialize_graph() {
    auto node_selection = [](uint32_t node, uint32_t attribute) -> bool {
      (void)attribute;
      (void)node;
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
    const MutableGraph *in = this->input_graph;
    const size_t n = this->graph->matrix_size;
    std::vector<uint32_t> xs, ys;
    xs.reserve(in->num_edges);
    ys.reserve(in->num_edges);
    for (size_t x = 0; x < n; x++)
      for (uint32_t y : *in->out_neighborhoods->at(x)) {
        xs.push_back((uint32_t)x);
        ys.push_back(y);
      }
    const std::vector<uint64_t> common = this->graph->intersect_rows(xs, ys);
    std::vector<uint64_t> through(n, 0);  // 2 t(x)
    for (size_t p = 0; p < xs.size(); p++) through[xs[p]] += common[p];
    num_lollipops = 0;
    for (size_t x = 0; x < n; x++) num_lollipops += (size_t)this->graph->row_lengths[x] * (through[x] / 2);
    common::stopClock("lollipop", start_time);
    std::cout << "Count: " << num_lollipops << std::endl;
  }
};

template <class T, class R>
application<T, R> *init_app(Parser input_data) {
  return new undirected_lollipop_counting<T, R>(input_data);
}
