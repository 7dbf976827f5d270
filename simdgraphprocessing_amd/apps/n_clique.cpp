// n_clique -- the app of apps/n_clique.cpp:1-107 on the device: the matrix is built
// with the orientation nbr < node and the recursion of apply_function down to
// query_depth (3 to 16) runs as one counting pass.
#define WRITE_VECTOR 1

#include "emptyheaded/emptyheaded.hpp"

template <class T, class R>
class n_clique : public application<T, R> {
 public:
  long num_cliques;
  size_t query_depth;

  explicit n_clique(Parser input_data)
      : application<T, R>(input_data), num_cliques(0), query_depth((size_t)input_data.n) {}
  ~n_clique() override { delete this->graph; }

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
    num_cliques = this->graph->clique_count((uint32_t)query_depth, nullptr, this->schedule);
    common::stopClock("n_cliques", start_time);
    std::cout << "Number of cliques: " << num_cliques << std::endl;
  }
};

template <class T, class R>
application<T, R> *init_app(Parser input_data) {
  return new n_clique<T, R>(input_data);
}
