// n_path -- the app of apps/n_path.cpp:1-95 on the device: the matrix keeps every entry,
// and the breadth-first search from the start row (each round: the frontier's rows
// OR-ed into the visited set, the new frontier = visited minus the previous visited)
// runs as one device search.  path_length is the number of rounds that found new
// vertices, capped at query_depth (--n; default 0xeffffff).
#define WRITE_VECTOR 1

#include "emptyheaded/emptyheaded.hpp"

template <class T, class R>
class n_path : public application<T, R> {
 public:
  size_t path_length;

  explicit n_path(Parser input_data) : application<T, R>(input_data), path_length(0) {}
  ~n_path() override { delete this->graph; }

  SparseMatrix<T, R> *materialize_graph() {
    auto node_selection = [](uint32_t node, uint32_t attribute) -> bool {
      (void)node;
      (void)attribute;
      return true;
    };
    auto edge_selection = [](uint32_t node, uint32_t nbr, uint32_t attribute) -> bool {
      (void)node;
      (void)nbr;
      (void)attribute;
      return true;
    };
    return SparseMatrix<T, R>::build(this->input_graph, node_selection, edge_selection, this->num_threads);
  }

  void run() override {
    this->graph = materialize_graph();
    const uint32_t internal_start =
        this->start_node == -1 ? this->graph->get_max_row_id() : this->graph->get_internal_id(this->start_node);
    std::cout << "Start node - External: " << this->graph->id_map[internal_start] << " Internal: " << internal_start
              << std::endl;
    const double start_time = common::startClock();
    const eh_bfs_info info = this->graph->bfs_path(internal_start, (uint32_t)this->query_depth);
    common::stopClock("pure bfs time", start_time);
    path_length = info.path_length;
    std::cout << "path length: " << path_length << std::endl;
  }
};

template <class T, class R>
application<T, R> *init_app(Parser input_data) {
  return new n_path<T, R>(input_data);
}
