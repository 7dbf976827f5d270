// emptyheaded.hpp -- the app framework (src/emptyheaded.hpp:1-85): application<T,R>,
// init_app<T,R>, and main() with the --layout switch (compiled out when
// GOOGLE_TEST is defined, as in the reference, so tests can include an app), plus the
// per-thread buffers and tables the reference's apps use (ParallelBuffer.hpp, Table.hpp).
//
// --layout=learned needs trained selector parameters: EH_SELECTOR_PARAMS names a
// file of 5508 little-endian float32 values (a raw dump, or a .npy whose header
// is skipped).
#pragma once

#include <cstdio>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "MutableGraph.hpp"
#include "ParallelBuffer.hpp"
#include "Parser.hpp"
#include "Set.hpp"
#include "SparseMatrix.hpp"
#include "Table.hpp"
#include "common.hpp"
#include "ops.hpp"
#include "runtime.hpp"

template <class T, class R>
class application {
 public:
  SparseMatrix<T, R> *graph = nullptr;
  MutableGraph *input_graph;
  size_t num_threads;
  std::string layout;
  size_t query_depth;
  long start_node;
  int schedule;

  explicit application(Parser input_data)
      : input_graph(input_data.input_graph), num_threads((size_t)input_data.num_threads),
        layout(input_data.layout), query_depth((size_t)input_data.n), start_node(input_data.start_node),
        schedule(input_data.schedule) {}
  virtual ~application() {}
  virtual void run() = 0;
};

template <class T, class R>
application<T, R> *init_app(Parser input_data);

namespace emptyheaded {

// load the learned selector's parameters into the shared context.  The selection serves the
// tuned triangle pass (only the rows it reads as an A-side or edge-centric B row are candidates);
// every other app, and the plain schedule, read every row's layout, so there every row held by
// another is a candidate.
inline void load_selector_params(const std::string &app, int schedule) {
  const char *path = std::getenv("EH_SELECTOR_PARAMS");
  if (!path) {
    std::cerr << "ERROR: --layout=learned needs EH_SELECTOR_PARAMS=<parameter file>" << std::endl;
    std::exit(1);
  }
  std::ifstream in(path, std::ios::binary);
  std::vector<char> raw((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  const size_t want = 4ull * EH_MLP_PARAMS;
  if (raw.size() < want) {
    std::cerr << "ERROR: " << path << " holds fewer than " << EH_MLP_PARAMS << " float32 values" << std::endl;
    std::exit(1);
  }
  const float *p = reinterpret_cast<const float *>(raw.data() + (raw.size() - want));
  Runtime &rt = Runtime::get();
  std::lock_guard<std::mutex> lock(rt.mutex());
  check(eh_selector_set_params(rt.ctx(), p), "selector parameters");
  const bool tuned_triangles = app == "undirected_triangle_counting" && schedule == EH_SCHED_TUNED;
  check(eh_selector_set_schedule(rt.ctx(), tuned_triangles ? 0 : -1, 0), "selector schedule");
}

template <class T>
inline int run_layout(const Parser &input_data) {
  application<T, T> *myapp = init_app<T, T>(input_data);
  myapp->run();
  delete myapp;
  return 0;
}

}  // namespace emptyheaded

#ifndef GOOGLE_TEST
int main(int argc, char *argv[]) {
  std::string s(argv[0]);
  const size_t slash = s.find_last_of('/');
  const std::string app = slash == std::string::npos ? s : s.substr(slash + 1);
  Parser input_data = input_parser::parse(argc, argv, app);
  common::alloc_scratch_space(512 * input_data.input_graph->max_nbrhood_size * sizeof(uint32_t),
                              (size_t)input_data.num_threads);
  const std::string &l = input_data.layout;
  if (l == "uint") return emptyheaded::run_layout<uinteger>(input_data);
  if (l == "bs") return emptyheaded::run_layout<bitset>(input_data);
  if (l == "pshort") return emptyheaded::run_layout<pshort>(input_data);
  if (l == "hybrid") return emptyheaded::run_layout<hybrid>(input_data);
  if (l == "hybrid3") return emptyheaded::run_layout<hybrid3>(input_data);
  if (l == "cost") return emptyheaded::run_layout<cost_hybrid>(input_data);
  if (l == "new_type") return emptyheaded::run_layout<new_type>(input_data);
  if (l == "bitset_new") return emptyheaded::run_layout<bitset_new>(input_data);
  if (l == "hybrid_block") return emptyheaded::run_layout<hybrid_block>(input_data);
  if (l == "learned") {
    emptyheaded::load_selector_params(app, input_data.schedule);
    return emptyheaded::run_layout<learned_hybrid>(input_data);
  }
  std::cout << "No valid layout entered." << std::endl;
  return 0;
}
#endif
