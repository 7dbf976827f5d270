// undirectedEdgeListToBinary -- the reference's converter
// (tools/undirectedEdgeListToBinary.cpp:5-35): read an undirected edge list once and
// write one binary graph per vertex ordering, <out dir>/u_<ordering>.bin, in the
// format MutableGraph::undirectedFromBinary reads.  The renumbering runs on the
// device.  An optional third argument seeds the orderings that shuffle.
#include <cstdlib>
#include <iostream>
#include <string>

#include "emptyheaded/MutableGraph.hpp"
#include "emptyheaded/common.hpp"

namespace {

MutableGraph *clone(const MutableGraph &g) {
  auto *rows = new std::vector<std::vector<uint32_t> *>();
  for (auto *r : *g.out_neighborhoods) rows->push_back(new std::vector<uint32_t>(*r));
  return new MutableGraph(g.num_nodes, g.num_edges, g.max_nbrhood_size, true, new std::vector<uint64_t>(*g.id_map),
                          nullptr, rows, rows, nullptr, nullptr);
}

struct Ordering {
  const char *name;
  eh_ordering id;
};

}  // namespace

int main(int argc, char *argv[]) {
  if (argc != 3 && argc != 4) {
    std::cout << "usage: " << argv[0] << " <input edge list> <output dir> [seed]" << std::endl;
    return 0;
  }
  const uint64_t seed = argc == 4 ? std::strtoull(argv[3], nullptr, 10) : 0;
  MutableGraph *input = MutableGraph::undirectedFromEdgeList(argv[1]);
  static const Ordering orderings[] = {
      {"shingles", EH_ORDER_SHINGLES}, {"the_game", EH_ORDER_THE_GAME},     {"bfs", EH_ORDER_BFS},
      {"degree", EH_ORDER_DEGREE},     {"rev_degree", EH_ORDER_REV_DEGREE}, {"strong_run", EH_ORDER_STRONG_RUN},
      {"random", EH_ORDER_RANDOM}};
  try {
    for (const Ordering &o : orderings) {
      MutableGraph *g = clone(*input);
      const double t = common::startClock();
      g->reorder(o.id, seed);
      common::stopClock(o.name, t);
      g->writeUndirectedToBinary(std::string(argv[2]) + "/u_" + o.name + ".bin");
      std::cout << o.name << " generated" << std::endl;
      delete g;
    }
  } catch (const emptyheaded::Error &e) {
    std::cout << "ERROR: " << e.what() << std::endl;
    return 1;
  }
  delete input;
  return 0;
}
