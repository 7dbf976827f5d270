// directedEdgeListToBinary -- the reference's directed converter
// (tools/directedEdgeListToBinary.cpp:5-17): read a directed edge list and write the
// binary file MutableGraph::directedFromBinary reads (out-rows, then in-rows, per
// node, ids in order of first appearance).  Host only; no device work.
#include <iostream>

#include "emptyheaded/MutableGraph.hpp"

int main(int argc, char *argv[]) {
  if (argc != 3) {
    std::cout << "usage: " << argv[0] << " <input edge list> <output file>" << std::endl;
    return 0;
  }
  MutableGraph *g = MutableGraph::directedFromEdgeList(argv[1]);
  std::cout << "Loaded edge list" << std::endl;
  g->writeDirectedToBinary(argv[2]);
  delete g;
  return 0;
}
