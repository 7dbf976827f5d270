// Parser.hpp -- Parser and input_parser::parse (src/Parser.hpp:10-180): the app
// command line.  Flags as in the reference: --graph=<path> --input_type=<binary|text>
// --t=<threads> --layout=<uint|pshort|bs|hybrid> [--n=<clique size / path depth>]
// [--start_node=<id>] [--asymmetric] (a directed graph); the reference's main also
// takes new_type and bitset_new (src/emptyheaded.hpp:60-65); this engine adds
// --layout=hybrid3|cost|learned|hybrid_block and --schedule=<tuned|plain>.  Attribute
// files are not read.
#pragma once

#include <getopt.h>

#include <cstdlib>
#include <iostream>
#include <string>

#include "MutableGraph.hpp"
#include "eh.h"

class Parser {
 public:
  int num_threads;
  bool attributes;
  int n;
  long start_node;
  MutableGraph *input_graph;
  std::string layout;
  int schedule = EH_SCHED_TUNED;

  Parser(int num_threads_in, bool attributes_in, int n_in, size_t start_node_in, MutableGraph *in_graph,
         std::string layout_in)
      : num_threads(num_threads_in), attributes(attributes_in), n(n_in), start_node((long)start_node_in),
        input_graph(in_graph), layout(layout_in) {}
};

namespace input_parser {

inline void printUsage(const std::string &app) {
  std::cout << "USAGE: ./application <OPTIONS>" << std::endl;
  std::cout << "OPTIONS: " << std::endl;
  std::cout << "\tREQUIRED: --graph=<path to graph> --input_type=<'binary','text'> --t=<# of threads>" << std::endl;
  std::cout << "\t\t--layout=<uint,pshort,bs,hybrid,new_type,bitset_new,hybrid3,cost,learned,hybrid_block>" << std::endl;
  std::cout << "\tOPTIONAL: --schedule=<tuned,plain>" << std::endl;
  std::cout << "\tOPTIONAL: --asymmetric (directed input), --input_type=text (edge list)" << std::endl;
  if (app == "n_clique") {
    std::cout << "\tOPTIONAL: --n=<size of cliques to find, 3 to 16, default is 4>" << std::endl;
    std::cout << "*takes a undirected graph as input*" << std::endl;
  }
  if (app == "n_path") {
    std::cout << "\tOPTIONAL: --n=<maximum path length> --start_node=<external id, default: the largest row>"
              << std::endl;
  }
  std::exit(0);
}

inline Parser parse(int argc, char *argv[], const std::string &app) {
  const char *graph_path = nullptr, *input_type = nullptr, *layout_type = nullptr, *schedule = nullptr;
  int num_threads = 0, n = -1;
  long start_node = -1;
  bool help = false, attributes = false, directed = false;
  static struct option long_options[] = {{"help", no_argument, 0, 'h'},
                                         {"attributes", required_argument, 0, 'a'},
                                         {"t", required_argument, 0, 't'},
                                         {"graph", required_argument, 0, 'g'},
                                         {"layout", required_argument, 0, 'l'},
                                         {"n", required_argument, 0, 'n'},
                                         {"start_node", required_argument, 0, 's'},
                                         {"input_type", required_argument, 0, 'f'},
                                         {"asymmetric", optional_argument, 0, 'm'},
                                         {"schedule", required_argument, 0, 'k'},
                                         {0, 0, 0, 0}};
  optind = 1;
  for (;;) {
    int idx = 0;
    const int c = getopt_long(argc, argv, "a:t:g:n:f:s:hl:k:", long_options, &idx);
    if (c == -1) break;
    switch (c) {
      case 'a': attributes = true; break;
      case 'h': help = true; break;
      case 't': num_threads = std::atoi(optarg); break;
      case 'g': graph_path = optarg; break;
      case 'f': input_type = optarg; break;
      case 'l': layout_type = optarg; break;
      case 'n': n = std::atoi(optarg); break;
      case 's': start_node = std::atol(optarg); break;
      case 'm': directed = true; break;
      case 'k': schedule = optarg; break;
      default: break;
    }
  }
  if (num_threads == 0 || !graph_path || !input_type || help || !layout_type) printUsage(app);
  if (n == -1 && (app == "n_clique" || app == "n_cycle")) n = 4;
  if (n == -1 && app == "n_path") n = 0xeffffff;
  const std::string kind(input_type);
  if (attributes || (kind != "binary" && kind != "text")) {
    std::cerr << "ERROR: this engine reads binary or text edge-list graphs without attributes" << std::endl;
    std::exit(1);
  }
  MutableGraph *g;
  if (directed) {
    std::cout << "Loading a directed graph" << std::endl;
    g = kind == "text" ? MutableGraph::directedFromEdgeList(graph_path) : MutableGraph::directedFromBinary(graph_path);
  } else {
    std::cout << "Loading a undirected graph" << std::endl;
    g = kind == "text" ? MutableGraph::undirectedFromEdgeList(graph_path)
                       : MutableGraph::undirectedFromBinary(graph_path);
  }
  Parser p(num_threads, false, n, (size_t)start_node, g, layout_type);
  if (schedule) {
    const std::string s(schedule);
    if (s != "tuned" && s != "plain") {
      std::cerr << "ERROR: --schedule must be tuned or plain" << std::endl;
      std::exit(1);
    }
    p.schedule = s == "plain" ? EH_SCHED_PLAIN : EH_SCHED_TUNED;
  }
  return p;
}

}  // namespace input_parser
