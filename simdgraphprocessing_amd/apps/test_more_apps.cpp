// The reference's tests of the other set_intersect apps (test/similar_nodes_test.cpp,
// test/symbiosity_test.cpp, test/undirected_lollipop_counting_test.cpp) on this engine:
//     test_more_apps <app> <facebook.bin | dfacebook.bin>
// app: similar_nodes (904), symbiosity (4039 undirected), symbiosity_directed (0),
// lollipop (713,455,740).  Each app defines init_app, so one is compiled in per binary
// through EH_APP_SOURCE.
#define GOOGLE_TEST
#include EH_APP_SOURCE

#include "check.hpp"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::cerr << "usage: " << argv[0] << " <case> <graph.bin>" << std::endl;
    return 2;
  }
  const std::string which = argv[1];
#if defined(EH_APP_SIMILAR)
  MutableGraph *g = MutableGraph::undirectedFromBinary(argv[2]);
  for (const char *layout : {"hybrid", "uint", "pshort", "bs"}) {
    Parser input_data(4, false, 0, 0, g, layout);
    if (std::string(layout) == "hybrid") {
      similar_nodes<hybrid, hybrid> app(input_data);
      app.run();
      EXPECT_EQ((size_t)904, app.count);
    } else if (std::string(layout) == "uint") {
      similar_nodes<uinteger, uinteger> app(input_data);
      app.run();
      EXPECT_EQ((size_t)904, app.count);
    } else if (std::string(layout) == "pshort") {
      similar_nodes<pshort, pshort> app(input_data);
      app.run();
      EXPECT_EQ((size_t)904, app.count);
    } else {
      similar_nodes<bitset, bitset> app(input_data);
      app.run();
      EXPECT_EQ((size_t)904, app.count);
    }
  }
#elif defined(EH_APP_SYMBIOSITY)
  const bool directed = which == "symbiosity_directed";
  MutableGraph *g = directed ? MutableGraph::directedFromBinary(argv[2]) : MutableGraph::undirectedFromBinary(argv[2]);
  Parser input_data(4, false, 0, 0, g, "hybrid");
  symbiosity<hybrid, hybrid> app(input_data);
  app.run();
  EXPECT_EQ((size_t)(directed ? 0 : 4039), app.count);
#else
  MutableGraph *g = MutableGraph::undirectedFromBinary(argv[2]);
  Parser input_data(4, false, 0, 0, g, "hybrid");
  undirected_lollipop_counting<hybrid, hybrid> app(input_data);
  app.run();
  EXPECT_EQ((size_t)713455740, app.num_lollipops);
#endif
  delete g;
  return TEST_RESULT();
}
