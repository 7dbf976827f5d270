// The reference's clique test (test/n_clique_test.cpp:1-12) on this engine, for
// every layout and both schedules, plus query_depth 3 (= triangles):
//     test_cliques <facebook.bin>
#define GOOGLE_TEST
#include "n_clique.cpp"

#include "check.hpp"

template <class T>
static void run_case(MutableGraph *g, const char *layout, int depth, long expected, int schedule) {
  Parser input_data(4, false, depth, 0, g, layout);
  input_data.schedule = schedule;
  n_clique<T, T> clique_app(input_data);
  clique_app.run();
  EXPECT_EQ(expected, clique_app.num_cliques);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "usage: test_cliques <facebook.bin>" << std::endl;
    return 2;
  }
  MutableGraph *inputGraph = MutableGraph::undirectedFromBinary(argv[1]);
  for (int schedule : {EH_SCHED_TUNED, EH_SCHED_PLAIN}) {
    run_case<hybrid>(inputGraph, "hybrid", 4, 30004668L, schedule);
    run_case<uinteger>(inputGraph, "uint", 4, 30004668L, schedule);
    run_case<pshort>(inputGraph, "pshort", 4, 30004668L, schedule);
    run_case<bitset>(inputGraph, "bs", 4, 30004668L, schedule);
    run_case<hybrid>(inputGraph, "hybrid", 3, 1612010L, schedule);
    run_case<new_type>(inputGraph, "new_type", 4, 30004668L, schedule);
    run_case<bitset_new>(inputGraph, "bitset_new", 4, 30004668L, schedule);
  }
  delete inputGraph;
  return TEST_RESULT();
}
