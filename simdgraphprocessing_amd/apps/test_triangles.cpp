// The reference's triangle test (test/undirected_triangle_counting_test.cpp:1-13)
// on this engine, for every layout and both schedules:
//     test_triangles <facebook.bin>
#define GOOGLE_TEST
#include "undirected_triangle_counting.cpp"

#include "check.hpp"

template <class T>
static void run_case(MutableGraph *g, const char *layout, int schedule) {
  Parser input_data(4, false, 0, 0, g, layout);
  input_data.schedule = schedule;
  undirected_triangle_counting<T, T> triangle_app(input_data);
  triangle_app.run();
  size_t expected_result = 1612010;
  EXPECT_EQ(expected_result, triangle_app.num_triangles);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "usage: test_triangles <facebook.bin>" << std::endl;
    return 2;
  }
  MutableGraph *inputGraph = MutableGraph::undirectedFromBinary(argv[1]);
  for (int schedule : {EH_SCHED_TUNED, EH_SCHED_PLAIN}) {
    run_case<hybrid>(inputGraph, "hybrid", schedule);
    run_case<uinteger>(inputGraph, "uint", schedule);
    run_case<pshort>(inputGraph, "pshort", schedule);
    run_case<bitset>(inputGraph, "bs", schedule);
    run_case<hybrid3>(inputGraph, "hybrid3", schedule);
    run_case<new_type>(inputGraph, "new_type", schedule);
    run_case<bitset_new>(inputGraph, "bitset_new", schedule);
    run_case<hybrid_block>(inputGraph, "hybrid_block", schedule);
  }
  delete inputGraph;
  return TEST_RESULT();
}
