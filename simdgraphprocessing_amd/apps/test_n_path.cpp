// The reference's BFS test (test/n_path_test.cpp:1-12) on this engine: dfacebook.bin as a
// directed graph, start node 0, depth 4 -> path length 4; plus the undirected facebook
// graph from its largest row with no depth limit -> 5 (the reference program's output),
// under several layouts:
//     test_n_path <dfacebook.bin> <facebook.bin>
#define GOOGLE_TEST
#include "n_path.cpp"

#include "check.hpp"

template <class T>
static void run_case(MutableGraph *g, const char *layout, int depth, long start, size_t expected) {
  Parser input_data(4, false, depth, (size_t)start, g, layout);
  n_path<T, T> app(input_data);
  app.run();
  EXPECT_EQ(expected, app.path_length);
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::cerr << "usage: test_n_path <dfacebook.bin> <facebook.bin>" << std::endl;
    return 2;
  }
  MutableGraph *d = MutableGraph::directedFromBinary(argv[1]);
  run_case<uinteger>(d, "uint", 4, 0, 4);
  run_case<hybrid>(d, "hybrid", 4, 0, 4);
  delete d;
  MutableGraph *u = MutableGraph::undirectedFromBinary(argv[2]);
  run_case<uinteger>(u, "uint", 0xeffffff, -1, 5);
  run_case<bitset>(u, "bs", 0xeffffff, -1, 5);
  run_case<pshort>(u, "pshort", 0xeffffff, -1, 5);
  run_case<hybrid>(u, "hybrid", 3, -1, 3);
  delete u;
  return TEST_RESULT();
}
