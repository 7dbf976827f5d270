// SparseMatrix::get_row and ops::set_intersect through the C++ surface, the way
// the reference's triangle loop uses them (apps/undirected_triangle_counting.cpp:45-56),
// checked against std::set_intersection over the host neighbourhoods, for every
// layout including the block layouts bitset_new / new_type:
//     test_set_ops <facebook.bin> [rows]
#define GOOGLE_TEST
#include <algorithm>
#include <iterator>

#include "emptyheaded/emptyheaded.hpp"
#include "check.hpp"

template <class T>
static void check_layout(MutableGraph *g, uint32_t rows) {
  auto all = [](uint32_t, uint32_t) { return true; };
  auto lower = [](uint32_t node, uint32_t nbr, uint32_t) { return nbr < node; };
  SparseMatrix<T, T> *m = SparseMatrix<T, T>::build(g, all, lower, 4);
  std::vector<uint8_t> out(64 + 16 * (g->max_nbrhood_size + 64) + 8 * (g->num_nodes / 64 + 8));
  auto kept = [&](uint32_t i) {
    std::vector<uint32_t> r;
    for (uint32_t j : *g->out_neighborhoods->at(i))
      if (j < i) r.push_back(j);
    return r;
  };
  size_t got_total = 0, want_total = 0;
  for (uint32_t i = 0; i < rows && i < g->num_nodes; i++) {
    Set<T> A = m->get_row(i);
    std::vector<uint32_t> a;
    A.foreach([&](uint32_t x) { a.push_back(x); });
    std::sort(a.begin(), a.end());  // new_type visits its uint part before its blocks
    const std::vector<uint32_t> ka = kept(i);
    EXPECT_EQ(ka.size(), a.size());
    EXPECT_EQ(true, a == ka);
    A.foreach([&](uint32_t j) {
      Set<T> B = m->get_row(j);
      Set<T> C(out.data());
      ops::set_intersect(&C, &A, &B);
      const std::vector<uint32_t> kb = kept(j);
      std::vector<uint32_t> want;
      std::set_intersection(ka.begin(), ka.end(), kb.begin(), kb.end(), std::back_inserter(want));
      std::vector<uint32_t> c;
      C.foreach([&](uint32_t x) { c.push_back(x); });
      std::sort(c.begin(), c.end());
      if (c != want) EXPECT_EQ(want.size(), c.size());
      got_total += C.cardinality;
      want_total += want.size();
    });
  }
  EXPECT_EQ(want_total, got_total);
  std::cout << "set_intersect over " << rows << " rows: " << got_total << " common elements" << std::endl;
  delete m;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "usage: test_set_ops <facebook.bin> [rows]" << std::endl;
    return 2;
  }
  const uint32_t rows = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 40;
  MutableGraph *g = MutableGraph::undirectedFromBinary(argv[1]);
  check_layout<hybrid>(g, rows);
  check_layout<pshort>(g, rows);
  check_layout<bitset>(g, rows);
  check_layout<uinteger>(g, rows);
  check_layout<hybrid3>(g, rows);
  check_layout<bitset_new>(g, rows);
  check_layout<new_type>(g, rows);
  check_layout<hybrid_block>(g, rows);
  delete g;
  return TEST_RESULT();
}
