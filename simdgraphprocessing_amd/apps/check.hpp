// check.hpp -- the two assertions the app tests use, in place of gtest's EXPECT_EQ
// (test/undirected_triangle_counting_test.cpp, test/n_clique_test.cpp).
#pragma once

#include <iostream>

static int eh_failures = 0;

#define EXPECT_EQ(expected, actual)                                                                 \
  do {                                                                                              \
    const auto e_ = (expected);                                                                     \
    const auto a_ = (actual);                                                                       \
    if (!(e_ == a_)) {                                                                              \
      std::cerr << __FILE__ << ":" << __LINE__ << ": expected " << e_ << ", got " << a_ << std::endl; \
      eh_failures++;                                                                                \
    }                                                                                               \
  } while (0)

#define TEST_RESULT() (eh_failures == 0 ? (std::cout << "PASSED" << std::endl, 0) : (std::cout << "FAILED" << std::endl, 1))
