// ParallelBuffer.hpp -- ParallelBuffer<T> (src/ParallelBuffer.hpp:4-24): one buffer of `size`
// T per thread, the output sets of the apps' ops::set_intersect calls.
//
// Same members and calls as the reference.  allocate(tid) keeps a buffer the thread already
// has (par_for_range may run its setup once per round, common.hpp), and the destructor frees
// what was allocated (the reference's leaks).
#pragma once

#include <cstddef>

template <class T>
class ParallelBuffer {
 public:
  size_t num_threads;
  size_t size;
  T **data;

  ParallelBuffer(size_t num_threads_in, size_t size_in)
      : num_threads(num_threads_in), size(size_in), data(new T *[num_threads_in ? num_threads_in : 1]()) {}
  ParallelBuffer(const ParallelBuffer &) = delete;
  ParallelBuffer &operator=(const ParallelBuffer &) = delete;
  ~ParallelBuffer() {
    for (size_t t = 0; t < num_threads; t++) delete[] data[t];
    delete[] data;
  }

  void allocate(size_t tid) {
    if (!data[tid]) data[tid] = new T[size];
  }
  void unallocate(size_t tid) {
    delete[] data[tid];
    data[tid] = nullptr;
  }
};
