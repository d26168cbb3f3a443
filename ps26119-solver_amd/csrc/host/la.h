// la.h — host-side sparse structures and a deterministic parallel loop for the setup and
// verification stages (the iteration itself runs on the GPU).
//
// Csr uses int32 row pointers: every matrix the GPU path accepts has nnz < 2^31, and the
// device kernels index with 32-bit offsets.
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <thread>
#include <vector>

#include "contract.h"

namespace ps26119::amd {

struct Csr {
  int rows = 0, cols = 0;
  std::vector<int> ptr;  // rows + 1
  std::vector<int> col;
  std::vector<double> val;
  std::int64_t nnz() const { return ptr.empty() ? 0 : ptr.back(); }
};

// Worker count for host loops: all cores, capped (results never depend on it).
int host_threads();
void set_host_threads(int t);  // 0 = all cores

// Runs f(begin, end) over [0, n) split into contiguous pieces on up to host_threads() threads.
// Every index is visited by exactly one call, so loops whose iterations write disjoint
// outputs are deterministic for any thread count.
void parallel_for(std::int64_t n, const std::function<void(std::int64_t, std::int64_t)>& f,
                  std::int64_t min_piece = 1 << 15);

// Aᵀ in CSR straight from the model's CSC arrays (rows of Aᵀ list entries in the model's order).
Csr csc_as_transpose(const Model& m);
// Transpose by counting sort; rows of the result list their entries by increasing column
// (the order the reference CPU path sums in, reference src/la/csr.cpp:19-36).
Csr transpose(const Csr& a);
// out = a·x, each row summed sequentially in CSR order with an fp64 accumulator.
void multiply(const Csr& a, const double* x, double* out);
// out = Aᵀy over the model's CSC arrays: each column summed sequentially in storage order (the
// bits of multiply(csc_as_transpose(m), y, out), without copying the matrix).
void column_products(const Model& m, const double* y, double* out);

}  // namespace ps26119::amd
