// Embedding-quality evaluators on the device (fuzzy cross-entropy,
// trustworthiness ranks).  The reference defines none.
#pragma once
#include "graph.hpp"

namespace umapgpu {

struct CeTerms {
  uint64_t n = 0;
  double attract = 0.0;      // sum over finite graph edges
  uint64_t edges = 0;        // finite graph edges counted
  double repulse_sum = 0.0;  // sum over kept random pairs of -log(1-v)
  uint64_t pairs = 0;        // kept random pairs
  double total = 0.0;        // attract + (n^2 - edges)/pairs * repulse_sum
};

// Key of the random pair stream for `seed`.
uint64_t ce_pair_key(uint64_t seed);

// Graph as device CSR (indptr u64[n+1], indices u32, data f32); y: n x dim
// device embedding.  Synchronous.
CeTerms fuzzy_cross_entropy(const uint64_t* indptr, const uint32_t* indices, const float* data, uint64_t n,
                            const float* y, uint64_t dim, float a, float b, uint64_t n_random, uint64_t seed,
                            hipStream_t s);

// Sum over queries and their k low-dimensional neighbours of
// max(rank - k, 0), ranks in the high-dimensional data X (n x dim).
// queries: device u32[nq]; nbrs: device u32[nq x k].  Synchronous.
double trustworthiness_penalty(const float* X, uint64_t n, uint64_t dim, const uint32_t* queries, uint64_t nq,
                               const uint32_t* nbrs, uint64_t k, hipStream_t s);

}  // namespace umapgpu
