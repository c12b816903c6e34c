// Exact brute-force kNN producer (benchmark and test inputs).
#pragma once
#include "common.hpp"

namespace umapgpu {

// X: n x dim row-major device array; outputs n x k (self at slot 0).
void knn_bruteforce(const float* X, uint64_t n, uint64_t dim, uint64_t k, uint32_t* out_idx, float* out_dist,
                    hipStream_t s);

}  // namespace umapgpu
