// Principal-component projection of device data: the cheap, non-learned
// initial layout (and the subsample-fit init of the learned initialisation).
#pragma once
#include <vector>

#include "common.hpp"

namespace umapgpu {

struct PcaResult {
  std::vector<double> mean;        // [D]
  std::vector<double> components;  // [out_dim x D], unit rows, largest |entry| positive
  std::vector<double> variances;   // [out_dim], descending
};

// Y (n x out_dim, device) = (X - mean) projected on the top out_dim principal
// axes of X (n x D, device, D <= 384).  Synchronous.
PcaResult pca_project(const float* X, uint64_t n, uint64_t D, uint64_t out_dim, float* Y, hipStream_t s);

// Eigen-decomposition of a symmetric D x D matrix (row-major, f64) by cyclic
// Jacobi rotations: eigenvalues descending, eigenvectors as rows of `vecs`.
void symmetric_eigen(std::vector<double> a, uint64_t D, std::vector<double>* vals, std::vector<double>* vecs);

}  // namespace umapgpu
