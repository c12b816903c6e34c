// Principal-component projection on the device.
//
// 1. column means (f64 block partials, summed on the host in a fixed order);
// 2. the centred covariance as 128 x 128 tiles: each block walks its share of
//    the rows, 32 at a time through LDS, and each thread accumulates an 8 x 8
//    corner of the tile in f32; per-block partials are summed in f64 by a
//    second kernel in block order, so the result does not depend on timing;
// 3. cyclic Jacobi on the host for the D x D eigenproblem (D <= 384);
// 4. projection: 32 rows per block staged through LDS, 8 lanes per row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pca.hpp"

namespace umapgpu {

namespace {

constexpr int kThreads = 256;
constexpr int kTile = 128;    // covariance tile edge
constexpr int kRows = 32;     // rows per LDS step
constexpr int kMaxD = 384;  // keeps the projection's LDS under 64 KiB
constexpr int kMaxOut = 8;

__global__ __launch_bounds__(kThreads) void col_sum_kernel(const float* __restrict__ X, uint64_t n, uint32_t D,
                                                           uint64_t chunk, double* __restrict__ part) {
  const uint64_t r0 = (uint64_t)blockIdx.x * chunk;
  const uint64_t r1 = r0 + chunk < n ? r0 + chunk : n;
  for (uint32_t c = threadIdx.x; c < D; c += kThreads) {
    double s = 0.0;
    for (uint64_t r = r0; r < r1; ++r) s += (double)X[r * D + c];
    part[(uint64_t)blockIdx.x * D + c] = s;
  }
}

// blockIdx.y enumerates the upper tile pairs (bi <= bj).
__global__ __launch_bounds__(kThreads) void cov_kernel(const float* __restrict__ X, uint64_t n, uint32_t D,
                                                       const float* __restrict__ mean,
                                                       const uint32_t* __restrict__ pairs, uint32_t npairs,
                                                       float* __restrict__ part) {
  __shared__ float xi[kRows][kTile];
  __shared__ float xj[kRows][kTile + 4];
  const uint32_t pr = pairs[blockIdx.y];
  const uint32_t bi = pr >> 16, bj = pr & 0xffffu;
  const int t = threadIdx.x, ti = t >> 4, tj = t & 15;
  float acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = 0.0f;
  for (uint64_t c = (uint64_t)blockIdx.x * kRows; c < n; c += (uint64_t)gridDim.x * kRows) {
    __syncthreads();
    for (int e = t; e < kRows * kTile; e += kThreads) {
      const int r = e / kTile, col = e % kTile;
      const uint64_t row = c + r;
      const uint32_t gi = bi * kTile + col, gj = bj * kTile + col;
      xi[r][col] = (row < n && gi < D) ? X[row * D + gi] - mean[gi] : 0.0f;
      xj[r][col] = (row < n && gj < D) ? X[row * D + gj] - mean[gj] : 0.0f;
    }
    __syncthreads();
    for (int r = 0; r < kRows; ++r) {
      float a[8], b[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        a[k] = xi[r][ti * 8 + k];
        b[k] = xj[r][tj * 8 + k];
      }
#pragma unroll
      for (int p = 0; p < 8; ++p)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[p][q] = acc[p][q] + a[p] * b[q];
    }
  }
  float* out = part + ((uint64_t)blockIdx.x * npairs + blockIdx.y) * (kTile * kTile);
#pragma unroll
  for (int p = 0; p < 8; ++p)
#pragma unroll
    for (int q = 0; q < 8; ++q) out[(ti * 8 + p) * kTile + tj * 8 + q] = acc[p][q];
}

// out[e] = sum over blocks b (in order) of part[b][e], e < len.
__global__ __launch_bounds__(kThreads) void sum_blocks_kernel(const float* __restrict__ part, uint32_t nblocks,
                                                              uint64_t len, double* __restrict__ out) {
  const uint64_t e = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= len) return;
  double s = 0.0;
  for (uint32_t b = 0; b < nblocks; ++b) s += (double)part[(uint64_t)b * len + e];
  out[e] = s;
}

__global__ __launch_bounds__(kThreads) void project_kernel(const float* __restrict__ X, uint64_t n, uint32_t D,
                                                           const float* __restrict__ mean,
                                                           const float* __restrict__ V, uint32_t O,
                                                           float* __restrict__ Y) {
  extern __shared__ float sh[];
  float* xs = sh;                          // [kRows][D + 1]
  float* mv = sh + kRows * (D + 1);        // [D]
  float* vv = mv + D;                      // [O][D]
  const int t = threadIdx.x;
  const uint64_t r0 = (uint64_t)blockIdx.x * kRows;
  for (uint32_t e = t; e < D; e += kThreads) mv[e] = mean[e];
  for (uint32_t e = t; e < O * D; e += kThreads) vv[e] = V[e];
  for (uint32_t e = t; e < kRows * D; e += kThreads) {
    const uint32_t r = e / D, d = e % D;
    xs[r * (D + 1) + d] = r0 + r < n ? X[(r0 + r) * D + d] : 0.0f;
  }
  __syncthreads();
  const int r = t >> 3, q = t & 7;
  float y[kMaxOut];
#pragma unroll
  for (int o = 0; o < kMaxOut; ++o) y[o] = 0.0f;
  for (uint32_t d = q; d < D; d += 8) {
    const float xc = xs[r * (D + 1) + d] - mv[d];
#pragma unroll
    for (int o = 0; o < kMaxOut; ++o)
      if (o < (int)O) y[o] = y[o] + xc * vv[o * D + d];
  }
#pragma unroll
  for (int o = 0; o < kMaxOut; ++o) {
    y[o] = y[o] + __shfl_xor(y[o], 1, 64);
    y[o] = y[o] + __shfl_xor(y[o], 2, 64);
    y[o] = y[o] + __shfl_xor(y[o], 4, 64);
  }
  if (q == 0 && r0 + r < n)
    for (uint32_t o = 0; o < O; ++o) Y[(r0 + r) * O + o] = y[o];
}

}  // namespace

void symmetric_eigen(std::vector<double> a, uint64_t D, std::vector<double>* vals, std::vector<double>* vecs) {
  // cyclic Jacobi (Golub & Van Loan, Matrix Computations, 8.5): rotate away
  // each off-diagonal entry in turn until the off-diagonal mass vanishes
  std::vector<double> v(D * D, 0.0);
  for (uint64_t i = 0; i < D; ++i) v[i * D + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (uint64_t i = 0; i < D; ++i) {
      diag += a[i * D + i] * a[i * D + i];
      for (uint64_t j = i + 1; j < D; ++j) off += a[i * D + j] * a[i * D + j];
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (uint64_t p = 0; p + 1 < D; ++p)
      for (uint64_t q = p + 1; q < D; ++q) {
        const double apq = a[p * D + q];
        if (apq == 0.0) continue;
        const double theta = (a[q * D + q] - a[p * D + p]) / (2.0 * apq);
        const double tt = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(tt * tt + 1.0), s = tt * c;
        for (uint64_t k = 0; k < D; ++k) {  // columns p, q
          const double akp = a[k * D + p], akq = a[k * D + q];
          a[k * D + p] = c * akp - s * akq;
          a[k * D + q] = s * akp + c * akq;
        }
        for (uint64_t k = 0; k < D; ++k) {  // rows p, q
          const double apk = a[p * D + k], aqk = a[q * D + k];
          a[p * D + k] = c * apk - s * aqk;
          a[q * D + k] = s * apk + c * aqk;
        }
        for (uint64_t k = 0; k < D; ++k) {  // eigenvector columns
          const double vkp = v[k * D + p], vkq = v[k * D + q];
          v[k * D + p] = c * vkp - s * vkq;
          v[k * D + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<uint64_t> order(D);
  for (uint64_t i = 0; i < D; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) { return a[x * D + x] > a[y * D + y]; });
  vals->assign(D, 0.0);
  vecs->assign(D * D, 0.0);
  for (uint64_t r = 0; r < D; ++r) {
    const uint64_t c = order[r];
    (*vals)[r] = a[c * D + c];
    uint64_t big = 0;
    for (uint64_t k = 0; k < D; ++k)
      if (std::fabs(v[k * D + c]) > std::fabs(v[big * D + c])) big = k;
    const double sg = v[big * D + c] < 0 ? -1.0 : 1.0;
    for (uint64_t k = 0; k < D; ++k) (*vecs)[r * D + k] = sg * v[k * D + c];
  }
}

PcaResult pca_project(const float* X, uint64_t n, uint64_t D, uint64_t out_dim, float* Y, hipStream_t s) {
  if (n < 2) throw Error(Err::InvalidArgument, "PCA needs at least 2 rows");
  if (D < 1 || D > (uint64_t)kMaxD) throw Error(Err::Unsupported, "PCA supports 1 <= D <= 384");
  if (out_dim < 1 || out_dim > std::min<uint64_t>(D, kMaxOut))
    throw Error(Err::InvalidArgument, "PCA output dimension must be in [1, min(D, 8)]");
  PcaResult res;
  // 1. means
  const uint64_t mblocks = std::max<uint64_t>(1, std::min<uint64_t>(1024, ceil_div(n, 1024)));
  const uint64_t mchunk = ceil_div(n, mblocks);
  DevBuf mpart(mblocks * D * sizeof(double));
  col_sum_kernel<<<(unsigned)mblocks, kThreads, 0, s>>>(X, n, (uint32_t)D, mchunk, mpart.as<double>());
  UMAP_LAUNCH_CHECK();
  std::vector<double> hm(mblocks * D);
  UMAP_HIP(hipMemcpyAsync(hm.data(), mpart.p, hm.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  UMAP_HIP(hipStreamSynchronize(s));
  res.mean.assign(D, 0.0);
  for (uint64_t c = 0; c < D; ++c) {
    double sum = 0.0;
    for (uint64_t b = 0; b < mblocks; ++b) sum += hm[b * D + c];
    res.mean[c] = sum / (double)n;
  }
  std::vector<float> meanf(D);
  for (uint64_t c = 0; c < D; ++c) meanf[c] = (float)res.mean[c];
  DevBuf dmean(D * 4);
  UMAP_HIP(hipMemcpyAsync(dmean.p, meanf.data(), D * 4, hipMemcpyHostToDevice, s));
  // 2. covariance tiles
  const uint32_t nt = (uint32_t)ceil_div(D, kTile);
  std::vector<uint32_t> pairs;
  for (uint32_t i = 0; i < nt; ++i)
    for (uint32_t j = i; j < nt; ++j) pairs.push_back((i << 16) | j);
  const uint32_t npairs = (uint32_t)pairs.size();
  DevBuf dpairs(npairs * 4);
  UMAP_HIP(hipMemcpyAsync(dpairs.p, pairs.data(), npairs * 4, hipMemcpyHostToDevice, s));
  const uint32_t cblocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(512, ceil_div(n, kRows)));
  const uint64_t tile_len = (uint64_t)npairs * kTile * kTile;
  DevBuf cpart(cblocks * tile_len * 4), csum(tile_len * 8);
  cov_kernel<<<dim3(cblocks, npairs), kThreads, 0, s>>>(X, n, (uint32_t)D, dmean.as<float>(), dpairs.as<uint32_t>(),
                                                        npairs, cpart.as<float>());
  UMAP_LAUNCH_CHECK();
  sum_blocks_kernel<<<(unsigned)ceil_div(tile_len, kThreads), kThreads, 0, s>>>(cpart.as<float>(), cblocks, tile_len,
                                                                                csum.as<double>());
  UMAP_LAUNCH_CHECK();
  std::vector<double> hc(tile_len);
  UMAP_HIP(hipMemcpyAsync(hc.data(), csum.p, tile_len * 8, hipMemcpyDeviceToHost, s));
  UMAP_HIP(hipStreamSynchronize(s));
  std::vector<double> cov(D * D, 0.0);
  for (uint32_t p = 0; p < npairs; ++p) {
    const uint32_t bi = pairs[p] >> 16, bj = pairs[p] & 0xffffu;
    for (int a = 0; a < kTile; ++a)
      for (int b = 0; b < kTile; ++b) {
        const uint64_t i = (uint64_t)bi * kTile + a, j = (uint64_t)bj * kTile + b;
        if (i < D && j < D) {
          const double v = hc[(uint64_t)p * kTile * kTile + a * kTile + b] / (double)n;
          cov[i * D + j] = v;
          cov[j * D + i] = v;
        }
      }
  }
  // 3. eigenvectors
  std::vector<double> vals, vecs;
  symmetric_eigen(cov, D, &vals, &vecs);
  res.variances.assign(vals.begin(), vals.begin() + out_dim);
  res.components.assign(vecs.begin(), vecs.begin() + out_dim * D);
  // 4. projection
  std::vector<float> vf(out_dim * D);
  for (uint64_t e = 0; e < out_dim * D; ++e) vf[e] = (float)res.components[e];
  DevBuf dv(out_dim * D * 4);
  UMAP_HIP(hipMemcpyAsync(dv.p, vf.data(), vf.size() * 4, hipMemcpyHostToDevice, s));
  const size_t lds = ((size_t)kRows * (D + 1) + D + out_dim * D) * 4;
  project_kernel<<<(unsigned)ceil_div(n, kRows), kThreads, lds, s>>>(X, n, (uint32_t)D, dmean.as<float>(),
                                                                     dv.as<float>(), (uint32_t)out_dim, Y);
  UMAP_LAUNCH_CHECK();
  UMAP_HIP(hipStreamSynchronize(s));
  return res;
}

}  // namespace umapgpu
