// Embedding-quality evaluators on the device.  The reference defines no
// quality measure; these are the ones the fit path's quality is measured by, computed on
// the GPU so that they stay cheap at millions of points:
//
// * fuzzy cross-entropy of an embedding against the symmetric fuzzy graph
//     sum_{(i,j) in graph} w log(w/v) + (1-w) log((1-w)/(1-v))
//   + (n^2 - nnz) / M * sum_{M random pairs} -log(1-v),
//   v = 1/(1 + a*|y_i-y_j|^(2b)), w and v clamped to [1e-12, 1-1e-12], all
//   in f64.  The M pairs come from the counter RNG (seeded), so a CPU
//   restatement draws the same ones.  Rows with a non-finite coordinate are
//   skipped, as are pairs with i == j.
// * trustworthiness ranks: for each query row q and each of its low-dimensional
//   neighbours l, the position of l in the ordering of all points by
//   (squared high-dimensional distance to q, index), q itself at position 0.
//
// Reductions go through fixed-size per-block partials summed in a fixed
// order, so results do not depend on scheduling.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>

#include "evaluate.hpp"

namespace umapgpu {

namespace {

constexpr int kThreads = 256;
constexpr unsigned kBlocks = 2048;  // partial slots per reduction

__device__ inline bool row_finite(const float* __restrict__ y, uint64_t i, uint32_t dim) {
  bool ok = true;
  for (uint32_t d = 0; d < dim; ++d) ok = ok && isfinite(y[i * dim + d]);
  return ok;
}

__device__ inline double clamp_p(double v) { return v < 1e-12 ? 1e-12 : (v > 1.0 - 1e-12 ? 1.0 - 1e-12 : v); }

__device__ inline double low_dim_v(const float* __restrict__ y, uint64_t i, uint64_t j, uint32_t dim, double a,
                                   double b) {
  double d2 = 0.0;
  for (uint32_t d = 0; d < dim; ++d) {
    const double df = (double)y[i * dim + d] - (double)y[j * dim + d];
    d2 += df * df;
  }
  return clamp_p(1.0 / (1.0 + a * pow(d2, b)));
}

template <typename T>
__device__ inline T block_sum(T v, T* lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kThreads / 64; ++i) r += lds[i];
  return r;
}

// Lane per row: the attractive term over the row's edges.
__global__ __launch_bounds__(kThreads) void ce_attract_kernel(const uint64_t* __restrict__ indptr,
                                                              const uint32_t* __restrict__ indices,
                                                              const float* __restrict__ data,
                                                              const float* __restrict__ y, uint64_t n,
                                                              uint32_t dim, double a, double b,
                                                              double* __restrict__ part_sum,
                                                              unsigned long long* __restrict__ part_cnt) {
  __shared__ double lds_d[kThreads / 64];
  __shared__ unsigned long long lds_u[kThreads / 64];
  double acc = 0.0;
  unsigned long long cnt = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
    if (!row_finite(y, i, dim)) continue;
    for (uint64_t e = indptr[i]; e < indptr[i + 1]; ++e) {
      const uint64_t j = indices[e];
      if (!row_finite(y, j, dim)) continue;
      const double w = clamp_p((double)data[e]);
      const double v = low_dim_v(y, i, j, dim, a, b);
      acc += w * log(w / v) + (1.0 - w) * log((1.0 - w) / (1.0 - v));
      ++cnt;
    }
  }
  const double s = block_sum(acc, lds_d);
  const unsigned long long c = block_sum(cnt, lds_u);
  if (threadIdx.x == 0) {
    part_sum[blockIdx.x] = s;
    part_cnt[blockIdx.x] = c;
  }
}

// The repulsive sample: pair p is (rng_index(key, 2p, n), rng_index(key, 2p+1, n)).
__global__ __launch_bounds__(kThreads) void ce_repulse_kernel(const float* __restrict__ y, uint64_t n,
                                                              uint32_t dim, double a, double b, uint64_t key,
                                                              uint64_t m, double* __restrict__ part_sum,
                                                              unsigned long long* __restrict__ part_cnt) {
  __shared__ double lds_d[kThreads / 64];
  __shared__ unsigned long long lds_u[kThreads / 64];
  double acc = 0.0;
  unsigned long long cnt = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x; p < m; p += (uint64_t)gridDim.x * kThreads) {
    const uint64_t i = rng_index(key, 2 * p, (uint32_t)n);
    const uint64_t j = rng_index(key, 2 * p + 1, (uint32_t)n);
    if (i == j || !row_finite(y, i, dim) || !row_finite(y, j, dim)) continue;
    acc += -log(1.0 - low_dim_v(y, i, j, dim, a, b));
    ++cnt;
  }
  const double s = block_sum(acc, lds_d);
  const unsigned long long c = block_sum(cnt, lds_u);
  if (threadIdx.x == 0) {
    part_sum[blockIdx.x] = s;
    part_cnt[blockIdx.x] = c;
  }
}

// One block: out[0] = sum of part_sum[0..np), out[1] = sum of part_cnt (as double).
__global__ __launch_bounds__(kThreads) void sum_partials_kernel(const double* __restrict__ part_sum,
                                                                const unsigned long long* __restrict__ part_cnt,
                                                                unsigned np, double* __restrict__ out_sum,
                                                                unsigned long long* __restrict__ out_cnt) {
  __shared__ double lds_d[kThreads / 64];
  __shared__ unsigned long long lds_u[kThreads / 64];
  double acc = 0.0;
  unsigned long long cnt = 0;
  for (unsigned i = threadIdx.x; i < np; i += kThreads) {
    acc += part_sum[i];
    cnt += part_cnt[i];
  }
  const double s = block_sum(acc, lds_d);
  const unsigned long long c = block_sum(cnt, lds_u);
  if (threadIdx.x == 0) {
    *out_sum = s;
    *out_cnt = c;
  }
}

// ---------------------------------------------------------------------------
// trustworthiness ranks
//
// A block takes kQB queries and one slice of the points.  Per query the k
// neighbour keys (d2(q,l), l) are sorted ascending in LDS; each point p != q
// adds one to bin pos(p) = #{neighbour keys below (d2(q,p), p)} unless it lies
// beyond the last key.  A neighbour's rank is then the running sum of the bins
// up to its own position, since the neighbour itself falls in its own bin.
// Bins go to global memory with one atomic per (query, bin) per block.

constexpr int kQB = 32;     // queries per block
constexpr int kMaxK = 32;   // neighbours per query
constexpr int kMaxDim = 256;

template <int DIM>
__device__ inline float sq_dist(const float* __restrict__ a, const float* __restrict__ b) {
  float acc = 0.0f;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const float df = a[d] - b[d];
    acc = acc + df * df;
  }
  return acc;
}

__device__ inline float sq_dist_dyn(const float* __restrict__ a, const float* __restrict__ b, uint32_t dim) {
  float acc = 0.0f;
  for (uint32_t d = 0; d < dim; ++d) {
    const float df = a[d] - b[d];
    acc = acc + df * df;
  }
  return acc;
}

__device__ inline bool key_less(float da, uint32_t ia, float db, uint32_t ib) {
  return da < db || (da == db && ia < ib);
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void trust_bins_kernel(const float* __restrict__ X, uint32_t n,
                                                              uint32_t dim, const uint32_t* __restrict__ queries,
                                                              uint32_t nq, const uint32_t* __restrict__ nbrs,
                                                              uint32_t k, uint32_t slice,
                                                              unsigned int* __restrict__ bins) {
  constexpr int D = DIM > 0 ? DIM : kMaxDim;
  __shared__ float qx[kQB][D + 1];
  __shared__ uint32_t qid[kQB];
  __shared__ float key_d[kQB][kMaxK];
  __shared__ uint32_t key_i[kQB][kMaxK];
  __shared__ unsigned int hist[kQB][kMaxK + 1];
  const uint32_t q0 = blockIdx.x * kQB;
  const uint32_t nqb = (nq - q0) < (uint32_t)kQB ? (nq - q0) : (uint32_t)kQB;
  const uint32_t ddim = DIM > 0 ? (uint32_t)DIM : dim;

  if (threadIdx.x < nqb) qid[threadIdx.x] = queries[q0 + threadIdx.x];
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < nqb * ddim; e += kThreads) {
    const uint32_t r = e / ddim, d = e % ddim;
    qx[r][d] = qid[r] < n ? X[(uint64_t)qid[r] * ddim + d] : 0.0f;
  }
  for (uint32_t e = threadIdx.x; e < (uint32_t)kQB * (kMaxK + 1); e += kThreads) (&hist[0][0])[e] = 0u;
  __syncthreads();
  // neighbour keys, then an insertion sort per query (k <= 32, one lane each)
  for (uint32_t e = threadIdx.x; e < nqb * k; e += kThreads) {
    const uint32_t r = e / k, j = e % k;
    const uint32_t l = nbrs[(uint64_t)(q0 + r) * k + j];
    float dl = INFINITY;
    if (l < n)
      dl = DIM > 0 ? sq_dist<D>(&qx[r][0], X + (uint64_t)l * ddim) : sq_dist_dyn(&qx[r][0], X + (uint64_t)l * ddim, ddim);
    key_d[r][j] = dl;
    key_i[r][j] = l;
  }
  __syncthreads();
  if (threadIdx.x < nqb) {
    const uint32_t r = threadIdx.x;
    for (uint32_t j = 1; j < k; ++j) {
      const float d = key_d[r][j];
      const uint32_t i = key_i[r][j];
      uint32_t t = j;
      while (t > 0 && key_less(d, i, key_d[r][t - 1], key_i[r][t - 1])) {
        key_d[r][t] = key_d[r][t - 1];
        key_i[r][t] = key_i[r][t - 1];
        --t;
      }
      key_d[r][t] = d;
      key_i[r][t] = i;
    }
  }
  __syncthreads();

  const uint64_t p_begin = (uint64_t)blockIdx.y * slice;
  const uint64_t p_end = p_begin + slice < n ? p_begin + slice : n;
  auto bin = [&](uint32_t r, float dp, uint32_t p) {
    const float dl = key_d[r][k - 1];
    const uint32_t il = key_i[r][k - 1];
    if (!(key_less(dp, p, dl, il) || (dp == dl && p == il))) return;  // beyond every neighbour
    uint32_t pos = 0;
    while (pos < k && key_less(key_d[r][pos], key_i[r][pos], dp, p)) ++pos;
    atomicAdd(&hist[r][pos], 1u);
  };
  for (uint64_t p = p_begin + threadIdx.x; p < p_end; p += kThreads) {
    if constexpr (DIM > 0) {
      float px[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) px[d] = X[p * DIM + d];
      for (uint32_t r = 0; r < nqb; ++r) {
        if ((uint32_t)p == qid[r] || qid[r] >= n) continue;
        float acc = 0.0f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const float df = qx[r][d] - px[d];
          acc = acc + df * df;
        }
        bin(r, acc, (uint32_t)p);
      }
    } else {
      for (uint32_t r = 0; r < nqb; ++r) {
        if ((uint32_t)p == qid[r] || qid[r] >= n) continue;
        bin(r, sq_dist_dyn(&qx[r][0], X + p * ddim, ddim), (uint32_t)p);
      }
    }
  }
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < nqb * k; e += kThreads) {
    const uint32_t r = e / k, j = e % k;
    const unsigned int h = hist[r][j];
    if (h) atomicAdd(&bins[(uint64_t)(q0 + r) * k + j], h);
  }
}

// rank of the j-th sorted neighbour = sum of bins 0..j; penalty per query.
__global__ __launch_bounds__(kThreads) void trust_penalty_kernel(const unsigned int* __restrict__ bins,
                                                                 uint32_t nq, uint32_t k, uint32_t kk,
                                                                 double* __restrict__ part_sum,
                                                                 unsigned long long* __restrict__ part_cnt) {
  __shared__ double lds_d[kThreads / 64];
  __shared__ unsigned long long lds_u[kThreads / 64];
  double acc = 0.0;
  for (uint32_t q = blockIdx.x * kThreads + threadIdx.x; q < nq; q += gridDim.x * kThreads) {
    unsigned long long run = 0;
    for (uint32_t j = 0; j < k; ++j) {
      run += bins[(uint64_t)q * k + j];
      if (run > kk) acc += (double)(run - kk);
    }
  }
  const double s = block_sum(acc, lds_d);
  const unsigned long long c = block_sum(0ull, lds_u);
  if (threadIdx.x == 0) {
    part_sum[blockIdx.x] = s;
    part_cnt[blockIdx.x] = c;
  }
}

struct Partials {
  DevBuf sum, cnt, out;
  Partials() : sum(kBlocks * sizeof(double)), cnt(kBlocks * sizeof(unsigned long long)), out(16) {}
  void total(unsigned np, double* s_out, unsigned long long* c_out, hipStream_t s) {
    sum_partials_kernel<<<1, kThreads, 0, s>>>(sum.as<double>(), cnt.as<unsigned long long>(), np,
                                               out.as<double>(), reinterpret_cast<unsigned long long*>(out.as<double>() + 1));
    UMAP_LAUNCH_CHECK();
    double h[2];
    UMAP_HIP(hipMemcpyAsync(h, out.p, 16, hipMemcpyDeviceToHost, s));
    UMAP_HIP(hipStreamSynchronize(s));
    *s_out = h[0];
    unsigned long long c;
    memcpy(&c, &h[1], 8);
    *c_out = c;
  }
};

}  // namespace

uint64_t ce_pair_key(uint64_t seed) { return sm_mix(seed + kSmGamma2); }

CeTerms fuzzy_cross_entropy(const uint64_t* indptr, const uint32_t* indices, const float* data, uint64_t n,
                            const float* y, uint64_t dim, float a, float b, uint64_t n_random, uint64_t seed,
                            hipStream_t s) {
  if (dim == 0) throw Error(Err::InvalidArgument, "embedding dimension must be positive");
  if (n >= 0xffffffffull) throw Error(Err::InvalidArgument, "evaluator needs n < 2^32");
  CeTerms r;
  r.n = n;
  if (n == 0) return r;
  Partials P;
  const unsigned grid_a = (unsigned)std::min<uint64_t>(kBlocks, ceil_div(n, kThreads));
  ce_attract_kernel<<<grid_a, kThreads, 0, s>>>(indptr, indices, data, y, n, (uint32_t)dim, (double)a, (double)b,
                                                P.sum.as<double>(), P.cnt.as<unsigned long long>());
  UMAP_LAUNCH_CHECK();
  unsigned long long c = 0;
  P.total(grid_a, &r.attract, &c, s);
  r.edges = c;
  if (n_random > 0) {
    const unsigned grid_r = (unsigned)std::min<uint64_t>(kBlocks, ceil_div(n_random, kThreads));
    ce_repulse_kernel<<<grid_r, kThreads, 0, s>>>(y, n, (uint32_t)dim, (double)a, (double)b, ce_pair_key(seed),
                                                  n_random, P.sum.as<double>(), P.cnt.as<unsigned long long>());
    UMAP_LAUNCH_CHECK();
    P.total(grid_r, &r.repulse_sum, &c, s);
    r.pairs = c;
  }
  const double nn = (double)n * (double)n;
  r.total = r.attract + (r.pairs ? (nn - (double)r.edges) / (double)r.pairs * r.repulse_sum : 0.0);
  return r;
}

double trustworthiness_penalty(const float* X, uint64_t n, uint64_t dim, const uint32_t* queries, uint64_t nq,
                               const uint32_t* nbrs, uint64_t k, hipStream_t s) {
  if (k < 1 || k > (uint64_t)kMaxK) throw Error(Err::Unsupported, "trustworthiness supports 1 <= k <= 32");
  if (dim < 1 || dim > (uint64_t)kMaxDim) throw Error(Err::Unsupported, "trustworthiness supports dim <= 256");
  if (n >= 0xffffffffull || nq >= 0xffffffffull) throw Error(Err::InvalidArgument, "evaluator needs n < 2^32");
  if (nq == 0) return 0.0;
  DevBuf bins(nq * k * sizeof(unsigned int));
  UMAP_HIP(hipMemsetAsync(bins.p, 0, bins.bytes, s));
  // slice points so that the grid holds a few thousand blocks
  const uint64_t qblocks = ceil_div(nq, kQB);
  uint64_t slices = std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(4096, qblocks), ceil_div(n, 4096)));
  const uint32_t slice = (uint32_t)ceil_div(n, slices);
  slices = ceil_div(n, slice);
  dim3 grid((unsigned)qblocks, (unsigned)slices);
#define UMAP_TRUST(DD)                                                                                       \
  trust_bins_kernel<DD><<<grid, kThreads, 0, s>>>(X, (uint32_t)n, (uint32_t)dim, queries, (uint32_t)nq, nbrs, \
                                                  (uint32_t)k, slice, bins.as<unsigned int>())
  switch (dim) {
    case 2: UMAP_TRUST(2); break;
    case 16: UMAP_TRUST(16); break;
    case 32: UMAP_TRUST(32); break;
    case 64: UMAP_TRUST(64); break;
    case 128: UMAP_TRUST(128); break;
    default: UMAP_TRUST(0); break;
  }
#undef UMAP_TRUST
  UMAP_LAUNCH_CHECK();
  Partials P;
  const unsigned grid_p = (unsigned)std::min<uint64_t>(kBlocks, ceil_div(nq, kThreads));
  trust_penalty_kernel<<<grid_p, kThreads, 0, s>>>(bins.as<unsigned int>(), (uint32_t)nq, (uint32_t)k,
                                                   (uint32_t)k, P.sum.as<double>(), P.cnt.as<unsigned long long>());
  UMAP_LAUNCH_CHECK();
  double pen = 0.0;
  unsigned long long c = 0;
  P.total(grid_p, &pen, &c, s);
  return pen;
}

}  // namespace umapgpu
