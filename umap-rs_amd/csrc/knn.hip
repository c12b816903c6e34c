// Exact brute-force Euclidean kNN: the caller-side producer the reference
// leaves to the user (README.md:37-43; examples/src/bin/mnist_demo.rs:203-229
// computes the same thing on the CPU).  Output row i: slot 0 is i itself at
// distance 0, then the k-1 nearest other points by ascending distance, ties
// broken by the lower index.  The query form (evaluators) takes a list of
// query rows and returns their k nearest other points, without the self slot;
// points with a non-finite coordinate are never neighbours.
//
// One thread per query; a block's 256 queries sweep all points in 64-point
// tiles staged in LDS, where every lane reads the same candidate at the same
// time (LDS broadcast).  Squared distances are sum((q-c)^2) in f32, so no
// cancellation; the top list stays sorted in registers and is only touched
// when some lane of the wave finds a closer point.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "knn.hpp"

namespace umapgpu {

namespace {

constexpr int kQ = 256;   // queries per block
constexpr int kTC = 64;   // candidates per LDS tile

// SELF: output slot 0 is the query itself and k-1 neighbours follow; else the
// k nearest others.  queries == nullptr means query t is row t.
template <int DIM, int KM, bool SELF>
__global__ __launch_bounds__(kQ) void knn_kernel(const float* __restrict__ X, uint32_t n, uint32_t k,
                                                 const uint32_t* __restrict__ queries, uint32_t nq,
                                                 uint32_t* __restrict__ out_i, float* __restrict__ out_d) {
  __shared__ float tile[kTC * DIM];
  const uint32_t t = blockIdx.x * kQ + threadIdx.x;
  const uint32_t q = t < nq ? (queries ? queries[t] : t) : 0xffffffffu;
  const bool active = q < n;  // query rows outside [0, n) get no neighbours
  float qv[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) qv[d] = active ? X[(uint64_t)q * DIM + d] : 0.0f;
  float bd[KM];
  uint32_t bi[KM];
#pragma unroll
  for (int i = 0; i < KM; ++i) {
    bd[i] = INFINITY;
    bi[i] = 0xffffffffu;
  }
  const uint32_t kk = SELF ? k - 1 : k;  // neighbours besides self
  float worst = INFINITY;
  for (uint32_t c0 = 0; c0 < n; c0 += kTC) {
    const uint32_t nc = (n - c0) < (uint32_t)kTC ? (n - c0) : (uint32_t)kTC;
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < nc * DIM; e += kQ) tile[e] = X[(uint64_t)c0 * DIM + e];
    __syncthreads();
    for (uint32_t c = 0; c < nc; ++c) {
      const float* cv = tile + c * DIM;
      float acc = 0.0f;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const float df = qv[d] - cv[d];
        acc = acc + df * df;
      }
      const uint32_t ci = c0 + c;
      const bool ins = active && ci != q && acc < worst;
      if (__ballot(ins)) {
        if (ins) {
          // insert into the list sorted by (distance, index)
          float nd = acc;
          uint32_t ni = ci;
#pragma unroll
          for (int i = 0; i < KM; ++i) {
            if (i < (int)kk) {
              const bool sw = nd < bd[i] || (nd == bd[i] && ni < bi[i]);
              const float td = bd[i];
              const uint32_t ti = bi[i];
              bd[i] = sw ? nd : td;
              bi[i] = sw ? ni : ti;
              nd = sw ? td : nd;
              ni = sw ? ti : ni;
            }
          }
          worst = (kk > 0) ? bd[0] : 0.0f;
#pragma unroll
          for (int i = 0; i < KM; ++i)
            if (i == (int)kk - 1) worst = bd[i];
        }
      }
    }
  }
  if (t >= nq) return;
  const uint64_t o = (uint64_t)t * k;
  if (SELF) {
    out_i[o] = q;
    if (out_d) out_d[o] = 0.0f;
  }
#pragma unroll
  for (int i = 0; i < KM; ++i)
    if (i < (int)kk) {
      out_i[o + (SELF ? 1 : 0) + i] = bi[i];
      if (out_d) out_d[o + (SELF ? 1 : 0) + i] = sqrtf(bd[i]);
    }
}

template <int DIM, bool SELF>
void launch_dim(const float* X, uint64_t n, uint64_t k, const uint32_t* queries, uint64_t nq, uint32_t* oi,
                float* od, hipStream_t s) {
  if (nq == 0) return;
  const unsigned grid = (unsigned)ceil_div(nq, kQ);
  if (k <= 16)
    knn_kernel<DIM, 16, SELF><<<grid, kQ, 0, s>>>(X, (uint32_t)n, (uint32_t)k, queries, (uint32_t)nq, oi, od);
  else if (k <= 32)
    knn_kernel<DIM, 32, SELF><<<grid, kQ, 0, s>>>(X, (uint32_t)n, (uint32_t)k, queries, (uint32_t)nq, oi, od);
  else
    throw Error(Err::Unsupported, "brute-force kNN supports k <= 32");
  UMAP_LAUNCH_CHECK();
}

}  // namespace

void knn_bruteforce(const float* X, uint64_t n, uint64_t dim, uint64_t k, uint32_t* oi, float* od, hipStream_t s) {
  if (k < 1 || k > n) throw Error(Err::InvalidArgument, "kNN needs 1 <= k <= n");
  if (n >= 0xffffffffull) throw Error(Err::InvalidArgument, "kNN needs n < 2^32");
  switch (dim) {
    case 2: launch_dim<2, true>(X, n, k, nullptr, n, oi, od, s); break;
    case 3: launch_dim<3, true>(X, n, k, nullptr, n, oi, od, s); break;
    case 8: launch_dim<8, true>(X, n, k, nullptr, n, oi, od, s); break;
    case 16: launch_dim<16, true>(X, n, k, nullptr, n, oi, od, s); break;
    case 20: launch_dim<20, true>(X, n, k, nullptr, n, oi, od, s); break;
    case 32: launch_dim<32, true>(X, n, k, nullptr, n, oi, od, s); break;
    case 64: launch_dim<64, true>(X, n, k, nullptr, n, oi, od, s); break;
    case 128: launch_dim<128, true>(X, n, k, nullptr, n, oi, od, s); break;
    default: throw Error(Err::Unsupported, "brute-force kNN supports dim in {2,3,8,16,20,32,64,128}");
  }
  UMAP_HIP(hipStreamSynchronize(s));
}

void knn_query(const float* X, uint64_t n, uint64_t dim, const uint32_t* queries, uint64_t nq, uint64_t k,
               uint32_t* oi, hipStream_t s) {
  if (k < 1 || k >= n) throw Error(Err::InvalidArgument, "kNN query needs 1 <= k < n");
  if (n >= 0xffffffffull || nq >= 0xffffffffull) throw Error(Err::InvalidArgument, "kNN needs n < 2^32");
  switch (dim) {
    case 1: launch_dim<1, false>(X, n, k, queries, nq, oi, nullptr, s); break;
    case 2: launch_dim<2, false>(X, n, k, queries, nq, oi, nullptr, s); break;
    case 3: launch_dim<3, false>(X, n, k, queries, nq, oi, nullptr, s); break;
    case 4: launch_dim<4, false>(X, n, k, queries, nq, oi, nullptr, s); break;
    case 8: launch_dim<8, false>(X, n, k, queries, nq, oi, nullptr, s); break;
    default: throw Error(Err::Unsupported, "kNN query supports dim in {1,2,3,4,8}");
  }
  UMAP_HIP(hipStreamSynchronize(s));
}

}  // namespace umapgpu
