// Exact brute-force Euclidean kNN: the caller-side producer the reference
// leaves to the user (README.md:37-43; examples/src/bin/mnist_demo.rs:203-229
// computes the same thing on the CPU).  Output row i: slot 0 is i itself at
// distance 0, then the k-1 nearest other points by ascending distance, ties
// broken by the lower index.
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

template <int DIM, int KM>
__global__ __launch_bounds__(kQ) void knn_kernel(const float* __restrict__ X, uint32_t n, uint32_t k,
                                                 uint32_t* __restrict__ out_i, float* __restrict__ out_d) {
  __shared__ float tile[kTC * DIM];
  const uint32_t q = blockIdx.x * kQ + threadIdx.x;
  const bool active = q < n;
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
  const uint32_t kk = k - 1;  // neighbours besides self
  float worst = INFINITY;
  for (uint32_t c0 = 0; c0 < n; c0 += kTC) {
    const uint32_t nc = (n - c0) < (uint32_t)kTC ? (n - c0) : (uint32_t)kTC;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < nc * DIM; t += kQ) tile[t] = X[(uint64_t)c0 * DIM + t];
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
  if (!active) return;
  out_i[(uint64_t)q * k] = q;
  out_d[(uint64_t)q * k] = 0.0f;
#pragma unroll
  for (int i = 0; i < KM; ++i)
    if (i < (int)kk) {
      out_i[(uint64_t)q * k + 1 + i] = bi[i];
      out_d[(uint64_t)q * k + 1 + i] = sqrtf(bd[i]);
    }
}

template <int DIM>
void launch_dim(const float* X, uint64_t n, uint64_t k, uint32_t* oi, float* od, hipStream_t s) {
  const unsigned grid = (unsigned)ceil_div(n, kQ);
  if (k <= 16)
    knn_kernel<DIM, 16><<<grid, kQ, 0, s>>>(X, (uint32_t)n, (uint32_t)k, oi, od);
  else if (k <= 32)
    knn_kernel<DIM, 32><<<grid, kQ, 0, s>>>(X, (uint32_t)n, (uint32_t)k, oi, od);
  else
    throw Error(Err::Unsupported, "brute-force kNN supports k <= 32");
  UMAP_LAUNCH_CHECK();
}

}  // namespace

void knn_bruteforce(const float* X, uint64_t n, uint64_t dim, uint64_t k, uint32_t* oi, float* od, hipStream_t s) {
  if (k < 1 || k > n) throw Error(Err::InvalidArgument, "kNN needs 1 <= k <= n");
  if (n >= 0xffffffffull) throw Error(Err::InvalidArgument, "kNN needs n < 2^32");
  switch (dim) {
    case 2: launch_dim<2>(X, n, k, oi, od, s); break;
    case 3: launch_dim<3>(X, n, k, oi, od, s); break;
    case 8: launch_dim<8>(X, n, k, oi, od, s); break;
    case 16: launch_dim<16>(X, n, k, oi, od, s); break;
    case 20: launch_dim<20>(X, n, k, oi, od, s); break;
    case 32: launch_dim<32>(X, n, k, oi, od, s); break;
    case 64: launch_dim<64>(X, n, k, oi, od, s); break;
    case 128: launch_dim<128>(X, n, k, oi, od, s); break;
    default: throw Error(Err::Unsupported, "brute-force kNN supports dim in {2,3,8,16,20,32,64,128}");
  }
  UMAP_HIP(hipStreamSynchronize(s));
}

}  // namespace umapgpu
