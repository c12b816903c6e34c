// Single-precision exp that reproduces glibc's expf bit for bit.
//
// The reference evaluates every membership weight and every sigma-bisection
// term with Rust's f32::exp, which on x86-64 Linux is glibc's expf; glibc picks
// its FMA build on every CPU with FMA (the reference's EPYC and this project's
// hosts).  That routine is the published table-driven method of Szabolcs Nagy
// (ARM optimized-routines, "exp2f/expf with 32-entry table", also in glibc
// sysdeps/ieee754/flt-32/e_expf.c): x*32/ln2 = k + r, exp(x) = 2^(k/32) *
// P(r) with a cubic P evaluated in double.  Reimplemented here from that
// description:
//   * the table holds the bits of 2^(i/32) rounded to double, minus i<<47,
//     computed from the definition (verified: all 32 entries equal the
//     correctly rounded powers);
//   * the cubic's coefficients are the method's published minimax values;
//   * every a*b+c of the FMA build is an explicit fma().
// Agreement with glibc is checked exhaustively on the host over every float
// in [-103.9, 88.5] (tests/test_libm_expf.py) and on the GPU against glibc
// values (tests/test_gpu_graph.py).  Compiles as host and device code.
#pragma once
#include <stdint.h>
#include <string.h>

#ifdef __HIPCC__
#define UMAP_HD __host__ __device__
#else
#define UMAP_HD
#include <math.h>
#endif

namespace umapgpu {

UMAP_HD inline double libm_bits_to_double(uint64_t u) {
  double d;
  memcpy(&d, &u, sizeof d);
  return d;
}
UMAP_HD inline uint64_t libm_double_to_bits(double d) {
  uint64_t u;
  memcpy(&u, &d, sizeof u);
  return u;
}

UMAP_HD inline uint64_t libm_exp2_table(uint32_t i) {
  // bits(2^(i/32)) - (i << 47)
  switch (i & 31u) {
    case 0: return 0x3ff0000000000000ULL;
    case 1: return 0x3fefd9b0d3158574ULL;
    case 2: return 0x3fefb5586cf9890fULL;
    case 3: return 0x3fef9301d0125b51ULL;
    case 4: return 0x3fef72b83c7d517bULL;
    case 5: return 0x3fef54873168b9aaULL;
    case 6: return 0x3fef387a6e756238ULL;
    case 7: return 0x3fef1e9df51fdee1ULL;
    case 8: return 0x3fef06fe0a31b715ULL;
    case 9: return 0x3feef1a7373aa9cbULL;
    case 10: return 0x3feedea64c123422ULL;
    case 11: return 0x3feece086061892dULL;
    case 12: return 0x3feebfdad5362a27ULL;
    case 13: return 0x3feeb42b569d4f82ULL;
    case 14: return 0x3feeab07dd485429ULL;
    case 15: return 0x3feea47eb03a5585ULL;
    case 16: return 0x3feea09e667f3bcdULL;
    case 17: return 0x3fee9f75e8ec5f74ULL;
    case 18: return 0x3feea11473eb0187ULL;
    case 19: return 0x3feea589994cce13ULL;
    case 20: return 0x3feeace5422aa0dbULL;
    case 21: return 0x3feeb737b0cdc5e5ULL;
    case 22: return 0x3feec49182a3f090ULL;
    case 23: return 0x3feed503b23e255dULL;
    case 24: return 0x3feee89f995ad3adULL;
    case 25: return 0x3feeff76f2fb5e47ULL;
    case 26: return 0x3fef199bdd85529cULL;
    case 27: return 0x3fef3720dcef9069ULL;
    case 28: return 0x3fef5818dcfba487ULL;
    case 29: return 0x3fef7c97337b9b5fULL;
    case 30: return 0x3fefa4afa2a490daULL;
    default: return 0x3fefd0765b6e4540ULL;
  }
}

UMAP_HD inline float libm_expf(float x) {
  // Special cases: glibc returns 0 below log(2^-150), inf above log(2^128),
  // and x+x for NaN (+inf stays +inf, -inf gives 0).
  if (!(x == x)) return x + x;
  if (x < -0x1.9fe368p6f) return 0.0f;
  if (x > 0x1.62e42ep6f) return (float)INFINITY;
  const double N = 32.0;
  const double inv_ln2_n = 0x1.71547652b82fep+0 * N;
  const double shift = 0x1.8p+52;
  const double c0 = 0x1.c6af84b912394p-5 / N / N / N;
  const double c1 = 0x1.ebfce50fac4f3p-3 / N / N;
  const double c2 = 0x1.62e42ff0c52d6p-1 / N;
  double xd = (double)x;
  double kd = fma(inv_ln2_n, xd, shift);
  uint64_t ki = libm_double_to_bits(kd);
  kd -= shift;
  double r = fma(inv_ln2_n, xd, -kd);
  uint64_t t = libm_exp2_table((uint32_t)(ki % 32u));
  t += ki << (52 - 5);
  double s = libm_bits_to_double(t);
  double z = fma(c0, r, c1);
  double r2 = r * r;
  double y = fma(c2, r, 1.0);
  y = fma(z, r2, y);
  y = y * s;
  return (float)y;
}

}  // namespace umapgpu
