// Compares the device libm restatement (umap-rs_amd/csrc/libm_glibc.hpp,
// compiled here as host code) with glibc's expf / powf over float bit ranges.
// Usage: libm_glibc_check exp <first_bits> <last_bits> <stride>
//        libm_glibc_check pow <first_bits> <last_bits> <stride> <y>
// Prints "<tested> <mismatches>".
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "libm_glibc.hpp"

int main(int argc, char** argv) {
  if (argc < 5) return 2;
  const bool is_pow = strcmp(argv[1], "pow") == 0;
  const uint64_t first = strtoull(argv[2], 0, 0), last = strtoull(argv[3], 0, 0), stride = strtoull(argv[4], 0, 0);
  const float y = is_pow ? strtof(argv[5], 0) : 0.0f;
  unsigned long long mism = 0, tested = 0;
#pragma omp parallel for reduction(+ : mism, tested) schedule(static)
  for (uint64_t u = first; u <= last; u += stride) {
    const uint32_t b = (uint32_t)u;
    float x;
    memcpy(&x, &b, 4);
    const float g = is_pow ? powf(x, y) : expf(x);
    const float m = is_pow ? umapgpu::libm_powf(x, y) : umapgpu::libm_expf(x);
    uint32_t gb, mb;
    memcpy(&gb, &g, 4);
    memcpy(&mb, &m, 4);
    tested++;
    if (gb != mb && !(g != g && m != m)) mism++;
  }
  printf("%llu %llu\n", tested, mism);
  return 0;
}
