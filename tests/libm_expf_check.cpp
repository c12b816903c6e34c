// Compares umapgpu::libm_expf (the device expf restatement, compiled here as
// host code) with glibc's expf over a range of float bit patterns.
// Usage: libm_expf_check <first_bits> <last_bits> <stride>; prints mismatches.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "libm_expf.hpp"

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  const uint64_t first = strtoull(argv[1], 0, 0), last = strtoull(argv[2], 0, 0), stride = strtoull(argv[3], 0, 0);
  unsigned long long mism = 0, tested = 0;
#pragma omp parallel for reduction(+ : mism, tested) schedule(static)
  for (uint64_t u = first; u <= last; u += stride) {
    uint32_t b = (uint32_t)u;
    float x;
    memcpy(&x, &b, 4);
    float g = expf(x), m = umapgpu::libm_expf(x);
    uint32_t gb, mb;
    memcpy(&gb, &g, 4);
    memcpy(&mb, &m, 4);
    tested++;
    if (gb != mb && !(g != g && m != m)) mism++;
  }
  printf("%llu %llu\n", tested, mism);
  return 0;
}
