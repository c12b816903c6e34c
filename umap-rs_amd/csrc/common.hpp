// Shared host/device utilities of libumapgpu: HIP error handling, device
// buffers, the counter RNG, and small device helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

namespace umapgpu {

// ---------------------------------------------------------------------------
// Errors.  Library code throws; the C-ABI layer converts to a status code and
// a last-error message.
enum class Err { InvalidArgument = 1, OutOfMemory = 2, Hip = 3, Overstep = 4, NoDevice = 5, Unsupported = 6 };

struct Error : std::runtime_error {
  Err code;
  Error(Err c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define UMAP_HIP(call)                                                  \
  do {                                                                  \
    hipError_t umap_e_ = (call);                                        \
    if (umap_e_ != hipSuccess) ::umapgpu::throw_hip(umap_e_, #call, __FILE__, __LINE__); \
  } while (0)

#define UMAP_LAUNCH_CHECK() UMAP_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Device buffer (owning, move-only).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n);
  void release();
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------------------
// Counter-based negative-sample RNG: SplitMix64's output function (Steele, Lea
// & Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014)
// evaluated at a counter built from (seed, epoch, edge, sample).  Replaces the
// reference's OS-seeded rand::rng() (layout/optimize_layout_euclidean.rs:269,
// :314), which no run can reproduce; this one makes every draw a pure function
// of its position, so GPU and CPU runs draw the same negatives.
constexpr uint64_t kSmGamma = 0x9e3779b97f4a7c15ULL;
constexpr uint64_t kSmGamma2 = 0xd1b54a32d192ed03ULL;

__host__ __device__ inline uint64_t sm_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rng_epoch_key(uint64_t seed, uint64_t epoch) {
  return sm_mix(seed + (epoch + 1) * kSmGamma);
}
__host__ __device__ inline uint64_t rng_edge_key(uint64_t epoch_key, uint64_t edge) {
  return sm_mix(epoch_key ^ ((edge + 1) * kSmGamma2));
}
// Uniform index in [0, n): floor(r * n / 2^64) from a 64-bit draw r
// (bias below n / 2^64).
__host__ __device__ inline uint32_t rng_index(uint64_t edge_key, uint64_t p, uint32_t n) {
  uint64_t r = sm_mix(edge_key + (p + 1) * kSmGamma);
  uint64_t hi = r >> 32, lo = r & 0xffffffffULL;
  uint64_t t = (lo * (uint64_t)n) >> 32;
  return (uint32_t)((hi * (uint64_t)n + t) >> 32);
}

// clip to [-4, 4] keeping NaN (utils/clip.rs:13-15, f32::clamp)
__host__ __device__ inline float clip4(float v) { return v < -4.0f ? -4.0f : (v > 4.0f ? 4.0f : v); }

// Rust `f64 as usize` saturates: NaN and negatives give 0.  Counts above
// 2^32-1 only arise from degenerate schedules and are bounded there.
__host__ __device__ inline uint32_t sat_count(double q) {
  if (!(q > 0.0)) return 0u;
  if (q >= 4294967295.0) return 0xffffffffu;
  return (uint32_t)q;
}

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace umapgpu
