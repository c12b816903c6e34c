// The reference's checkpoint bytes: bincode 1.3 (`bincode::serialize` /
// `bincode::deserialize`, the crate's dev-dependency, umap-rs/Cargo.toml:24)
// over the serde derives of Optimizer (optimizer.rs:27-56), LearnedManifold
// (manifold.rs:14-38), FittedUmap (embedding.rs:438-443) and UmapConfig
// (config.rs:8-235).  Host-only.
//
// bincode 1.3's default (legacy) encoding, as its documentation states:
//   integers fixed width, little endian; usize as u64;
//   f32 / f64 as their IEEE bits, little endian;
//   bool as one byte 0 / 1; Option as one tag byte 0 (None) / 1 (Some) then
//   the value; enum as a u32 variant index, then the variant's fields;
//   struct and tuple fields in declaration order, nothing between them;
//   sequences as a u64 length, then the elements.
// ndarray 0.17 (feature "serde") writes an array as the struct
// {v: u8 = 1, dim: [usize; D] (a tuple), data: sequence in logical order};
// sprs 0.11 (feature "serde") writes CsMatI<f32, u32, usize> as
// {storage: CompressedStorage (CSR = variant 0, CSC = 1), nrows: usize,
//  ncols: usize, indptr: IndPtr {storage: Vec<usize>}, indices: Vec<u32>,
//  data: Vec<f32>}.
// The decoder rejects what those crates' deserializers reject: a truncated
// input, an array version other than 1, an array whose data length is not
// its shape's product, a bool or Option tag other than 0 / 1, an unknown enum
// variant, and a graph that is not a valid CSR matrix.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "umapgpu.h"

namespace umapgpu {
namespace bincode {

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "bincode's integers are little endian");

// LearnedManifold (manifold.rs:15-38)
struct Manifold {
  uint64_t nrows = 0, ncols = 0;  // the graph's shape (n_vertices x n_vertices)
  std::vector<uint64_t> indptr;
  std::vector<uint32_t> indices;
  std::vector<float> data;
  std::vector<float> sigmas, rhos;
  uint64_t n_vertices = 0;
  float a = 0, b = 0;
};

// Optimizer (optimizer.rs:28-56)
struct Optimizer {
  Manifold manifold;
  std::vector<uint32_t> head, tail;
  std::vector<double> epochs_per_sample;
  uint64_t rows = 0, cols = 0;  // embedding shape
  std::vector<float> embedding;
  std::vector<double> epoch_of_next_sample, epoch_of_next_negative_sample, epochs_per_negative_sample;
  uint64_t current_epoch = 0, total_epochs = 0;
  float gamma = 0, initial_alpha = 0;
  double negative_sample_rate = 0;
  uint32_t metric_type = 0;  // MetricType (metric.rs:7-12): Euclidean = 0, Generic = 1
};

// FittedUmap (embedding.rs:439-443)
struct Fitted {
  uint64_t rows = 0, cols = 0;
  std::vector<float> embedding;
  Manifold manifold;
  umapgpu_config config{};
};

class Writer {
 public:
  std::vector<uint8_t> out;
  void raw(const void* p, uint64_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    out.insert(out.end(), b, b + n);
  }
  void u8(uint8_t v) { raw(&v, 1); }
  void u32(uint32_t v) { raw(&v, 4); }
  void u64(uint64_t v) { raw(&v, 8); }
  void f32(float v) { raw(&v, 4); }
  void f64(double v) { raw(&v, 8); }
  void boolean(bool v) { u8(v ? 1 : 0); }
  void opt_f32(bool has, float v) {
    u8(has ? 1 : 0);
    if (has) f32(v);
  }
  void opt_u64(bool has, uint64_t v) {
    u8(has ? 1 : 0);
    if (has) u64(v);
  }
  template <typename T>
  void vec(const std::vector<T>& v) {
    u64(v.size());
    raw(v.data(), v.size() * sizeof(T));
  }
  // ndarray Array1<T>
  template <typename T>
  void array1(const std::vector<T>& v) {
    u8(1);
    u64(v.size());
    vec(v);
  }
  // ndarray Array2<T>, row-major
  template <typename T>
  void array2(const std::vector<T>& v, uint64_t rows, uint64_t cols) {
    u8(1);
    u64(rows);
    u64(cols);
    vec(v);
  }
};

class Reader {
 public:
  Reader(const uint8_t* p, uint64_t n) : p_(p), n_(n) {}
  uint64_t remaining() const { return n_ - pos_; }
  void raw(void* dst, uint64_t n, const char* what) {
    if (remaining() < n) fail(std::string("unexpected end of input while reading ") + what);
    std::memcpy(dst, p_ + pos_, n);
    pos_ += n;
  }
  uint8_t u8(const char* what) {
    uint8_t v;
    raw(&v, 1, what);
    return v;
  }
  uint32_t u32(const char* what) {
    uint32_t v;
    raw(&v, 4, what);
    return v;
  }
  uint64_t u64(const char* what) {
    uint64_t v;
    raw(&v, 8, what);
    return v;
  }
  float f32(const char* what) {
    float v;
    raw(&v, 4, what);
    return v;
  }
  double f64(const char* what) {
    double v;
    raw(&v, 8, what);
    return v;
  }
  bool boolean(const char* what) {
    const uint8_t v = u8(what);
    if (v > 1) fail(std::string("invalid value: integer `") + std::to_string(v) + "`, expected a boolean (" + what + ")");
    return v == 1;
  }
  bool option_tag(const char* what) {
    const uint8_t v = u8(what);
    if (v > 1) fail(std::string("invalid Option tag ") + std::to_string(v) + " (" + what + ")");
    return v == 1;
  }
  template <typename T>
  std::vector<T> vec(const char* what) {
    const uint64_t len = u64(what);
    if (len > remaining() / sizeof(T)) fail(std::string("unexpected end of input while reading ") + what);
    std::vector<T> v(len);
    raw(v.data(), len * sizeof(T), what);
    return v;
  }
  template <typename T>
  std::vector<T> array1(const char* what) {
    version(what);
    const uint64_t dim = u64(what);
    std::vector<T> v = vec<T>(what);
    if (v.size() != dim) fail(std::string("data and dimension must match in size (") + what + ")");
    return v;
  }
  template <typename T>
  std::vector<T> array2(uint64_t* rows, uint64_t* cols, const char* what) {
    version(what);
    *rows = u64(what);
    *cols = u64(what);
    std::vector<T> v = vec<T>(what);
    if (*cols != 0 && *rows > UINT64_MAX / *cols) fail(std::string("array shape overflows (") + what + ")");
    if (v.size() != *rows * *cols) fail(std::string("data and dimension must match in size (") + what + ")");
    return v;
  }
  [[noreturn]] static void fail(const std::string& m) { throw Error(Err::InvalidArgument, "bincode: " + m); }

 private:
  void version(const char* what) {
    const uint8_t v = u8(what);
    if (v != 1) fail(std::string("unknown array version: ") + std::to_string(v) + " (" + what + ")");
  }
  const uint8_t* p_;
  uint64_t n_, pos_ = 0;
};

// Encoded sizes, so that a caller can size its buffer without the data
// being gathered.
inline uint64_t array1_bytes(uint64_t len, uint64_t elem) { return 1 + 8 + 8 + len * elem; }
inline uint64_t array2_bytes(uint64_t rows, uint64_t cols, uint64_t elem) { return 1 + 16 + 8 + rows * cols * elem; }
inline uint64_t manifold_bytes(uint64_t n, uint64_t nnz) {
  return 4 + 8 + 8 + (8 + 8 * (n + 1)) + (8 + 4 * nnz) + (8 + 4 * nnz) + 2 * array1_bytes(n, 4) + 8 + 4 + 4;
}
inline uint64_t optimizer_bytes(uint64_t n, uint64_t nnz, uint64_t edges, uint64_t dim) {
  return manifold_bytes(n, nnz) + 2 * array1_bytes(edges, 4) + 4 * array1_bytes(edges, 8) + array2_bytes(n, dim, 4) +
         8 + 8 + 4 + 4 + 8 + 4;
}
inline uint64_t config_bytes(const umapgpu_config& c) {
  return 8 + 4 + 4 + (1 + (c.has_a ? 4 : 0)) + (1 + (c.has_b ? 4 : 0)) + 8 + 4 + 4 +
         (1 + (c.has_disconnection_distance ? 4 : 0)) + 1 + (1 + (c.has_n_epochs ? 8 : 0)) + 4 + 8 + 4;
}
inline uint64_t fitted_bytes(uint64_t n, uint64_t nnz, uint64_t dim, const umapgpu_config& c) {
  return array2_bytes(n, dim, 4) + manifold_bytes(n, nnz) + config_bytes(c);
}

inline void write_manifold(Writer& w, const Manifold& m) {
  // graph: CsMatI<f32, u32, usize>
  w.u32(0);  // CompressedStorage::CSR
  w.u64(m.nrows);
  w.u64(m.ncols);
  w.vec(m.indptr);
  w.vec(m.indices);
  w.vec(m.data);
  w.array1(m.sigmas);
  w.array1(m.rhos);
  w.u64(m.n_vertices);
  w.f32(m.a);
  w.f32(m.b);
}

inline Manifold read_manifold(Reader& r) {
  Manifold m;
  const uint32_t storage = r.u32("graph.storage");
  if (storage > 1) Reader::fail("unknown variant index " + std::to_string(storage) + " of CompressedStorage");
  if (storage != 0) Reader::fail("the manifold graph must be a CSR matrix");
  m.nrows = r.u64("graph.nrows");
  m.ncols = r.u64("graph.ncols");
  m.indptr = r.vec<uint64_t>("graph.indptr");
  m.indices = r.vec<uint32_t>("graph.indices");
  m.data = r.vec<float>("graph.data");
  // the structure checks of a sprs CSR matrix
  if (m.indptr.size() != m.nrows + 1) Reader::fail("graph: indptr length must be nrows + 1");
  if (m.indices.size() != m.data.size()) Reader::fail("graph: indices and data lengths differ");
  if (m.indptr[0] != 0 || m.indptr[m.nrows] != m.indices.size()) Reader::fail("graph: indptr must span the data");
  for (uint64_t i = 0; i < m.nrows; ++i) {
    if (m.indptr[i + 1] < m.indptr[i]) Reader::fail("graph: indptr must be non-decreasing");
    for (uint64_t e = m.indptr[i]; e < m.indptr[i + 1]; ++e) {
      if (m.indices[e] >= m.ncols) Reader::fail("graph: column index out of range");
      if (e > m.indptr[i] && m.indices[e] <= m.indices[e - 1]) Reader::fail("graph: rows must be sorted");
    }
  }
  m.sigmas = r.array1<float>("sigmas");
  m.rhos = r.array1<float>("rhos");
  m.n_vertices = r.u64("n_vertices");
  m.a = r.f32("a");
  m.b = r.f32("b");
  if (m.nrows != m.n_vertices || m.ncols != m.n_vertices || m.sigmas.size() != m.n_vertices ||
      m.rhos.size() != m.n_vertices)
    Reader::fail("the graph, sigmas and rhos must all have n_vertices rows");
  return m;
}

inline void write_optimizer(Writer& w, const Optimizer& o) {
  write_manifold(w, o.manifold);
  w.array1(o.head);
  w.array1(o.tail);
  w.array1(o.epochs_per_sample);
  w.array2(o.embedding, o.rows, o.cols);
  w.array1(o.epoch_of_next_sample);
  w.array1(o.epoch_of_next_negative_sample);
  w.array1(o.epochs_per_negative_sample);
  w.u64(o.current_epoch);
  w.u64(o.total_epochs);
  w.f32(o.gamma);
  w.f32(o.initial_alpha);
  w.f64(o.negative_sample_rate);
  w.u32(o.metric_type);
}

inline Optimizer read_optimizer(Reader& r) {
  Optimizer o;
  o.manifold = read_manifold(r);
  o.head = r.array1<uint32_t>("head");
  o.tail = r.array1<uint32_t>("tail");
  o.epochs_per_sample = r.array1<double>("epochs_per_sample");
  o.embedding = r.array2<float>(&o.rows, &o.cols, "embedding");
  o.epoch_of_next_sample = r.array1<double>("epoch_of_next_sample");
  o.epoch_of_next_negative_sample = r.array1<double>("epoch_of_next_negative_sample");
  o.epochs_per_negative_sample = r.array1<double>("epochs_per_negative_sample");
  o.current_epoch = r.u64("current_epoch");
  o.total_epochs = r.u64("total_epochs");
  o.gamma = r.f32("gamma");
  o.initial_alpha = r.f32("initial_alpha");
  o.negative_sample_rate = r.f64("negative_sample_rate");
  o.metric_type = r.u32("metric_type");
  if (o.metric_type > 1) Reader::fail("unknown variant index " + std::to_string(o.metric_type) + " of MetricType");
  const uint64_t e = o.head.size();
  if (o.tail.size() != e || o.epochs_per_sample.size() != e || o.epoch_of_next_sample.size() != e ||
      o.epoch_of_next_negative_sample.size() != e || o.epochs_per_negative_sample.size() != e)
    Reader::fail("the per-edge arrays must have one length");
  if (o.rows != o.manifold.n_vertices) Reader::fail("the embedding must have n_vertices rows");
  return o;
}

// UmapConfig (config.rs:206-224) with its three parameter structs.
inline void write_config(Writer& w, const umapgpu_config& c) {
  w.u64(c.n_components);
  w.f32(c.min_dist);  // ManifoldParams (config.rs:9-41)
  w.f32(c.spread);
  w.opt_f32(c.has_a != 0, c.a);
  w.opt_f32(c.has_b != 0, c.b);
  w.u64(c.n_neighbors);  // GraphParams (:59-115)
  w.f32(c.local_connectivity);
  w.f32(c.set_op_mix_ratio);
  w.opt_f32(c.has_disconnection_distance != 0, c.disconnection_distance);
  w.boolean(c.symmetrize != 0);
  w.opt_u64(c.has_n_epochs != 0, c.n_epochs);  // OptimizationParams (:134-169)
  w.f32(c.learning_rate);
  w.u64(c.negative_sample_rate);
  w.f32(c.repulsion_strength);
}

inline umapgpu_config read_config(Reader& r) {
  umapgpu_config c{};
  c.n_components = r.u64("config.n_components");
  c.min_dist = r.f32("config.manifold.min_dist");
  c.spread = r.f32("config.manifold.spread");
  c.has_a = r.option_tag("config.manifold.a");
  c.a = c.has_a ? r.f32("config.manifold.a") : 0.0f;
  c.has_b = r.option_tag("config.manifold.b");
  c.b = c.has_b ? r.f32("config.manifold.b") : 0.0f;
  c.n_neighbors = r.u64("config.graph.n_neighbors");
  c.local_connectivity = r.f32("config.graph.local_connectivity");
  c.set_op_mix_ratio = r.f32("config.graph.set_op_mix_ratio");
  c.has_disconnection_distance = r.option_tag("config.graph.disconnection_distance");
  c.disconnection_distance = c.has_disconnection_distance ? r.f32("config.graph.disconnection_distance") : 0.0f;
  c.symmetrize = r.boolean("config.graph.symmetrize");
  c.has_n_epochs = r.option_tag("config.optimization.n_epochs");
  c.n_epochs = c.has_n_epochs ? r.u64("config.optimization.n_epochs") : 0;
  c.learning_rate = r.f32("config.optimization.learning_rate");
  c.negative_sample_rate = r.u64("config.optimization.negative_sample_rate");
  c.repulsion_strength = r.f32("config.optimization.repulsion_strength");
  return c;
}

inline void write_fitted(Writer& w, const Fitted& f) {
  w.array2(f.embedding, f.rows, f.cols);
  write_manifold(w, f.manifold);
  write_config(w, f.config);
}

inline Fitted read_fitted(Reader& r) {
  Fitted f;
  f.embedding = r.array2<float>(&f.rows, &f.cols, "embedding");
  f.manifold = read_manifold(r);
  f.config = read_config(r);
  if (f.rows != f.manifold.n_vertices) Reader::fail("the embedding must have n_vertices rows");
  return f;
}

}  // namespace bincode
}  // namespace umapgpu
