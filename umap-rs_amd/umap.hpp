// umap.hpp -- C++17 facade over libumapgpu's C ABI (include/umapgpu.h) that
// mirrors the public API of the umap-rs crate (0.4.5): the same type and
// method names, argument meanings, defaults and panic texts.  Reference
// items (paths relative to its tree):
//
//   UmapConfig, ManifoldParams, GraphParams, OptimizationParams  config.rs
//   Metric, MetricType, EuclideanMetric                          metric.rs, distances.rs
//   Umap::{new, with_metrics, learn_manifold, fit}               embedding.rs:73-337
//   LearnedManifold::{n_vertices, curve_params, graph}           manifold.rs:41-54
//   Optimizer::{new, step_epochs, current_epoch, total_epochs,
//               remaining_epochs, embedding, manifold, into_fitted}  optimizer.rs:74-403
//   FittedUmap::{embedding, into_embedding, manifold, config}    embedding.rs:462-493
//
// Where the reference panics, this facade throws umap::Panic carrying the
// same message.  Ownership follows the reference: Optimizer's constructor
// takes the manifold by value (moved), into_fitted() consumes the optimizer.
// Header-only; link with -lumapgpu.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <optional>
#include <typeinfo>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "umapgpu.h"

namespace umap {

// A reference panic (validation, overstep, shape mismatch).
struct Panic : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// A device failure (HIP error, out of memory, no device).
struct GpuError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(umapgpu_status s) {
  if (s == UMAPGPU_OK) return;
  const std::string msg = umapgpu_last_error();
  if (s == UMAPGPU_INVALID_ARGUMENT || s == UMAPGPU_OVERSTEP) throw Panic(msg);
  throw GpuError("[status " + std::to_string((int)s) + "] " + msg);
}

// ---------------------------------------------------------------- config.rs
struct ManifoldParams {  // config.rs:9-52
  float min_dist = 0.1f;
  float spread = 1.0f;
  std::optional<float> a;
  std::optional<float> b;
};
struct GraphParams {  // config.rs:59-127
  size_t n_neighbors = 15;
  float local_connectivity = 1.0f;
  float set_op_mix_ratio = 1.0f;
  std::optional<float> disconnection_distance;
  bool symmetrize = true;
};
struct OptimizationParams {  // config.rs:134-180
  std::optional<size_t> n_epochs;
  float learning_rate = 1.0f;
  size_t negative_sample_rate = 5;
  float repulsion_strength = 1.0f;
};
struct UmapConfig {  // config.rs:206-235
  size_t n_components = 2;
  ManifoldParams manifold;
  GraphParams graph;
  OptimizationParams optimization;

  umapgpu_config to_c() const {
    umapgpu_config c;
    umapgpu_config_default(&c);
    c.n_components = n_components;
    c.min_dist = manifold.min_dist;
    c.spread = manifold.spread;
    c.has_a = manifold.a.has_value();
    c.a = manifold.a.value_or(0.0f);
    c.has_b = manifold.b.has_value();
    c.b = manifold.b.value_or(0.0f);
    c.n_neighbors = graph.n_neighbors;
    c.local_connectivity = graph.local_connectivity;
    c.set_op_mix_ratio = graph.set_op_mix_ratio;
    c.has_disconnection_distance = graph.disconnection_distance.has_value();
    c.disconnection_distance = graph.disconnection_distance.value_or(0.0f);
    c.symmetrize = graph.symmetrize;
    c.has_n_epochs = optimization.n_epochs.has_value();
    c.n_epochs = optimization.n_epochs.value_or(0);
    c.learning_rate = optimization.learning_rate;
    c.negative_sample_rate = optimization.negative_sample_rate;
    c.repulsion_strength = optimization.repulsion_strength;
    return c;
  }
  static UmapConfig from_c(const umapgpu_config& c) {
    UmapConfig u;
    u.n_components = c.n_components;
    u.manifold.min_dist = c.min_dist;
    u.manifold.spread = c.spread;
    if (c.has_a) u.manifold.a = c.a;
    if (c.has_b) u.manifold.b = c.b;
    u.graph.n_neighbors = c.n_neighbors;
    u.graph.local_connectivity = c.local_connectivity;
    u.graph.set_op_mix_ratio = c.set_op_mix_ratio;
    if (c.has_disconnection_distance) u.graph.disconnection_distance = c.disconnection_distance;
    u.graph.symmetrize = c.symmetrize != 0;
    if (c.has_n_epochs) u.optimization.n_epochs = c.n_epochs;
    u.optimization.learning_rate = c.learning_rate;
    u.optimization.negative_sample_rate = c.negative_sample_rate;
    u.optimization.repulsion_strength = c.repulsion_strength;
    return u;
  }
};

// ------------------------------------------------------- arrays (ndarray)
// Borrowed row-major 2-D view (ArrayView2): any row stride >= cols.
template <typename T>
struct ArrayView2 {
  const T* data = nullptr;
  size_t rows = 0, cols = 0, row_stride = 0;
  ArrayView2() = default;
  ArrayView2(const T* d, size_t r, size_t c) : data(d), rows(r), cols(c), row_stride(c) {}
  ArrayView2(const T* d, size_t r, size_t c, size_t stride) : data(d), rows(r), cols(c), row_stride(stride) {}
  size_t shape(int axis) const { return axis == 0 ? rows : cols; }
};

// Owned row-major 2-D array (Array2<f32>).
struct Array2f {
  std::vector<float> data;
  size_t rows = 0, cols = 0;
  Array2f() = default;
  Array2f(size_t r, size_t c, float v = 0.0f) : data(r * c, v), rows(r), cols(c) {}
  float& operator()(size_t i, size_t j) { return data[i * cols + j]; }
  float operator()(size_t i, size_t j) const { return data[i * cols + j]; }
  ArrayView2<float> view() const { return ArrayView2<float>(data.data(), rows, cols); }
  std::vector<size_t> shape() const { return {rows, cols}; }
};

// --------------------------------------------------------------- metric.rs
enum class MetricType { Euclidean = UMAPGPU_METRIC_EUCLIDEAN, Generic = UMAPGPU_METRIC_GENERIC };  // :7-12

class Metric {  // trait Metric (metric.rs:19-62)
 public:
  virtual ~Metric() = default;
  virtual MetricType metric_type() const { return MetricType::Generic; }
  // distance between a and b; writes d distance / d a into grad[0..dim)
  virtual float distance(const float* a, const float* b, size_t dim, float* grad) const = 0;
  virtual float disconnection_threshold() const { return std::numeric_limits<float>::infinity(); }
  virtual std::optional<float> squared_distance(const float*, const float*, size_t) const { return std::nullopt; }
};

class EuclideanMetric : public Metric {  // distances.rs:10-50
 public:
  MetricType metric_type() const override { return MetricType::Euclidean; }
  float distance(const float* x, const float* y, size_t dim, float* grad) const override {
    float sum_sq = 0.0f;
    for (size_t i = 0; i < dim; ++i) {
      const float d = x[i] - y[i];
      sum_sq = sum_sq + d * d;
    }
    const float dist = std::sqrt(sum_sq);
    const float denom = dist + 1e-6f;
    for (size_t i = 0; i < dim; ++i) grad[i] = (x[i] - y[i]) / denom;
    return dist;
  }
  std::optional<float> squared_distance(const float* x, const float* y, size_t dim) const override {
    float s = 0.0f;
    for (size_t i = 0; i < dim; ++i) {
      const float d = x[i] - y[i];
      s = s + d * d;
    }
    return s;
  }
};

namespace detail {
inline float metric_trampoline(const float* a, const float* b, uint64_t dim, float* grad, void* user) {
  return static_cast<const Metric*>(user)->distance(a, b, (size_t)dim, grad);
}
}  // namespace detail

// -------------------------------------------------------------- manifold.rs
// CsMatI<f32, u32, usize> on the host (umap/fuzzy_simplicial_set.rs:18).
struct SparseMat {
  size_t n = 0;
  std::vector<uint64_t> indptr;
  std::vector<uint32_t> indices;
  std::vector<float> data;
  size_t nnz() const { return indptr.empty() ? 0 : (size_t)indptr.back(); }
  std::pair<size_t, size_t> shape() const { return {n, n}; }
};

namespace detail {
// bincode::serialize through a *_to_bincode entry point: the size, then the bytes.
template <typename H, typename F>
std::vector<uint8_t> to_bincode(F fn, H h) {
  uint64_t n = 0;
  check(fn(h, nullptr, 0, &n));
  std::vector<uint8_t> out(n);
  check(fn(h, out.data(), n, &n));
  return out;
}
}  // namespace detail

class LearnedManifold {  // manifold.rs:15-54, device resident
 public:
  LearnedManifold() = default;
  explicit LearnedManifold(umapgpu_manifold* h) : h_(h) {}
  LearnedManifold(LearnedManifold&& o) noexcept : h_(o.h_), borrowed_(o.borrowed_) { o.h_ = nullptr; }
  LearnedManifold& operator=(LearnedManifold&& o) noexcept {
    if (this != &o) {
      reset();
      h_ = o.h_;
      borrowed_ = o.borrowed_;
      o.h_ = nullptr;
    }
    return *this;
  }
  LearnedManifold(const LearnedManifold& o) : h_(nullptr) {  // #[derive(Clone)]
    umapgpu_manifold* c = nullptr;
    check(umapgpu_manifold_clone(o.handle(), &c));
    h_ = c;
  }
  ~LearnedManifold() { reset(); }

  static LearnedManifold borrowed(const umapgpu_manifold* h) {
    LearnedManifold m(const_cast<umapgpu_manifold*>(h));
    m.borrowed_ = true;
    return m;
  }

  size_t n_vertices() const {
    uint64_t nv = 0;
    check(umapgpu_manifold_info(handle(), &nv, nullptr, nullptr, nullptr));
    return (size_t)nv;
  }
  std::pair<float, float> curve_params() const {
    float a = 0, b = 0;
    check(umapgpu_manifold_info(handle(), nullptr, nullptr, &a, &b));
    return {a, b};
  }
  SparseMat graph() const {
    uint64_t nv = 0, nnz = 0;
    check(umapgpu_manifold_info(handle(), &nv, &nnz, nullptr, nullptr));
    SparseMat g;
    g.n = (size_t)nv;
    g.indptr.resize(nv + 1);
    g.indices.resize(nnz);
    g.data.resize(nnz);
    check(umapgpu_manifold_copy_graph(handle(), g.indptr.data(), nnz ? g.indices.data() : nullptr,
                                      nnz ? g.data.data() : nullptr));
    return g;
  }
  std::pair<std::vector<float>, std::vector<float>> sigmas_rhos() const {
    const size_t n = n_vertices();
    std::vector<float> s(n), r(n);
    check(umapgpu_manifold_copy_sigmas_rhos(handle(), s.data(), r.data()));
    return {s, r};
  }
  // bincode::serialize(&manifold) / bincode::deserialize (manifold.rs:14)
  std::vector<uint8_t> to_bincode() const { return detail::to_bincode(umapgpu_manifold_to_bincode, handle()); }
  static LearnedManifold from_bincode(const std::vector<uint8_t>& bytes) {
    umapgpu_manifold* m = nullptr;
    check(umapgpu_manifold_from_bincode(bytes.data(), bytes.size(), &m));
    return LearnedManifold(m);
  }
  umapgpu_manifold* handle() const {
    if (!h_) throw Panic("LearnedManifold was moved and can no longer be used");
    return h_;
  }
  umapgpu_manifold* release() {
    umapgpu_manifold* h = handle();
    h_ = nullptr;
    return h;
  }

 private:
  void reset() {
    if (h_ && !borrowed_) umapgpu_manifold_destroy(h_);
    h_ = nullptr;
  }
  umapgpu_manifold* h_ = nullptr;
  bool borrowed_ = false;
};

// ------------------------------------------------------------- FittedUmap
class FittedUmap {  // embedding.rs:439-523
 public:
  explicit FittedUmap(umapgpu_fitted* h) : h_(h) {}
  FittedUmap(FittedUmap&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  FittedUmap(const FittedUmap&) = delete;
  ~FittedUmap() {
    if (h_) umapgpu_fitted_destroy(h_);
  }
  Array2f embedding() const {
    uint64_t n = 0, d = 0;
    check(umapgpu_fitted_info(h(), &n, &d));
    Array2f e((size_t)n, (size_t)d);
    check(umapgpu_fitted_copy_embedding(h(), e.data.data()));
    return e;
  }
  Array2f into_embedding() && {
    Array2f e = embedding();
    umapgpu_fitted_destroy(h_);
    h_ = nullptr;
    return e;
  }
  LearnedManifold manifold() const { return LearnedManifold::borrowed(umapgpu_fitted_manifold(h())); }
  UmapConfig config() const {
    umapgpu_config c;
    check(umapgpu_fitted_config(h(), &c));
    return UmapConfig::from_c(c);
  }
  // bincode::serialize(&fitted) / bincode::deserialize (embedding.rs:438)
  std::vector<uint8_t> to_bincode() const { return detail::to_bincode(umapgpu_fitted_to_bincode, h()); }
  static FittedUmap from_bincode(const std::vector<uint8_t>& bytes) {
    umapgpu_fitted* f = nullptr;
    check(umapgpu_fitted_from_bincode(bytes.data(), bytes.size(), &f));
    return FittedUmap(f);
  }
  // embedding.rs:516-523: not implemented in the reference either
  Array2f transform(ArrayView2<float>, ArrayView2<uint32_t>, ArrayView2<float>) const {
    throw Panic("not yet implemented: Transform not yet implemented - contributions welcome!");
  }

 private:
  umapgpu_fitted* h() const {
    if (!h_) throw Panic("FittedUmap was moved and can no longer be used");
    return h_;
  }
  umapgpu_fitted* h_ = nullptr;
};

// -------------------------------------------------------------- Optimizer
class Optimizer {  // optimizer.rs:28-403
 public:
  // Optimizer::new (optimizer.rs:74-290); the manifold is moved in.
  Optimizer(LearnedManifold manifold, const Array2f& init, size_t total_epochs, const UmapConfig& config,
            MetricType metric_type, uint64_t seed = 0) {
    const umapgpu_config c = config.to_c();
    umapgpu_optimizer* o = nullptr;
    check(umapgpu_optimizer_new(manifold.handle(), init.data.data(), init.rows, init.cols, total_epochs, &c,
                                (int32_t)metric_type, seed, &o));
    manifold.release();  // consumed by the library
    h_ = o;
  }
  Optimizer(Optimizer&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  Optimizer(const Optimizer&) = delete;
  // bincode::serialize(&opt) / bincode::deserialize (optimizer.rs:27); the
  // checkpoint has no RNG state, so `seed` keys the restored negative sampling.
  std::vector<uint8_t> to_bincode() const { return detail::to_bincode(umapgpu_optimizer_to_bincode, h()); }
  static Optimizer from_bincode(const std::vector<uint8_t>& bytes, uint64_t seed = 0) {
    umapgpu_optimizer* o = nullptr;
    check(umapgpu_optimizer_from_bincode(bytes.data(), bytes.size(), seed, &o));
    return Optimizer(o);
  }
  ~Optimizer() {
    if (h_) umapgpu_optimizer_destroy(h_);
  }

  // Optimizer::step_epochs (optimizer.rs:297-366).  MetricType::Generic with
  // EuclideanMetric itself (not a subclass) runs the generic loop on the
  // device with the metric built in (n_components <= 8); any other metric
  // runs the reference's serial generic loop on the host through the callback.
  void step_epochs(size_t n, const Metric& output_metric) {
    int32_t mt = 0;
    check(umapgpu_optimizer_export_params(h(), nullptr, nullptr, nullptr, nullptr, nullptr, &mt, nullptr));
    uint64_t dim = 0;
    check(umapgpu_optimizer_info(h(), nullptr, nullptr, &dim));
    if (mt == UMAPGPU_METRIC_GENERIC && typeid(output_metric) == typeid(EuclideanMetric) && dim <= 8)
      check(umapgpu_optimizer_step_epochs_builtin(h(), n, UMAPGPU_BUILTIN_METRIC_EUCLIDEAN));
    else if (mt == UMAPGPU_METRIC_GENERIC)
      check(umapgpu_optimizer_step_epochs_generic(h(), n, &detail::metric_trampoline,
                                                  const_cast<Metric*>(&output_metric)));
    else
      check(umapgpu_optimizer_step_epochs(h(), n));
  }
  void set_mode(umapgpu_sgd_mode mode) { check(umapgpu_optimizer_set_mode(h(), (int32_t)mode)); }
  size_t current_epoch() const { return (size_t)umapgpu_optimizer_current_epoch(h()); }
  size_t total_epochs() const { return (size_t)umapgpu_optimizer_total_epochs(h()); }
  size_t remaining_epochs() const { return (size_t)umapgpu_optimizer_remaining_epochs(h()); }
  Array2f embedding() const {
    uint64_t ne = 0, nv = 0, d = 0;
    check(umapgpu_optimizer_info(h(), &ne, &nv, &d));
    Array2f e((size_t)nv, (size_t)d);
    check(umapgpu_optimizer_copy_embedding(h(), e.data.data()));
    return e;
  }
  LearnedManifold manifold() const { return LearnedManifold::borrowed(umapgpu_optimizer_manifold(h())); }
  // Optimizer::into_fitted (optimizer.rs:397-403); consumes the optimizer.
  FittedUmap into_fitted(const UmapConfig& config) && {
    const umapgpu_config c = config.to_c();
    umapgpu_fitted* f = nullptr;
    check(umapgpu_optimizer_into_fitted(h(), &c, &f));
    h_ = nullptr;
    return FittedUmap(f);
  }
  umapgpu_optimizer* handle() const { return h(); }

 private:
  explicit Optimizer(umapgpu_optimizer* h) : h_(h) {}
  umapgpu_optimizer* h() const {
    if (!h_) throw Panic("Optimizer was moved and can no longer be used");
    return h_;
  }
  umapgpu_optimizer* h_ = nullptr;
};

// ------------------------------------------------------------------- Umap
class Umap {  // embedding.rs:58-430
 public:
  explicit Umap(UmapConfig config)  // Umap::new (:73-79)
      : config_(std::move(config)), metric_(std::make_shared<EuclideanMetric>()),
        output_metric_(std::make_shared<EuclideanMetric>()) {}
  static Umap with_metrics(UmapConfig config, std::shared_ptr<Metric> metric,
                           std::shared_ptr<Metric> output_metric) {  // :98-108
    Umap u(std::move(config));
    u.metric_ = std::move(metric);
    u.output_metric_ = std::move(output_metric);
    return u;
  }

  // Umap::learn_manifold (:149-240); data values are never read (:155).
  LearnedManifold learn_manifold(ArrayView2<float> data, ArrayView2<uint32_t> knn_indices,
                                 ArrayView2<float> knn_dists) const {
    const umapgpu_config c = config_.to_c();
    umapgpu_manifold* m = nullptr;
    check(umapgpu_learn_manifold(&c, data.rows, knn_indices.data, knn_indices.rows, knn_indices.cols,
                                 knn_indices.row_stride, knn_dists.data, knn_dists.rows, knn_dists.cols,
                                 knn_dists.row_stride, metric_->disconnection_threshold(), &m));
    return LearnedManifold(m);
  }

  // Umap::fit (:274-337)
  FittedUmap fit(ArrayView2<float> data, ArrayView2<uint32_t> knn_indices, ArrayView2<float> knn_dists,
                 ArrayView2<float> init, uint64_t seed = 0) const {
    const umapgpu_config c = config_.to_c();
    const MetricType mt = output_metric_->metric_type();
    std::vector<float> packed;
    const float* ip = init.data;
    if (init.row_stride != init.cols) {
      packed.resize(init.rows * init.cols);
      for (size_t i = 0; i < init.rows; ++i)
        for (size_t j = 0; j < init.cols; ++j) packed[i * init.cols + j] = init.data[i * init.row_stride + j];
      ip = packed.data();
    }
    umapgpu_fitted* f = nullptr;
    check(umapgpu_fit(&c, data.rows, knn_indices.data, knn_indices.rows, knn_indices.cols, knn_indices.row_stride,
                      knn_dists.data, knn_dists.rows, knn_dists.cols, knn_dists.row_stride, ip, init.rows, init.cols,
                      metric_->disconnection_threshold(), (int32_t)mt,
                      mt == MetricType::Generic ? &detail::metric_trampoline : nullptr,
                      const_cast<Metric*>(output_metric_.get()), seed, &f));
    return FittedUmap(f);
  }

  const UmapConfig& config() const { return config_; }

 private:
  UmapConfig config_;
  std::shared_ptr<Metric> metric_, output_metric_;
};

}  // namespace umap
