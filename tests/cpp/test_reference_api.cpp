// The reference's unit tests (umap-rs/src/tests.rs:14-215) against the C++
// facade umap-rs_amd/umap.hpp.  The fixture is generate_test_data
// (tests.rs:14-38) with a fixed seed: ring neighbours (i+j+1) % n, unsorted
// U[0,2) distances, init U[0,10).  Exit status 0 when every test passes.
#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "umap.hpp"

using namespace umap;

struct TestData {
  std::vector<float> data;
  std::vector<uint32_t> knn_indices;
  std::vector<float> knn_dists;
  Array2f init;
  size_t n = 50, f = 20, k = 10;
  ArrayView2<float> data_view() const { return {data.data(), n, f}; }
  ArrayView2<uint32_t> idx_view() const { return {knn_indices.data(), n, k}; }
  ArrayView2<float> dist_view() const { return {knn_dists.data(), n, k}; }
};

static TestData generate_test_data(unsigned seed) {  // tests.rs:14-38
  TestData t;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u01(0.0f, 1.0f);
  t.data.resize(t.n * t.f);
  for (auto& v : t.data) v = u01(rng);
  t.knn_indices.resize(t.n * t.k);
  t.knn_dists.resize(t.n * t.k);
  for (size_t i = 0; i < t.n; ++i)
    for (size_t j = 0; j < t.k; ++j) {
      t.knn_indices[i * t.k + j] = (uint32_t)((i + j + 1) % t.n);
      t.knn_dists[i * t.k + j] = u01(rng) * 2.0f;
    }
  t.init = Array2f(t.n, 2);
  for (auto& v : t.init.data) v = u01(rng) * 10.0f;
  return t;
}

static UmapConfig config10() {
  UmapConfig c;
  c.graph.n_neighbors = 10;
  return c;
}

#define REQUIRE(cond)                                                   \
  do {                                                                  \
    if (!(cond)) throw std::runtime_error(std::string("check failed: ") + #cond); \
  } while (0)

static void test_high_level_fit() {  // tests.rs:40-63
  auto t = generate_test_data(1);
  UmapConfig config = config10();
  config.optimization.n_epochs = 10;
  Umap umap(config);
  FittedUmap fitted = umap.fit(t.data_view(), t.idx_view(), t.dist_view(), t.init.view());
  Array2f e = fitted.embedding();
  REQUIRE((e.shape() == std::vector<size_t>{50, 2}));
  for (float x : e.data) REQUIRE(std::isfinite(x) || std::isnan(x));
}

static void test_learn_manifold() {  // tests.rs:65-79
  auto t = generate_test_data(2);
  Umap umap(config10());
  LearnedManifold manifold = umap.learn_manifold(t.data_view(), t.idx_view(), t.dist_view());
  REQUIRE(manifold.n_vertices() == 50);
  REQUIRE(manifold.graph().nnz() > 0);
  auto ab = manifold.curve_params();
  REQUIRE(ab.first > 0.0f && ab.second > 0.0f);
}

static void test_optimizer_step_epochs() {  // tests.rs:81-110
  auto t = generate_test_data(3);
  UmapConfig config = config10();
  Umap umap(config);
  LearnedManifold manifold = umap.learn_manifold(t.data_view(), t.idx_view(), t.dist_view());
  const size_t total_epochs = 20;
  EuclideanMetric metric;
  Optimizer opt(std::move(manifold), t.init, total_epochs, config, metric.metric_type());
  REQUIRE(opt.current_epoch() == 0);
  REQUIRE(opt.remaining_epochs() == 20);
  opt.step_epochs(5, metric);
  REQUIRE(opt.current_epoch() == 5);
  REQUIRE(opt.remaining_epochs() == 15);
  opt.step_epochs(15, metric);
  REQUIRE(opt.current_epoch() == 20);
  REQUIRE(opt.remaining_epochs() == 0);
  FittedUmap fitted = std::move(opt).into_fitted(config);
  REQUIRE((fitted.embedding().shape() == std::vector<size_t>{50, 2}));
}

static void test_optimizer_overstep_panics() {  // tests.rs:112-130 (#[should_panic])
  auto t = generate_test_data(4);
  UmapConfig config = config10();
  Umap umap(config);
  LearnedManifold manifold = umap.learn_manifold(t.data_view(), t.idx_view(), t.dist_view());
  EuclideanMetric metric;
  Optimizer opt(std::move(manifold), t.init, 20, config, metric.metric_type());
  opt.step_epochs(18, metric);
  try {
    opt.step_epochs(5, metric);
  } catch (const Panic& p) {
    REQUIRE(std::string(p.what()) == "Cannot step 5 epochs: would exceed total_epochs 20 (current: 18)");
    return;
  }
  throw std::runtime_error("expected a panic");
}

static void test_checkpoint_serialization() {  // tests.rs:132-162
  auto t = generate_test_data(5);
  UmapConfig config = config10();
  Umap umap(config);
  LearnedManifold manifold = umap.learn_manifold(t.data_view(), t.idx_view(), t.dist_view());
  EuclideanMetric metric;
  Optimizer opt(std::move(manifold), t.init, 50, config, metric.metric_type());
  opt.step_epochs(20, metric);
  // Serialize (bincode::serialize) and deserialize (bincode::deserialize)
  std::vector<uint8_t> serialized = opt.to_bincode();
  Optimizer opt2 = Optimizer::from_bincode(serialized);
  REQUIRE(opt2.current_epoch() == 20);
  REQUIRE(opt2.remaining_epochs() == 30);
  // Continue training
  opt2.step_epochs(30, metric);
  FittedUmap fitted = std::move(opt2).into_fitted(config);
  REQUIRE((fitted.embedding().shape() == std::vector<size_t>{50, 2}));
}

static void test_manifold_serialization() {  // tests.rs:164-188
  auto t = generate_test_data(6);
  Umap umap(config10());
  LearnedManifold manifold = umap.learn_manifold(t.data_view(), t.idx_view(), t.dist_view());
  auto ab = manifold.curve_params();
  std::vector<uint8_t> serialized = manifold.to_bincode();
  LearnedManifold manifold2 = LearnedManifold::from_bincode(serialized);
  REQUIRE(manifold2.curve_params() == ab);
  REQUIRE(manifold2.n_vertices() == manifold.n_vertices());
  REQUIRE(manifold2.graph().nnz() == manifold.graph().nnz());
}

static void test_fitted_umap_serialization() {  // tests.rs:190-215
  auto t = generate_test_data(7);
  UmapConfig config = config10();
  config.optimization.n_epochs = 10;
  Umap umap(config);
  FittedUmap fitted = umap.fit(t.data_view(), t.idx_view(), t.dist_view(), t.init.view());
  std::vector<uint8_t> serialized = fitted.to_bincode();
  FittedUmap fitted2 = FittedUmap::from_bincode(serialized);
  REQUIRE(fitted2.embedding().shape() == fitted.embedding().shape());
  REQUIRE(fitted2.manifold().n_vertices() == fitted.manifold().n_vertices());
}

static void test_validation_panic_text() {  // embedding.rs:346-351
  auto t = generate_test_data(8);
  UmapConfig config = config10();
  config.graph.set_op_mix_ratio = 1.5f;
  try {
    Umap(config).learn_manifold(t.data_view(), t.idx_view(), t.dist_view());
  } catch (const Panic& p) {
    REQUIRE(std::string(p.what()) == "set_op_mix_ratio must be between 0.0 and 1.0, got 1.5");
    return;
  }
  throw std::runtime_error("expected a panic");
}

int main() {
  const std::vector<std::pair<const char*, std::function<void()>>> tests = {
      {"test_high_level_fit", test_high_level_fit},
      {"test_learn_manifold", test_learn_manifold},
      {"test_optimizer_step_epochs", test_optimizer_step_epochs},
      {"test_optimizer_overstep_panics", test_optimizer_overstep_panics},
      {"test_checkpoint_serialization", test_checkpoint_serialization},
      {"test_manifold_serialization", test_manifold_serialization},
      {"test_fitted_umap_serialization", test_fitted_umap_serialization},
      {"test_validation_panic_text", test_validation_panic_text},
  };
  int failed = 0;
  for (auto& [name, fn] : tests) {
    try {
      fn();
      std::printf("PASS %s\n", name);
    } catch (const std::exception& e) {
      std::printf("FAIL %s: %s\n", name, e.what());
      ++failed;
    }
    std::fflush(stdout);
  }
  return failed ? 1 : 0;
}
