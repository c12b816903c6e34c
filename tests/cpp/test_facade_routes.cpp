// The C++ facade's metric routing (umap-rs_amd/umap.hpp, Optimizer::step_epochs):
// MetricType::Generic with EuclideanMetric runs the generic loop on the
// device with the metric built in; any other metric (here a subclass with
// the same distance) runs the reference's host loop through the callback.
// In SERIAL mode both are the reference's sequential order with the same
// arithmetic, so they must agree bit for bit; the default parallel epoch is
// a different order, so it must not.  Exit status 0 when every check passes.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "umap.hpp"

using namespace umap;

struct UserEuclidean : EuclideanMetric {};  // a user's metric: not the library's own type

static LearnedManifold ring_manifold(size_t n, size_t k, const UmapConfig& cfg, std::mt19937& rng) {
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  std::vector<float> data(n * 4);
  for (auto& v : data) v = u(rng);
  std::vector<uint32_t> idx(n * k);
  std::vector<float> dist(n * k);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < k; ++j) {
      idx[i * k + j] = (uint32_t)((i + j) % n);  // self first, then the ring
      dist[i * k + j] = j == 0 ? 0.0f : 0.1f * (float)j + 0.05f * u(rng);
    }
  return Umap(cfg).learn_manifold({data.data(), n, 4}, {idx.data(), n, k}, {dist.data(), n, k});
}

static bool same_bits(const Array2f& a, const Array2f& b) {
  return a.data.size() == b.data.size() && std::memcmp(a.data.data(), b.data.data(), a.data.size() * 4) == 0;
}

int main() {
  const size_t n = 300, k = 10;
  std::mt19937 rng(5);
  UmapConfig cfg;
  cfg.graph.n_neighbors = k;
  LearnedManifold m = ring_manifold(n, k, cfg, rng);
  Array2f init(n, 2);
  std::uniform_real_distribution<float> u(0.0f, 10.0f);
  for (auto& v : init.data) v = u(rng);
  EuclideanMetric builtin;
  UserEuclidean user;
  int failed = 0;
  {
    Optimizer dev(LearnedManifold(m), init, 8, cfg, MetricType::Generic, 7);
    Optimizer host(LearnedManifold(m), init, 8, cfg, MetricType::Generic, 7);
    dev.set_mode(UMAPGPU_SGD_SERIAL);
    dev.step_epochs(4, builtin);  // device, serial generic loop
    host.step_epochs(4, user);    // host loop, metric called back
    const bool ok = same_bits(dev.embedding(), host.embedding());
    std::printf("%s serial generic: device loop == host callback loop\n", ok ? "PASS" : "FAIL");
    failed += !ok;
  }
  {
    Optimizer par(LearnedManifold(m), init, 8, cfg, MetricType::Generic, 7);
    Optimizer host(LearnedManifold(m), init, 8, cfg, MetricType::Generic, 7);
    par.step_epochs(4, builtin);  // device, default parallel epoch
    host.step_epochs(4, user);
    const bool ok = !same_bits(par.embedding(), host.embedding());
    std::printf("%s default generic epoch runs on the device (differs from the sequential loop)\n",
                ok ? "PASS" : "FAIL");
    failed += !ok;
  }
  std::fflush(stdout);
  return failed ? 1 : 0;
}
