// Learned initialisation network: a D -> 256 -> 256 -> O ReLU MLP trained on
// the embedding a plain fit produces for a subsample, then used to predict the
// initial layout of every point.  Forward and backward run on bf16 MFMA with
// f32 accumulation; master weights and Adam moments are f32.  The reference
// has no counterpart (its fits start from the caller's init).
#pragma once
#include <vector>

#include "common.hpp"

namespace umapgpu {

constexpr int kMlpHidden = 256;
constexpr int kMlpMaxOut = 8;

struct MlpConfig {
  uint64_t in_dim = 0, out_dim = 2, batch = 4096;
  float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  uint64_t seed = 0;
};

// Flat f32 parameter vector: W1 [256 x D], b1 [256], W2 [256 x 256], b2 [256],
// W3 [O x 256], b3 [O] (rows are output units).
struct MlpLayout {
  uint64_t D = 0, O = 0, Dp = 0;  // Dp: D rounded up to 64 (bf16 copy row length)
  uint64_t w1 = 0, b1 = 0, w2 = 0, b2 = 0, w3 = 0, b3 = 0, total = 0;
  static MlpLayout make(uint64_t D, uint64_t O);
};

struct MlpState {
  MlpConfig cfg;
  MlpLayout lay;
  uint64_t bcap = 0;   // batch capacity (multiple of 64)
  uint64_t step = 0;   // Adam steps taken
  double train_ms = 0.0;
  float yscale = 1.0f;
  DevBuf w, m, v;                 // f32[P]
  DevBuf w1b, w2b, w2tb;          // bf16 copies: [256 x Dp], [256 x 256], transposed [256 x 256]
  DevBuf xmean, xinv, ymean;      // f32[D], f32[D], f32[O]
  DevBuf xt, h1t, h2t, dh1t, dh2t;  // bf16 feature-major batch buffers [rows x bcap]
  DevBuf dyt;                     // f32 [O x bcap]
  DevBuf part;                    // f32 [splits x P] weight-gradient partials
  DevBuf grad;                    // f32 [P]
  DevBuf loss_part;               // f32 per forward tile
  DevBuf rows;                    // u32 batch row ids
};

void mlp_create(const MlpConfig& c, MlpState* st, hipStream_t s);
void mlp_set_params(MlpState* st, const float* host, hipStream_t s);  // resets Adam
void mlp_get_params(const MlpState& st, float* host, hipStream_t s);
void mlp_fit_normalization(MlpState* st, const float* X, const float* Y, uint64_t n, hipStream_t s);
void mlp_set_normalization(MlpState* st, const float* xmean, const float* xinv, const float* ymean, float yscale,
                           hipStream_t s);
void mlp_get_normalization(const MlpState& st, float* xmean, float* xinv, float* ymean, float* yscale,
                           hipStream_t s);
// Gradient of the batch made of rows 0..n-1 (n <= batch), no update.
double mlp_gradients(MlpState* st, const float* X, const float* Y, uint64_t n, float* grad_host, hipStream_t s);
// `epochs` passes over rows 0..n-1 in shuffled batches with Adam; returns the
// mean loss of the last epoch.
double mlp_train(MlpState* st, const float* X, const float* Y, uint64_t n, uint64_t epochs, hipStream_t s);
void mlp_predict(const MlpState& st, const float* X, uint64_t n, float* Y, hipStream_t s);

}  // namespace umapgpu
