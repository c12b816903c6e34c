/*
 * umap_oracle.c -- CPU restatement of the umap-rs fit path, used ONLY as test
 * infrastructure (the checker for the HIP implementation and the CPU baseline
 * leg of bench.py).  Nothing in the product library links or calls this file.
 *
 * Every function restates one stage of the reference crate (umap-rs 0.4.5),
 * following its operation order exactly: plain IEEE f32/f64, glibc libm
 * (expf/powf/logf/log2f, which is what Rust's f32::exp/powf/ln/log2 lower to on
 * Linux), and no FMA contraction (build with -ffp-contract=off; Rust never
 * contracts).  Citations are file:line inside the reference tree.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   find_ab_params      -- finite-difference gradient check of the GD step and
 *                          long-run convergence to scipy.optimize.curve_fit.
 *   smooth_knn_dist     -- closed forms (k=2 -> sigma=1, k=3 closed form,
 *                          constant rows hit the floor), the psum invariant,
 *                          an FP64 twin.
 *   membership / set ops -- dense brute force (numpy) for n <= 500.
 *   optimizer_new       -- closed forms of threshold, eps = max/w, [0,10] scaling.
 *   sgd epochs          -- hand-computed single-edge known answers, clip
 *                          saturation, serial vs Hogwild statistics.
 *   ndarray mean order  -- parity unpinned (ndarray 0.17 is not present; the
 *                          8-way unrolled fold is restated from its published
 *                          source and only affects the sigma floor).
 *   negative-sample RNG -- parity unpinned against the reference (rand 0.9
 *                          ThreadRng is OS-seeded); the build uses its own
 *                          counter RNG, restated identically here.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#include <stdio.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* constants: umap/constants.rs:4,7                                           */
static const float SMOOTH_K_TOLERANCE = 1e-5f;
static const float MIN_K_DIST_SCALE = 1e-3f;

ORC_API void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------- */
/* ndarray 0.17 `sum()` on a contiguous slice: eight-accumulator unrolled fold
 * (ndarray numeric_util::unrolled_fold), `mean()` = sum / len.
 * Used by smooth_knn_dist.rs:74 (whole array) and :177 (one row).           */
ORC_API float orc_unrolled_sum_f32(const float *xs, uint64_t len) {
  float acc = 0.0f;
  float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f, p3 = 0.0f, p4 = 0.0f, p5 = 0.0f, p6 = 0.0f, p7 = 0.0f;
  while (len >= 8) {
    p0 = p0 + xs[0];
    p1 = p1 + xs[1];
    p2 = p2 + xs[2];
    p3 = p3 + xs[3];
    p4 = p4 + xs[4];
    p5 = p5 + xs[5];
    p6 = p6 + xs[6];
    p7 = p7 + xs[7];
    xs += 8;
    len -= 8;
  }
  acc = acc + (p0 + p4);
  acc = acc + (p1 + p5);
  acc = acc + (p2 + p6);
  acc = acc + (p3 + p7);
  for (uint64_t i = 0; i < len && i < 7; i++) acc = acc + xs[i];
  return acc;
}

/* mean of an (n x k) array with row stride `stride`: contiguous arrays fold as
 * one slice; strided views fold row by row and add the row sums in order.   */
ORC_API float orc_ndarray_mean_2d(const float *x, uint64_t n, uint64_t k, uint64_t stride) {
  if (n * k == 0) return NAN;
  float s;
  if (stride == k || n == 1) {
    s = orc_unrolled_sum_f32(x, n * k);
  } else {
    s = 0.0f;
    for (uint64_t i = 0; i < n; i++) s = s + orc_unrolled_sum_f32(x + i * stride, k);
  }
  return s / (float)(n * k);
}

/* ------------------------------------------------------------------------- */
/* find_ab_params: umap/find_ab_params.rs:9-82 (f32 gradient descent).        */
ORC_API void orc_find_ab_params(float spread, float min_dist, int32_t n_iterations, float *a_out,
                                float *b_out, int32_t *iters_run) {
  enum { NP = 300 };
  float xv[NP], yv[NP];
  for (int i = 0; i < NP; i++) xv[i] = ((spread * 3.0f) * (float)i) / ((float)NP - 1.0f); /* :18-21 */
  for (int i = 0; i < NP; i++) {                                                         /* :24-32 */
    float x = xv[i];
    yv[i] = (x < min_dist) ? 1.0f : expf(-(x - min_dist) / spread);
  }
  float a = 1.5f, b = 0.9f; /* :37-38 */
  const float learning_rate = 0.01f;
  int it;
  for (it = 0; it < n_iterations; it++) { /* :44 */
    float grad_a = 0.0f, grad_b = 0.0f, total_error = 0.0f;
    for (int i = 0; i < NP; i++) {
      float x = xv[i], y_true = yv[i];
      float x_2b = powf(x, 2.0f * b);
      float denom = 1.0f + a * x_2b;
      float y_pred = 1.0f / denom;
      float error = y_pred - y_true;
      total_error = total_error + error * error;
      grad_a = grad_a + (2.0f * error) * (-x_2b / (denom * denom)); /* :64 */
      if (x > 0.0f) {
        float ln_x = logf(x);
        grad_b = grad_b + (2.0f * error) * ((((-2.0f * a) * x_2b) * ln_x) / (denom * denom)); /* :69 */
      }
    }
    a = a - (learning_rate * grad_a) / (float)NP; /* :74-75 */
    b = b - (learning_rate * grad_b) / (float)NP;
    a = a < 0.001f ? 0.001f : (a > 10.0f ? 10.0f : a); /* :78-79 */
    b = b < 0.001f ? 0.001f : (b > 10.0f ? 10.0f : b);
    if (total_error / (float)NP < 1e-7f) { it++; break; } /* :82 */
  }
  *a_out = a;
  *b_out = b;
  if (iters_run) *iters_run = it;
}

/* ------------------------------------------------------------------------- */
/* smooth_knn_dist: umap/smooth_knn_dist.rs:60-198 (all f32).                 */
static void smooth_row(const float *row, uint64_t n_neighbors, float target, int32_t n_iter,
                       float local_connectivity, float mean_distances, float *sigma_out,
                       float *rho_out) {
  float lo = 0.0f, hi = INFINITY, mid = 1.0f;
  uint64_t non_zero_count = 0;
  float max_non_zero = 0.0f;
  for (uint64_t j = 0; j < n_neighbors; j++) { /* :90-97 */
    float d = row[j];
    if (d > 0.0f) {
      non_zero_count++;
      if (d > max_non_zero) max_non_zero = d;
    }
  }
  float rho = 0.0f;
  /* Rust `as usize` saturates: negative/NaN -> 0 */
  uint64_t local_conn_usize = (local_connectivity > 0.0f) ? (uint64_t)local_connectivity : 0;
  if (non_zero_count >= local_conn_usize) { /* :103 */
    float fl = floorf(local_connectivity);
    uint64_t index = (fl > 0.0f) ? (uint64_t)fl : 0;
    float interpolation = local_connectivity - fl;
    float nth_minus_1 = 0.0f, nth = 0.0f;
    uint64_t count = 0;
    for (uint64_t j = 0; j < n_neighbors; j++) { /* :111-124 */
      float d = row[j];
      if (d > 0.0f) {
        if (count + 1 == index) nth_minus_1 = d;
        if (count + 1 == index + 1) nth = d;
        count++;
        if (count > index + 1) break;
      }
    }
    if (index > 0) { /* :126-139 */
      rho = nth_minus_1;
      if (interpolation > SMOOTH_K_TOLERANCE) rho = rho + interpolation * (nth - nth_minus_1);
    } else {
      float first_non_zero = 0.0f;
      for (uint64_t j = 0; j < n_neighbors; j++)
        if (row[j] > 0.0f) {
          first_non_zero = row[j];
          break;
        }
      rho = interpolation * first_non_zero;
    }
  } else if (non_zero_count > 0) {
    rho = max_non_zero; /* :140-142 */
  }
  for (int32_t it = 0; it < n_iter; it++) { /* :145-171 */
    float psum = 0.0f;
    for (uint64_t j = 1; j < n_neighbors; j++) {
      float d = row[j] - rho;
      if (d > 0.0f)
        psum = psum + expf(-(d / mid));
      else
        psum = psum + 1.0f;
    }
    if (fabsf(psum - target) < SMOOTH_K_TOLERANCE) break;
    if (psum > target) {
      hi = mid;
      mid = (lo + hi) / 2.0f;
    } else {
      lo = mid;
      if (hi == INFINITY)
        mid = mid * 2.0f;
      else
        mid = (lo + hi) / 2.0f;
    }
  }
  float sigma = mid;
  if (rho > 0.0f) { /* :176-183 */
    float mean_ith = orc_unrolled_sum_f32(row, n_neighbors) / (float)n_neighbors;
    if (sigma < MIN_K_DIST_SCALE * mean_ith) sigma = MIN_K_DIST_SCALE * mean_ith;
  } else if (sigma < MIN_K_DIST_SCALE * mean_distances) {
    sigma = MIN_K_DIST_SCALE * mean_distances;
  }
  *sigma_out = sigma;
  *rho_out = rho;
}

ORC_API void orc_smooth_knn_dist(const float *dists, uint64_t n, uint64_t n_neighbors, uint64_t stride,
                                 uint64_t k, int32_t n_iter, float local_connectivity, float bandwidth,
                                 float *sigmas, float *rhos) {
  float target = log2f((float)k) * bandwidth; /* :71 */
  float mean_distances = orc_ndarray_mean_2d(dists, n, n_neighbors, stride); /* :74 */
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t i = 0; i < n; i++)
    smooth_row(dists + i * stride, n_neighbors, target, n_iter, local_connectivity, mean_distances,
               &sigmas[i], &rhos[i]);
}

/* ------------------------------------------------------------------------- */
/* compute_membership_strength: umap/fuzzy_simplicial_set.rs:294-306         */
static inline float membership(float d, float rho, float sigma) {
  if (d - rho <= 0.0f || sigma == 0.0f) return 1.0f;
  return expf(-(d - rho) / sigma);
}

/* build_membership_csr: umap/fuzzy_simplicial_set.rs:171-292.
 * The DashSet of disconnected slots (embedding.rs:177-185) is the predicate
 * dist >= disconnection_distance.  Returns nnz; *indptr/indices/data are
 * malloc'ed (free with orc_free).                                           */
ORC_API uint64_t orc_membership_csr(uint64_t n, uint64_t k, const uint32_t *idx, const float *dists,
                                    uint64_t stride, float disconnection_distance, const float *sigmas,
                                    const float *rhos, uint64_t **indptr_out, uint32_t **indices_out,
                                    float **data_out) {
  uint64_t *indptr = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
  indptr[0] = 0;
  for (uint64_t i = 0; i < n; i++) { /* :182-216 */
    uint64_t c = 0;
    for (uint64_t j = 0; j < k; j++) {
      float d = dists[i * stride + j];
      if (d >= disconnection_distance) continue;
      uint64_t ki = idx[i * stride + j];
      if (ki == i || ki >= n) continue;
      if (membership(d, rhos[i], sigmas[i]) != 0.0f) c++;
    }
    indptr[i + 1] = indptr[i] + c;
  }
  uint64_t nnz = indptr[n];
  uint32_t *indices = (uint32_t *)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
  float *data = (float *)malloc((nnz ? nnz : 1) * sizeof(float));
  for (uint64_t i = 0; i < n; i++) { /* :232-255 */
    uint64_t off = indptr[i];
    for (uint64_t j = 0; j < k; j++) {
      float d = dists[i * stride + j];
      if (d >= disconnection_distance) continue;
      uint32_t ki = idx[i * stride + j];
      if ((uint64_t)ki == i || (uint64_t)ki >= n) continue;
      float v = membership(d, rhos[i], sigmas[i]);
      if (v != 0.0f) {
        indices[off] = ki;
        data[off] = v;
        off++;
      }
    }
    /* insertion sort by column, :264-282 */
    uint64_t s = indptr[i], len = indptr[i + 1] - indptr[i];
    for (uint64_t a = 1; a < len; a++) {
      uint64_t m = a;
      while (m > 0 && indices[s + m - 1] > indices[s + m]) {
        uint32_t ti = indices[s + m - 1];
        indices[s + m - 1] = indices[s + m];
        indices[s + m] = ti;
        float tv = data[s + m - 1];
        data[s + m - 1] = data[s + m];
        data[s + m] = tv;
        m--;
      }
    }
  }
  *indptr_out = indptr;
  *indices_out = indices;
  *data_out = data;
  return nnz;
}

/* csr_get: umap/fuzzy_simplicial_set.rs:373-383 (binary search in a row) */
static float csr_get(const uint64_t *indptr, const uint32_t *indices, const float *data, uint64_t row,
                     uint32_t col) {
  uint64_t lo = indptr[row], hi = indptr[row + 1];
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (indices[mid] < col)
      lo = mid + 1;
    else if (indices[mid] > col)
      hi = mid;
    else
      return data[mid];
  }
  return 0.0f;
}

/* apply_set_operations_parallel: umap/fuzzy_simplicial_set.rs:392-552, with
 * build_csc_structure (:310-370).                                           */
ORC_API uint64_t orc_set_operations(uint64_t n, const uint64_t *indptr, const uint32_t *indices,
                                    const float *data, float set_op_mix_ratio, uint64_t **indptr_out,
                                    uint32_t **indices_out, float **data_out) {
  const float mix = set_op_mix_ratio;
  const float prod_coeff = 1.0f - 2.0f * set_op_mix_ratio; /* :394 */
  uint64_t nnz_in = indptr[n];
  /* CSC structure: column counts, pointers, sequential fill (:316-362) */
  uint64_t *cptr = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
  for (uint64_t e = 0; e < nnz_in; e++) cptr[indices[e] + 1]++;
  for (uint64_t c = 0; c < n; c++) cptr[c + 1] += cptr[c];
  uint32_t *crow = (uint32_t *)malloc((nnz_in ? nnz_in : 1) * sizeof(uint32_t));
  uint64_t *cofs = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
  for (uint64_t r = 0; r < n; r++)
    for (uint64_t e = indptr[r]; e < indptr[r + 1]; e++) {
      uint32_t c = indices[e];
      crow[cptr[c] + cofs[c]] = (uint32_t)r;
      cofs[c]++;
    }
  free(cofs);
  /* row counts (:410-446) */
  uint64_t *optr = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
  optr[0] = 0;
  for (uint64_t r = 0; r < n; r++) {
    uint64_t count = 0;
    for (uint64_t e = indptr[r]; e < indptr[r + 1]; e++) {
      uint32_t col = indices[e];
      float val_rc = data[e];
      float val_cr = csr_get(indptr, indices, data, col, (uint32_t)r);
      float fv = mix * val_rc + mix * val_cr + prod_coeff * val_rc * val_cr;
      if (fv != 0.0f) count++;
    }
    for (uint64_t t = cptr[r]; t < cptr[r + 1]; t++) {
      uint32_t c = crow[t];
      if (csr_get(indptr, indices, data, r, c) != 0.0f) continue;
      float val_cr = csr_get(indptr, indices, data, c, (uint32_t)r);
      float fv = mix * val_cr;
      if (fv != 0.0f) count++;
    }
    optr[r + 1] = optr[r] + count;
  }
  uint64_t nnz = optr[n];
  uint32_t *oind = (uint32_t *)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
  float *odat = (float *)malloc((nnz ? nnz : 1) * sizeof(float));
  /* fill (:475-516) and row sort (:524-542) */
  for (uint64_t r = 0; r < n; r++) {
    uint64_t off = optr[r];
    for (uint64_t e = indptr[r]; e < indptr[r + 1]; e++) {
      uint32_t col = indices[e];
      float val_rc = data[e];
      float val_cr = csr_get(indptr, indices, data, col, (uint32_t)r);
      float fv = mix * val_rc + mix * val_cr + prod_coeff * val_rc * val_cr;
      if (fv != 0.0f) {
        oind[off] = col;
        odat[off] = fv;
        off++;
      }
    }
    for (uint64_t t = cptr[r]; t < cptr[r + 1]; t++) {
      uint32_t c = crow[t];
      if (csr_get(indptr, indices, data, r, c) != 0.0f) continue;
      float val_cr = csr_get(indptr, indices, data, c, (uint32_t)r);
      float fv = mix * val_cr;
      if (fv != 0.0f) {
        oind[off] = c;
        odat[off] = fv;
        off++;
      }
    }
    uint64_t s = optr[r], len = optr[r + 1] - optr[r];
    for (uint64_t a = 1; a < len; a++) {
      uint64_t m = a;
      while (m > 0 && oind[s + m - 1] > oind[s + m]) {
        uint32_t ti = oind[s + m - 1];
        oind[s + m - 1] = oind[s + m];
        oind[s + m] = ti;
        float tv = odat[s + m - 1];
        odat[s + m - 1] = odat[s + m];
        odat[s + m] = tv;
        m--;
      }
    }
  }
  free(cptr);
  free(crow);
  *indptr_out = optr;
  *indices_out = oind;
  *data_out = odat;
  return nnz;
}

/* Rows whose weights sum to exactly 0 (embedding.rs:216-222, :327-334); the
 * Rust f32 `iter().sum()` is a sequential fold.  flags may be NULL.          */
ORC_API uint64_t orc_zero_rows(uint64_t n, const uint64_t *indptr, const float *data, uint8_t *flags) {
  uint64_t cnt = 0;
  for (uint64_t r = 0; r < n; r++) {
    float s = -0.0f;
    for (uint64_t e = indptr[r]; e < indptr[r + 1]; e++) s = s + data[e];
    int z = (s == 0.0f);
    if (flags) flags[r] = (uint8_t)z;
    cnt += (uint64_t)z;
  }
  return cnt;
}

/* ------------------------------------------------------------------------- */
/* Optimizer::new preprocessing: optimizer.rs:74-290 and
 * make_epochs_per_sample.rs:21-52.  Outputs malloc'ed edge arrays; the
 * embedding is normalised in place.                                          */
ORC_API uint64_t orc_optimizer_prepare(uint64_t n, const uint64_t *indptr, const uint32_t *indices,
                                       const float *data, uint64_t total_epochs, uint64_t negative_sample_rate,
                                       float *embedding, uint64_t dim, uint32_t **head_out, uint32_t **tail_out,
                                       double **eps_out, double **eons_out, double **epns_out,
                                       double **eonns_out, float *threshold_out) {
  uint64_t nnz = indptr[n];
  float max_val = 0.0f; /* :90-94, Rust f32::max ignores NaN */
  for (uint64_t e = 0; e < nnz; e++) max_val = fmaxf(max_val, data[e]);
  uint64_t default_epochs = (n <= 10000) ? 500 : 200; /* :96 */
  uint64_t threshold_epochs = (total_epochs > 10) ? total_epochs : default_epochs;
  float threshold = max_val / (float)threshold_epochs; /* :102 */
  if (threshold_out) *threshold_out = threshold;
  uint64_t total = 0;
  for (uint64_t e = 0; e < nnz; e++)
    if (data[e] >= threshold) total++;
  uint32_t *head = (uint32_t *)malloc((total ? total : 1) * 4);
  uint32_t *tail = (uint32_t *)malloc((total ? total : 1) * 4);
  float *w = (float *)malloc((total ? total : 1) * 4);
  uint64_t o = 0;
  for (uint64_t r = 0; r < n; r++) /* :139-159 */
    for (uint64_t e = indptr[r]; e < indptr[r + 1]; e++)
      if (data[e] >= threshold) {
        head[o] = (uint32_t)r;
        tail[o] = indices[e];
        w[o] = data[e];
        o++;
      }
  /* make_epochs_per_sample */
  float max_weight = -INFINITY;
  for (uint64_t e = 0; e < total; e++) max_weight = fmaxf(max_weight, w[e]);
  if (max_weight <= 0.0f) max_weight = 1.0f;
  double n_epochs_f64 = (double)total_epochs, max_w64 = (double)max_weight;
  double *eps = (double *)malloc((total ? total : 1) * 8);
  double *eons = (double *)malloc((total ? total : 1) * 8);
  double *epns = (double *)malloc((total ? total : 1) * 8);
  double *eonns = (double *)malloc((total ? total : 1) * 8);
  double neg_rate = (double)negative_sample_rate;
  for (uint64_t e = 0; e < total; e++) {
    double ns = n_epochs_f64 * ((double)w[e] / max_w64);
    eps[e] = (ns > 0.0) ? n_epochs_f64 / ns : -1.0;
    eons[e] = eps[e];               /* :242 */
    epns[e] = eps[e] / neg_rate;    /* :249-254 */
    eonns[e] = epns[e];             /* :261-268 */
  }
  free(w);
  /* normalisation to [0, 10] per dimension: :184-230 */
  float *mins = (float *)malloc(dim * 4), *maxs = (float *)malloc(dim * 4), *scales = (float *)malloc(dim * 4);
  for (uint64_t d = 0; d < dim; d++) {
    mins[d] = INFINITY;
    maxs[d] = -INFINITY;
  }
  for (uint64_t i = 0; i < n; i++)
    for (uint64_t d = 0; d < dim; d++) {
      mins[d] = fminf(mins[d], embedding[i * dim + d]);
      maxs[d] = fmaxf(maxs[d], embedding[i * dim + d]);
    }
  for (uint64_t d = 0; d < dim; d++) {
    float range = maxs[d] - mins[d];
    scales[d] = (range > 0.0f) ? 10.0f / range : 0.0f;
  }
  for (uint64_t i = 0; i < n; i++)
    for (uint64_t d = 0; d < dim; d++)
      if (scales[d] > 0.0f) embedding[i * dim + d] = (embedding[i * dim + d] - mins[d]) * scales[d];
  free(mins);
  free(maxs);
  free(scales);
  *head_out = head;
  *tail_out = tail;
  *eps_out = eps;
  *eons_out = eons;
  *epns_out = epns;
  *eonns_out = eonns;
  return total;
}

/* ------------------------------------------------------------------------- */
/* Counter-based negative-sample RNG.  The reference draws from rand 0.9's
 * OS-seeded ThreadRng (optimize_layout_euclidean.rs:269,314), which no run can
 * reproduce; this build replaces it with SplitMix64 (Steele, Lea & Flood 2014,
 * "Fast splittable pseudorandom number generators") evaluated at a counter:
 *   epoch_key = mix(seed + (epoch+1)*G)         (per epoch)
 *   edge_key  = mix(epoch_key ^ ((edge+1)*G2))  (per edge activation)
 *   draw p    = mix(edge_key + (p+1)*G)         (per negative sample)
 * and maps the 64-bit draw to [0,n) as floor(r * n / 2^64).                 */
#define SM_G 0x9e3779b97f4a7c15ULL
#define SM_G2 0xd1b54a32d192ed03ULL
static inline uint64_t sm_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
ORC_API uint64_t orc_rng_epoch_key(uint64_t seed, uint64_t epoch) { return sm_mix(seed + (epoch + 1) * SM_G); }
ORC_API uint64_t orc_rng_edge_key(uint64_t epoch_key, uint64_t edge) { return sm_mix(epoch_key ^ ((edge + 1) * SM_G2)); }
ORC_API uint32_t orc_rng_index(uint64_t edge_key, uint64_t p, uint32_t n) {
  uint64_t r = sm_mix(edge_key + (p + 1) * SM_G);
  uint64_t hi = r >> 32, lo = r & 0xffffffffULL;
  uint64_t t = (lo * (uint64_t)n) >> 32;
  return (uint32_t)((hi * (uint64_t)n + t) >> 32);
}

/* clip: utils/clip.rs:13-15 (f32::clamp keeps NaN) */
static inline float clip4(float v) { return v < -4.0f ? -4.0f : (v > 4.0f ? 4.0f : v); }

/* Rust `f64 as usize` saturates (NaN/negative -> 0). */
static inline uint64_t sat_u64(double q) {
  if (!(q > 0.0)) return 0;
  if (q >= 4294967295.0) return 4294967295ULL; /* bounded: a larger count is a degenerate schedule */
  return (uint64_t)q;
}

/* Force denominators: the reference's a*d2^b*d2 + 1 (ref_den = 1,
 * optimize_layout_euclidean.rs:293, :329) or umap-learn's a*d2^b + 1. */
static int g_ref_den = 1;

/* Test hook for the single-epoch entry points: 1 = reference denominators. */
ORC_API void orc_set_force_denominator(int32_t reference) { g_ref_den = reference ? 1 : 0; }

/* One edge of the Euclidean SGD epoch: optimize_layout_euclidean.rs:277-344
 * (parallel body; the serial body :168-231 is the same arithmetic).         */
static inline void sgd_edge(float *head_emb, float *tail_emb, uint64_t dim, uint32_t j, uint32_t k, uint64_t i,
                            uint32_t n_vertices, const double *eps, double *eons, const double *epns,
                            double *eonns, float a, float b, float gamma, float alpha, uint64_t epoch,
                            uint64_t epoch_key, uint64_t *n_pos, uint64_t *n_neg_drawn) {
  if (!(eons[i] <= (double)epoch)) return;
  float *cur = head_emb + (uint64_t)j * dim;
  float *oth = tail_emb + (uint64_t)k * dim;
  float d2 = 0.0f;
  for (uint64_t d = 0; d < dim; d++) {
    float diff = cur[d] - oth[d];
    d2 = d2 + diff * diff;
  }
  float gc = 0.0f;
  if (d2 > 0.0f) {
    float pb = powf(d2, b);
    gc = (((-2.0f * a) * b) * pb) / d2;
    gc = gc / ((a * pb) * (g_ref_den ? d2 : 1.0f) + 1.0f);
  }
  for (uint64_t d = 0; d < dim; d++) {
    float diff = cur[d] - oth[d];
    float g = clip4(gc * diff);
    cur[d] = cur[d] + g * alpha;
    oth[d] = oth[d] - g * alpha; /* move_other = true (optimizer.rs:338) */
  }
  eons[i] = eons[i] + eps[i];
  uint64_t n_neg = sat_u64(((double)epoch - eonns[i]) / epns[i]);
  uint64_t ekey = orc_rng_edge_key(epoch_key, i);
  for (uint64_t p = 0; p < n_neg; p++) {
    uint32_t kk = orc_rng_index(ekey, p, n_vertices);
    if (kk == j) continue;
    float *o2 = tail_emb + (uint64_t)kk * dim;
    float dd = 0.0f;
    for (uint64_t d = 0; d < dim; d++) {
      float diff = cur[d] - o2[d];
      dd = dd + diff * diff;
    }
    float g2 = 0.0f;
    if (dd > 0.0f) {
      float pb = powf(dd, b);
      g2 = ((2.0f * gamma) * b) / ((0.001f + dd) * ((a * pb) * (g_ref_den ? dd : 1.0f) + 1.0f));
    }
    if (g2 > 0.0f)
      for (uint64_t d = 0; d < dim; d++) {
        float diff = cur[d] - o2[d];
        float g = clip4(g2 * diff);
        cur[d] = cur[d] + g * alpha;
      }
  }
  eonns[i] = eonns[i] + (double)n_neg * epns[i];
  *n_pos += 1;
  *n_neg_drawn += n_neg;
}

/* One epoch in serial edge order (the deterministic mode). */
ORC_API void orc_sgd_epoch_serial(float *head_emb, float *tail_emb, uint64_t dim, const uint32_t *head,
                                  const uint32_t *tail, uint64_t n_edges, uint32_t n_vertices, const double *eps,
                                  double *eons, const double *epns, double *eonns, float a, float b, float gamma,
                                  float alpha, uint64_t epoch, uint64_t seed, uint64_t *counts) {
  uint64_t ekey = orc_rng_epoch_key(seed, epoch);
  uint64_t np = 0, nn = 0;
  for (uint64_t i = 0; i < n_edges; i++)
    sgd_edge(head_emb, tail_emb, dim, head[i], tail[i], i, n_vertices, eps, eons, epns, eonns, a, b, gamma, alpha,
             epoch, ekey, &np, &nn);
  if (counts) {
    counts[0] += np;
    counts[1] += nn;
  }
}

/* One epoch, Hogwild over OpenMP threads with contiguous edge chunks (the
 * restatement of the Rayon `into_par_iter` loop, :268-346): racy plain f32
 * loads/stores on both embeddings, exactly as the reference allows.       */
ORC_API void orc_sgd_epoch_hogwild(float *head_emb, float *tail_emb, uint64_t dim, const uint32_t *head,
                                   const uint32_t *tail, uint64_t n_edges, uint32_t n_vertices, const double *eps,
                                   double *eons, const double *epns, double *eonns, float a, float b, float gamma,
                                   float alpha, uint64_t epoch, uint64_t seed, uint64_t *counts) {
  uint64_t ekey = orc_rng_epoch_key(seed, epoch);
  uint64_t np = 0, nn = 0;
#pragma omp parallel for schedule(static) reduction(+ : np, nn)
  for (uint64_t i = 0; i < n_edges; i++)
    sgd_edge(head_emb, tail_emb, dim, head[i], tail[i], i, n_vertices, eps, eons, epns, eonns, a, b, gamma, alpha,
             epoch, ekey, &np, &nn);
  if (counts) {
    counts[0] += np;
    counts[1] += nn;
  }
}

/* One epoch of the build's row-owner synchronous mode (the device default on
 * symmetric schedules).  Not a reference function: it restates the device
 * algorithm so that its output can be checked bit for bit.  Every read comes
 * from h0/t0, copies of the head and tail arrays taken at the start of the
 * epoch.  For each active edge i = (j,k) of row j, in edge order:
 *   own sample  h = h0[j] + clip(gc(h0[j], t0[k]) * (h0[j] - t0[k])) * alpha,
 *               then its negative samples on h against t0 (rs:278-341);
 *   twin tail   dt = -clip(gc(h0[k], t0[j]) * (h0[k] - t0[j])) * alpha, the
 *               tail side of the twin edge (k,j) (rs:284-306), which the
 *               owner of row k evaluates from the same values for its head;
 * and the row is written once: head[j] = h0[j] + ((0 + dh_1) + dh_2 ...),
 * tail[j] = ((t0[j] + dt_1) + dt_2) ...; with one shared array,
 * head[j] = (h0[j] + sum dh) + (t - t0[j]).                                 */
static float euclid_dist_grad(const float *x, const float *y, uint64_t dim, float *grad);
static void sgd_epoch_generic_serial(float *head_emb, float *tail_emb, uint64_t dim, const uint32_t *head,
                                     const uint32_t *tail, uint64_t n_edges, uint32_t n_vertices, const double *eps,
                                     double *eons, const double *epns, double *eonns, float a, float b, float gamma,
                                     float alpha, uint64_t epoch, uint64_t seed, uint64_t *counts);

/* The generic-metric forms (optimize_layout_generic.rs:64-118) with the
 * Euclidean metric's distance and gradient (distances.rs:25-41). */
static inline float generic_attract_coeff(float dist, float a, float b) {
  float w_l = (dist > 0.0f) ? 1.0f / (1.0f + a * powf(dist, 2.0f * b)) : 1.0f;
  return ((2.0f * b) * (w_l - 1.0f)) / (dist + 1e-6f);
}
static inline float generic_repel_coeff(float dist, float a, float b, float gamma) {
  float wl = (dist > 0.0f) ? 1.0f / (1.0f + a * powf(dist, 2.0f * b)) : 1.0f;
  return (((gamma * 2.0f) * b) * wl) / (dist + 1e-6f);
}

static void sgd_epoch_sync(float *head_emb, float *tail_emb, int shared, int generic, uint64_t dim,
                           const uint32_t *head, const uint32_t *tail, uint64_t n_edges, uint32_t n_vertices,
                           const double *eps, double *eons, const double *epns, double *eonns, float a, float b,
                           float gamma, float alpha, uint64_t epoch, uint64_t seed, uint64_t *counts) {
  if (dim > 16) {
    fprintf(stderr, "orc_sgd_epoch_sync: dim %llu > 16\n", (unsigned long long)dim);
    abort();
  }
  const uint64_t nel = (uint64_t)n_vertices * dim;
  float *h0 = (float *)malloc((nel ? nel : 1) * sizeof(float));
  float *t0 = shared ? h0 : (float *)malloc((nel ? nel : 1) * sizeof(float));
  memcpy(h0, head_emb, nel * sizeof(float));
  if (!shared) memcpy(t0, tail_emb, nel * sizeof(float));
  uint64_t *rp = (uint64_t *)calloc((uint64_t)n_vertices + 1, sizeof(uint64_t));
  for (uint64_t i = 0; i < n_edges; i++) rp[head[i] + 1]++;
  for (uint64_t r = 0; r < n_vertices; r++) rp[r + 1] += rp[r];
  const uint64_t ekey = orc_rng_epoch_key(seed, epoch);
  const float m2ab = (-2.0f * a) * b, g2b = (2.0f * gamma) * b;
  uint64_t np = 0, nn = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : np, nn)
  for (uint64_t j = 0; j < n_vertices; j++) {
    float acc[16], t[16], h[16], dif[16];
    int any = 0;
    for (uint64_t d = 0; d < dim; d++) {
      acc[d] = 0.0f;
      t[d] = t0[j * dim + d];
    }
    for (uint64_t i = rp[j]; i < rp[j + 1]; i++) {
      if (!(eons[i] <= (double)epoch)) continue;
      any = 1;
      const uint32_t k = tail[i];
      eons[i] = eons[i] + eps[i];
      const uint64_t n_neg = sat_u64(((double)epoch - eonns[i]) / epns[i]);
      eonns[i] = eonns[i] + (double)n_neg * epns[i];
      const uint64_t ik = orc_rng_edge_key(ekey, i);
      if (generic) {
        float g1[16], g2v[16];
        const float dist = euclid_dist_grad(h0 + j * dim, t0 + (uint64_t)k * dim, dim, g1);
        const float gc = generic_attract_coeff(dist, a, b);
        for (uint64_t d = 0; d < dim; d++) h[d] = h0[j * dim + d] + clip4(gc * g1[d]) * alpha;
        for (uint64_t p = 0; p < n_neg; p++) {
          const uint32_t kk = orc_rng_index(ik, p, n_vertices);
          if (kk == j) continue;
          const float dd = euclid_dist_grad(h, t0 + (uint64_t)kk * dim, dim, g1);
          const float gcn = generic_repel_coeff(dd, a, b, gamma);
          for (uint64_t d = 0; d < dim; d++) h[d] = h[d] + clip4(gcn * g1[d]) * alpha;
        }
        for (uint64_t d = 0; d < dim; d++) acc[d] = acc[d] + (h[d] - h0[j * dim + d]);
        /* twin (k,j): current = h0[k], other = t0[j]; the other side moves by
         * clip(gc * grad_other) * alpha, grad_other from distance(other, current) */
        const float dk = euclid_dist_grad(h0 + (uint64_t)k * dim, t0 + j * dim, dim, g1);
        (void)euclid_dist_grad(t0 + j * dim, h0 + (uint64_t)k * dim, dim, g2v);
        const float gck = generic_attract_coeff(dk, a, b);
        for (uint64_t d = 0; d < dim; d++) t[d] = t[d] + clip4(gck * g2v[d]) * alpha;
      } else {
        float d2 = 0.0f;
        for (uint64_t d = 0; d < dim; d++) {
          dif[d] = h0[j * dim + d] - t0[(uint64_t)k * dim + d];
          d2 = d2 + dif[d] * dif[d];
        }
        float gc = 0.0f;
        if (d2 > 0.0f) {
          const float pb = powf(d2, b);
          gc = (m2ab * pb) / d2;
          gc = gc / ((a * pb) * (g_ref_den ? d2 : 1.0f) + 1.0f);
        }
        for (uint64_t d = 0; d < dim; d++) h[d] = h0[j * dim + d] + clip4(gc * dif[d]) * alpha;
        for (uint64_t p = 0; p < n_neg; p++) {
          const uint32_t kk = orc_rng_index(ik, p, n_vertices);
          if (kk == j) continue;
          float dd = 0.0f;
          for (uint64_t d = 0; d < dim; d++) {
            dif[d] = h[d] - t0[(uint64_t)kk * dim + d];
            dd = dd + dif[d] * dif[d];
          }
          float g2 = 0.0f;
          if (dd > 0.0f) {
            const float pb = powf(dd, b);
            g2 = g2b / ((0.001f + dd) * ((a * pb) * (g_ref_den ? dd : 1.0f) + 1.0f));
          }
          if (g2 > 0.0f)
            for (uint64_t d = 0; d < dim; d++) h[d] = h[d] + clip4(g2 * dif[d]) * alpha;
        }
        for (uint64_t d = 0; d < dim; d++) acc[d] = acc[d] + (h[d] - h0[j * dim + d]);
        d2 = 0.0f;
        for (uint64_t d = 0; d < dim; d++) {
          dif[d] = h0[(uint64_t)k * dim + d] - t0[j * dim + d];
          d2 = d2 + dif[d] * dif[d];
        }
        gc = 0.0f;
        if (d2 > 0.0f) {
          const float pb = powf(d2, b);
          gc = (m2ab * pb) / d2;
          gc = gc / ((a * pb) * (g_ref_den ? d2 : 1.0f) + 1.0f);
        }
        for (uint64_t d = 0; d < dim; d++) t[d] = t[d] + (-(clip4(gc * dif[d]) * alpha));
      }
      np += 1;
      nn += n_neg;
    }
    if (!any) continue;
    for (uint64_t d = 0; d < dim; d++) {
      const float hv = h0[j * dim + d] + acc[d];
      if (shared) {
        head_emb[j * dim + d] = hv + (t[d] - t0[j * dim + d]);
      } else {
        head_emb[j * dim + d] = hv;
        tail_emb[j * dim + d] = t[d];
      }
    }
  }
  free(rp);
  if (!shared) free(t0);
  free(h0);
  if (counts) {
    counts[0] += np;
    counts[1] += nn;
  }
}

ORC_API void orc_sgd_epoch_sync(float *head_emb, float *tail_emb, int shared, uint64_t dim, const uint32_t *head,
                                const uint32_t *tail, uint64_t n_edges, uint32_t n_vertices, const double *eps,
                                double *eons, const double *epns, double *eonns, float a, float b, float gamma,
                                float alpha, uint64_t epoch, uint64_t seed, uint64_t *counts) {
  sgd_epoch_sync(head_emb, tail_emb, shared, 0, dim, head, tail, n_edges, n_vertices, eps, eons, epns, eonns, a, b,
                 gamma, alpha, epoch, seed, counts);
}

/* The schedule recurrence of every epoch alone (optimize_layout_euclidean.rs:
 * 277 activity test, :308 eons += eps, :310-311 n_neg = (epoch - eonns) /
 * epns as usize, :343 eonns += n_neg * epns), epochs [start, start + n) for
 * every edge.  It draws no random numbers and moves no embedding, so it gives
 * the exact positive and negative counts and the final schedule of any mode.
 * counts[0] += positive updates, counts[1] += negative samples. */
ORC_API void orc_schedule_replay(const double *eps, const double *epns, double *eons, double *eonns,
                                 uint64_t n_edges, uint64_t start_epoch, uint64_t n_epochs, uint64_t *counts) {
  uint64_t np = 0, nn = 0;
#pragma omp parallel for schedule(static) reduction(+ : np, nn)
  for (uint64_t i = 0; i < n_edges; i++) {
    double a = eons[i], c = eonns[i];
    for (uint64_t e = start_epoch; e < start_epoch + n_epochs; e++) {
      if (!(a <= (double)e)) continue;
      a = a + eps[i];
      const uint64_t k = sat_u64(((double)e - c) / epns[i]);
      c = c + (double)k * epns[i];
      np += 1;
      nn += k;
    }
    eons[i] = a;
    eonns[i] = c;
  }
  counts[0] += np;
  counts[1] += nn;
}

/* Optimizer::step_epochs (optimizer.rs:297-366): tail snapshot per call,
 * alpha = lr * (1 - epoch/total) in f32.  mode 0 = serial, 1 = Hogwild, 2 = the
 * build's synchronous row-owner epoch (orc_sgd_epoch_sync), 3 = that epoch with
 * the generic-metric forms, 4 = the serial generic-metric epoch
 * (optimize_layout_generic.rs:35-124, Euclidean metric).
 * semantics: bit 1 = one shared array instead of the tail snapshot, bit 2 =
 * umap-learn's force denominators (both opt-in departures; 0 = reference). */
ORC_API void orc_step_epochs(float *embedding, uint64_t n, uint64_t dim, const uint32_t *head, const uint32_t *tail,
                             uint64_t n_edges, const double *eps, double *eons, const double *epns, double *eonns,
                             float a, float b, float gamma, float initial_alpha, uint64_t start_epoch,
                             uint64_t n_epochs, uint64_t total_epochs, uint64_t seed, int32_t mode,
                             uint64_t *counts, int32_t semantics) {
  uint64_t nel = n * dim;
  const int shared = (semantics & 1) != 0;
  g_ref_den = (semantics & 2) ? 0 : 1;
  float *copy = shared ? embedding : (float *)malloc((nel > 0 ? nel : 1) * sizeof(float));
  if (!shared) memcpy(copy, embedding, n * dim * sizeof(float));
  for (uint64_t e = start_epoch; e < start_epoch + n_epochs; e++) {
    float alpha = initial_alpha * (1.0f - ((float)e / (float)total_epochs));
    if (mode == 0)
      orc_sgd_epoch_serial(embedding, copy, dim, head, tail, n_edges, (uint32_t)n, eps, eons, epns, eonns, a, b,
                           gamma, alpha, e, seed, counts);
    else if (mode == 2 || mode == 3)
      sgd_epoch_sync(embedding, copy, shared, mode == 3, dim, head, tail, n_edges, (uint32_t)n, eps, eons, epns,
                     eonns, a, b, gamma, alpha, e, seed, counts);
    else if (mode == 4)
      sgd_epoch_generic_serial(embedding, copy, dim, head, tail, n_edges, (uint32_t)n, eps, eons, epns, eonns, a, b,
                               gamma, alpha, e, seed, counts);
    else
      orc_sgd_epoch_hogwild(embedding, copy, dim, head, tail, n_edges, (uint32_t)n, eps, eons, epns, eonns, a, b,
                            gamma, alpha, e, seed, counts);
  }
  if (!shared) free(copy);
  g_ref_den = 1;
}

/* Generic-metric epoch: layout/optimize_layout_generic.rs:35-124 with the
 * Euclidean metric's distance/gradient (distances.rs:25-41).  Serial only. */
static float euclid_dist_grad(const float *x, const float *y, uint64_t dim, float *grad) {
  float s = 0.0f;
  for (uint64_t i = 0; i < dim; i++) {
    float diff = x[i] - y[i];
    s = s + diff * diff;
  }
  float dist = sqrtf(s);
  float denom = dist + 1e-6f;
  for (uint64_t i = 0; i < dim; i++) grad[i] = (x[i] - y[i]) / denom;
  return dist;
}

static void sgd_epoch_generic_serial(float *head_emb, float *tail_emb, uint64_t dim, const uint32_t *head,
                                     const uint32_t *tail, uint64_t n_edges, uint32_t n_vertices, const double *eps,
                                     double *eons, const double *epns, double *eonns, float a, float b, float gamma,
                                     float alpha, uint64_t epoch, uint64_t seed, uint64_t *counts) {
  uint64_t ekey = orc_rng_epoch_key(seed, epoch);
  uint64_t np = 0, nn = 0;
  float *g1 = (float *)malloc(dim * 4), *g2 = (float *)malloc(dim * 4);
  for (uint64_t i = 0; i < n_edges; i++) {
    if (!(eons[i] <= (double)epoch)) continue;
    uint32_t j = head[i], k = tail[i];
    float *cur = head_emb + (uint64_t)j * dim, *oth = tail_emb + (uint64_t)k * dim;
    float dist = euclid_dist_grad(cur, oth, dim, g1);
    (void)euclid_dist_grad(oth, cur, dim, g2);
    float w_l = (dist > 0.0f) ? 1.0f / (1.0f + a * powf(dist, 2.0f * b)) : 1.0f;
    float gc = ((2.0f * b) * (w_l - 1.0f)) / (dist + 1e-6f);
    for (uint64_t d = 0; d < dim; d++) {
      float g = clip4(gc * g1[d]);
      cur[d] = cur[d] + g * alpha;
      g = clip4(gc * g2[d]);
      oth[d] = oth[d] + g * alpha;
    }
    eons[i] = eons[i] + eps[i];
    uint64_t n_neg = sat_u64(((double)epoch - eonns[i]) / epns[i]);
    uint64_t ek = orc_rng_edge_key(ekey, i);
    for (uint64_t p = 0; p < n_neg; p++) {
      uint32_t kk = orc_rng_index(ek, p, n_vertices);
      if (kk == j) continue;
      float *o2 = tail_emb + (uint64_t)kk * dim;
      float dd = euclid_dist_grad(cur, o2, dim, g1);
      float wl = (dd > 0.0f) ? 1.0f / (1.0f + a * powf(dd, 2.0f * b)) : 1.0f;
      float gcn = (((gamma * 2.0f) * b) * wl) / (dd + 1e-6f);
      for (uint64_t d = 0; d < dim; d++) {
        float g = clip4(gcn * g1[d]);
        cur[d] = cur[d] + g * alpha;
      }
    }
    eonns[i] = eonns[i] + (double)n_neg * epns[i];
    np += 1;
    nn += n_neg;
  }
  free(g1);
  free(g2);
  if (counts) {
    counts[0] += np;
    counts[1] += nn;
  }
}

ORC_API void orc_sgd_epoch_generic_euclid(float *head_emb, float *tail_emb, uint64_t dim, const uint32_t *head,
                                          const uint32_t *tail, uint64_t n_edges, uint32_t n_vertices,
                                          const double *eps, double *eons, const double *epns, double *eonns,
                                          float a, float b, float gamma, float alpha, uint64_t epoch, uint64_t seed) {
  sgd_epoch_generic_serial(head_emb, tail_emb, dim, head, tail, n_edges, n_vertices, eps, eons, epns, eonns, a, b,
                           gamma, alpha, epoch, seed, NULL);
}

ORC_API int32_t orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
