// Runtime-size matrix helpers behind parsmooth._utils (reference
// parsmooth/_utils.py): one lane per batch item, each item's matrices row-major
// and contiguous in device memory.  They serve the Python utilities and their
// tests; the passes hold their blocks in registers at compile-time sizes
// (psm_linalg.hpp) and use the same arithmetic.
#pragma once

#include "psm_linalg.hpp"

namespace psm {

constexpr int kUtilWG = 64;

// tria (_utils.py:22-24): A (rows x cols) triangularised in place from the right
// by Householder reflections, as tria_inplace does at compile-time sizes: on exit
// its first min(rows, cols) columns hold L (lower, non-negative diagonal) with
// L L^T = A A^T and the other columns are zero.
template <typename R>
__global__ __launch_bounds__(kUtilWG) void k_util_tria(int64_t nb, int rows, int cols,
                                                       R* __restrict__ A) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  R* M = A + b * (int64_t)rows * cols;
  const int nr = rows < cols ? rows : cols;
  for (int i = 0; i < nr; ++i) {
    R* Mi = M + (int64_t)i * cols;
    const R alpha = Mi[i];
    R sigma = R(0);
    for (int j = i + 1; j < cols; ++j) sigma = fma(Mi[j], Mi[j], sigma);
    const R nrm = sqrt_(fma(alpha, alpha, sigma));
    // v = x - ||x|| e0 without cancellation (Parlett's formula)
    const R v0 = (alpha > R(0)) ? -sigma * rcp_(alpha + nrm) : alpha - nrm;
    const R vv = fma(v0, v0, sigma);
    const R f = (vv > R(0)) ? R(2) * rcp_(vv) : R(0);
    for (int k = i + 1; k < rows; ++k) {
      R* Mk = M + (int64_t)k * cols;
      R dot = v0 * Mk[i];
      for (int j = i + 1; j < cols; ++j) dot = fma(Mi[j], Mk[j], dot);
      const R t = f * dot;
      Mk[i] = fma(-t, v0, Mk[i]);
      for (int j = i + 1; j < cols; ++j) Mk[j] = fma(-t, Mi[j], Mk[j]);
    }
    Mi[i] = nrm;
    for (int j = i + 1; j < cols; ++j) Mi[j] = R(0);
  }
}

// _cholesky_update / cholesky_update_many (_utils.py:13-20, 39-81): the lower
// factor L (n x n, in place) of L L^T + mult v v^T for each of the item's nv
// vectors in turn, by the column recurrence of cholesky_update; after each
// vector the entries above the diagonal are dropped and non-finite entries set
// to zero.  omega: n values of scratch per item.
template <typename R>
__global__ __launch_bounds__(kUtilWG) void k_util_chol_update(int64_t nb, int n, int nv,
                                                              R* __restrict__ L,
                                                              const R* __restrict__ V,
                                                              R* __restrict__ omega, R mult) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  R* Lb = L + b * (int64_t)n * n;
  R* w = omega + b * (int64_t)n;
  for (int k = 0; k < nv; ++k) {
    const R* v = V + (b * (int64_t)nv + k) * n;
    for (int i = 0; i < n; ++i) w[i] = v[i];
    R bb = R(1);
    for (int j = 0; j < n; ++j) {
      const R d = Lb[(int64_t)j * n + j];
      const R wj = w[j];
      const R new_d = sqrt_(d * d + mult / bb * wj * wj);
      const R gamma = d * d * bb + mult * wj * wj;
      const R r = wj / d;
      const R s = mult * wj / gamma;
      for (int i = 0; i < n; ++i) w[i] = fma(-r, Lb[(int64_t)i * n + j], w[i]);
      for (int i = j + 1; i < n; ++i) {
        R& x = Lb[(int64_t)i * n + j];
        x = new_d * (x / d + s * w[i]);
      }
      Lb[(int64_t)j * n + j] = new_d;
      bb = bb + mult * r * r;
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        R& x = Lb[(int64_t)i * n + j];
        x = (j > i || !isfinite(x)) ? R(0) : x;
      }
  }
}

// mvn_loglikelihood (_utils.py:152-160): log N(x; 0, L L^T) for each item, by
// forward substitution, in FP64 whatever the input type.  y: n doubles of
// scratch per item.
template <typename R>
__global__ __launch_bounds__(kUtilWG) void k_util_mvn_ll(int64_t nb, int n, const R* __restrict__ x,
                                                         const R* __restrict__ L,
                                                         double* __restrict__ out,
                                                         double* __restrict__ y) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const R* Lb = L + b * (int64_t)n * n;
  const R* xb = x + b * (int64_t)n;
  double* yb = y + b * (int64_t)n;
  double quad = 0.0, logdet = 0.0;
  for (int i = 0; i < n; ++i) {
    double s = (double)xb[i];
    for (int j = 0; j < i; ++j) s = fma(-(double)Lb[(int64_t)i * n + j], yb[j], s);
    const double d = (double)Lb[(int64_t)i * n + i];
    const double yi = s / d;
    yb[i] = yi;
    quad = fma(yi, yi, quad);
    logdet += log(fabs(d));
  }
  out[b] = -0.5 * quad - logdet - 0.5 * n * 1.8378770664093453;  // log(2 pi)
}

}  // namespace psm
