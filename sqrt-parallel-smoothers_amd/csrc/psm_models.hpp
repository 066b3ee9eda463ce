// Built-in state-space model functions and the per-point linearizations
// (extended / sigma-point statistical linear regression) evaluated on device.
//
// A Python callable cannot run inside a kernel, so the library carries a small
// registry of the functions the reference's own tests, notebooks and
// benchmarks use (tests/_lgssm.py:5-39, tests/bearings/bearings_utils.py:7-69,
// tests/test_linearization.py:157), each with an analytic Jacobian.
#pragma once
#include "psm_linalg.hpp"
#include "psm_dual.hpp"
#include "psm.h"

namespace psm {

// Device view of a psm_model_t (see include/psm.h).
// S is the storage type of the arrays, R the arithmetic type the values are
// read as (a wider R recomputes an FP32 linearization in FP64).
template <typename R, typename S = R>
struct ModelView {
  int kind;
  const S* params;   // model parameters (device)
  const S* q_mean;   // additive noise mean (device, nout)
  const S* q_chol;   // noise chol (sqrt form) or covariance (covariance form), nout x nout
  // PSM_MODEL_TABULATED: the row (linearization point) and sigma point read,
  // and the number of sigma points per row (0: rows hold f(m) and its Jacobian)
  int64_t row = 0;
  int pt = 0;
  int npts = 0;
  PSM_DEV R par(int i) const { return R(params[i]); }
  PSM_DEV R par64(int64_t i) const { return R(params[i]); }
  PSM_DEV R qm(int i) const { return R(q_mean[i]); }
  PSM_DEV R qc(int i) const { return R(q_chol[i]); }
};

PSM_DEV double atan2_(double y, double x) { return atan2(y, x); }
PSM_DEV float atan2_(float y, float x) { return atan2f(y, x); }
PSM_DEV double sin_(double x) { return sin(x); }
PSM_DEV float sin_(float x) { return sinf(x); }
PSM_DEV double cos_(double x) { return cos(x); }
PSM_DEV float cos_(float x) { return cosf(x); }
PSM_DEV double exp_(double x) { return exp(x); }
PSM_DEV float exp_(float x) { return expf(x); }
PSM_DEV double atan_(double x) { return atan(x); }
PSM_DEV float atan_(float x) { return atanf(x); }

// tangent rules of the model functions
template <typename B> PSM_DEV DualT<B> sin_(DualT<B> a) { return {sin_(a.v), cos_(a.v) * a.d}; }
template <typename B> PSM_DEV DualT<B> cos_(DualT<B> a) { return {cos_(a.v), -sin_(a.v) * a.d}; }
template <typename B> PSM_DEV DualT<B> exp_(DualT<B> a) {
  const B e = exp_(a.v);
  return {e, e * a.d};
}
template <typename B> PSM_DEV DualT<B> atan_(DualT<B> a) { return {atan_(a.v), a.d / (B(1) + a.v * a.v)}; }
template <typename B> PSM_DEV DualT<B> atan2_(DualT<B> y, DualT<B> x) {
  const B q = x.v * x.v + y.v * y.v;
  return {atan2_(y.v, x.v), (x.v * y.d - y.v * x.d) / q};
}

// f(x) and, when JAC, its Jacobian J (NOUT x NIN).
template <typename R, int NIN, int NOUT, bool JAC, typename S>
PSM_DEV void model_eval(const ModelView<R, S>& md, const R (&x)[NIN], R (&y)[NOUT], R (&J)[NOUT][NIN]) {
  switch (md.kind) {
    case PSM_MODEL_TABULATED:
    case PSM_MODEL_CONDITIONAL: {
      // values evaluated by the caller: per row either f at each sigma point
      // (npts x NOUT) or f(m) followed by the Jacobian (NOUT + NOUT x NIN);
      // a conditional-moments model adds its conditional factor (NOUT x NOUT)
      // after each value block
      const int V = md.kind == PSM_MODEL_CONDITIONAL ? NOUT * NOUT : 0;
      if (md.npts > 0) {
        const int64_t base = (md.row * md.npts + md.pt) * (NOUT + V);
#pragma unroll
        for (int i = 0; i < NOUT; ++i) y[i] = md.par64(base + i);
      } else {
        const int64_t base = md.row * (NOUT + NOUT * NIN + V);
#pragma unroll
        for (int i = 0; i < NOUT; ++i) {
          y[i] = md.par64(base + i);
          if (JAC) {
#pragma unroll
            for (int j = 0; j < NIN; ++j) J[i][j] = md.par64(base + NOUT + i * NIN + j);
          }
        }
      }
      break;
    }
    case PSM_MODEL_LINEAR: {
      // x -> A x, A row-major NOUT x NIN
#pragma unroll
      for (int i = 0; i < NOUT; ++i) {
        R s = R(0);
#pragma unroll
        for (int j = 0; j < NIN; ++j) {
          const R a = md.par(i * NIN + j);
          s = fma(a, x[j], s);
          if (JAC) J[i][j] = a;
        }
        y[i] = s;
      }
      break;
    }
    case PSM_MODEL_COORDINATED_TURN: {
      if constexpr (NIN == 5 && NOUT == 5) {
        // tests/bearings/bearings_utils.py:7-44: branch at |w| < 1e-6 where
        // sin(wt)/w -> dt and (cos(wt)-1)/w -> 0 with zero w-derivative.
        const R dt = md.par(0);
        const R w = x[4];
        const R c = cos_(w * dt);
        const R s = sin_(w * dt);
        const bool small = abs_(w) < R(1e-6);
        const R a = small ? dt : s / w;
        const R bb = small ? R(0) : (c - R(1)) / w;
        y[0] = x[0] + a * x[2] - bb * x[3];
        y[1] = x[1] + bb * x[2] + a * x[3];
        y[2] = c * x[2] + s * x[3];
        y[3] = -s * x[2] + c * x[3];
        y[4] = x[4];
        if (JAC) {
          const R da = small ? R(0) : (dt * c - a) / w;
          const R db = small ? R(0) : (-dt * s - bb) / w;
          const R dc = -dt * s, ds = dt * c;
#pragma unroll
          for (int i = 0; i < 5; ++i)
#pragma unroll
            for (int j = 0; j < 5; ++j) J[i][j] = R(0);
          J[0][0] = R(1); J[0][2] = a;  J[0][3] = -bb; J[0][4] = da * x[2] - db * x[3];
          J[1][1] = R(1); J[1][2] = bb; J[1][3] = a;   J[1][4] = db * x[2] + da * x[3];
          J[2][2] = c;    J[2][3] = s;  J[2][4] = dc * x[2] + ds * x[3];
          J[3][2] = -s;   J[3][3] = c;  J[3][4] = -ds * x[2] + dc * x[3];
          J[4][4] = R(1);
        }
      }
      break;
    }
    case PSM_MODEL_BEARINGS: {
      if constexpr (NOUT == 2 && NIN >= 2) {
        // tests/bearings/bearings_utils.py:47-69 (no angle wrapping)
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const R dx = x[0] - md.par(2 * r);
          const R dy = x[1] - md.par(2 * r + 1);
          y[r] = atan2_(dy, dx);
          if (JAC) {
            const R q = dx * dx + dy * dy;
#pragma unroll
            for (int j = 0; j < NIN; ++j) J[r][j] = R(0);
            J[r][0] = -dy / q;
            J[r][1] = dx / q;
          }
        }
      }
      break;
    }
    default: {
      // elementwise sin / cos / exp / arctan (NIN == NOUT)
      if constexpr (NIN == NOUT) {
#pragma unroll
        for (int i = 0; i < NOUT; ++i) {
          R v, d;
          const R xi = x[i];
          if (md.kind == PSM_MODEL_SIN) { v = sin_(xi); d = cos_(xi); }
          else if (md.kind == PSM_MODEL_COS) { v = cos_(xi); d = -sin_(xi); }
          else if (md.kind == PSM_MODEL_EXP) { v = exp_(xi); d = v; }
          else { v = atan_(xi); d = R(1) / (R(1) + xi * xi); }
          y[i] = v;
          if (JAC) {
#pragma unroll
            for (int j = 0; j < NIN; ++j) J[i][j] = (i == j) ? d : R(0);
          }
        }
      }
      break;
    }
  }
}

// Conditional factor (chol in square-root form, covariance otherwise) of a
// PSM_MODEL_CONDITIONAL row: at the sigma point md.pt, or at the mean.
template <typename R, int NIN, int NOUT, typename S>
PSM_DEV void model_cond_factor(const ModelView<R, S>& md, R (&V)[NOUT][NOUT]) {
  constexpr int W = NOUT * NOUT;
  const int64_t base = md.npts > 0 ? (md.row * md.npts + md.pt) * (NOUT + W) + NOUT
                                   : md.row * (NOUT + NOUT * NIN + W) + NOUT + NOUT * NIN;
#pragma unroll
  for (int i = 0; i < NOUT; ++i)
#pragma unroll
    for (int j = 0; j < NOUT; ++j) V[i][j] = md.par64(base + i * NOUT + j);
}

template <typename R, int NIN, int NOUT, typename S>
PSM_DEV void model_value(const ModelView<R, S>& md, const R (&x)[NIN], R (&y)[NOUT]) {
  R J[NOUT][NIN];
  model_eval<R, NIN, NOUT, false>(md, x, y, J);
}

// Sigma points whose regression columns join the square-root factor in one
// triangularisation.
constexpr int kSigmaBatch = 4;

// Sigma-point table (unit points xi_p, weights) in device memory, stored in
// FP64 for every compute type: the statistical linear regression subtracts
// F P F^T from the sigma-point covariance, and weights rounded to FP32 (1/10
// is 0.1000000015) no longer sum to one, which leaves a residual of ~1e-8 of
// the covariance -- more than a small process noise (3e-9 on the
// bearings-only model) -- in the regression noise factor.
template <typename R, typename S = double>
struct SigmaView {
  int npts;
  const S* xi;  // npts x n
  const S* wm;
  const S* wc;
  PSM_DEV R pxi(int i) const { return R(xi[i]); }
  PSM_DEV R pwm(int i) const { return R(wm[i]); }
  PSM_DEV R pwc(int i) const { return R(wc[i]); }
};

// ----------------------------------------------------------------------------
// Extended (first-order Taylor) linearization, parsmooth/linearization/_extended.py:59-70:
// F = df/dx(m), chol/cov passed through, b = f(m) - F m + m_q.  The nominal
// covariance is not used.
// ----------------------------------------------------------------------------
template <typename R, int NIN, int NOUT, typename S>
PSM_DEV void linearize_extended(const ModelView<R, S>& md, const R (&m)[NIN], R (&F)[NOUT][NIN],
                                R (&C)[NOUT][NOUT], R (&b)[NOUT]) {
  R y[NOUT];
  model_eval<R, NIN, NOUT, true>(md, m, y, F);
#pragma unroll
  for (int i = 0; i < NOUT; ++i) {
    R s = y[i];
#pragma unroll
    for (int j = 0; j < NIN; ++j) s = fma(-F[i][j], m[j], s);
    b[i] = s + md.qm(i);
#pragma unroll
    for (int j = 0; j < NOUT; ++j) C[i][j] = md.qc(i * NOUT + j);
  }
  // conditional moments (parsmooth/linearization/_extended.py:43-56): the
  // conditional factor at the mean replaces the noise
  if (md.kind == PSM_MODEL_CONDITIONAL) model_cond_factor<R, NIN, NOUT>(md, C);
}

// ----------------------------------------------------------------------------
// Sigma-point statistical linear regression, parsmooth/linearization/_sigma_points.py:63-100.
// Two passes over the points (the function is re-evaluated instead of storing
// npts outputs): pass 1 forms the weighted mean, pass 2 the cross-covariance
// and, in square-root form, streams sqrt(wc_p)(f_p - m_f) into a running
// triangular factor that starts from the noise factor, which equals
// tria([tria(sqrt_Phi) | chol_q]) of the reference.  The factor is then
// downdated by the columns of F L (rank-one downdates, multiplier -1).
// SQRT=false returns Phi - F P F^T + Q symmetrised (:80-85).
// ----------------------------------------------------------------------------
template <typename R, int NIN, int NOUT, bool SQRT, typename S>
PSM_DEV void linearize_sigma(const ModelView<R, S>& md, const SigmaView<R>& sp, const R (&m)[NIN],
                             const R (&L)[NIN][NIN], R (&F)[NOUT][NIN], R (&C)[NOUT][NOUT],
                             R (&b)[NOUT]) {
  R mf[NOUT];
  set_zero(mf);
  for (int p = 0; p < sp.npts; ++p) {
    R chi[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      R s = m[i];
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
      chi[i] = s;
    }
    R f[NOUT];
    ModelView<R, S> mp = md;
    mp.pt = p;
    model_value<R, NIN, NOUT>(mp, chi, f);
    const R w = sp.pwm(p);
#pragma unroll
    for (int i = 0; i < NOUT; ++i) mf[i] = fma(w, f[i], mf[i]);
  }
  R Psi[NIN][NOUT];
  set_zero(Psi);
  R Acc[NOUT][NOUT + 1];   // running factor | new column
  R Phi[NOUT][NOUT];
  if constexpr (SQRT) {
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) Acc[i][j] = md.qc(i * NOUT + j);
    {
      R sq[NOUT][NOUT];
#pragma unroll
      for (int i = 0; i < NOUT; ++i)
#pragma unroll
        for (int j = 0; j < NOUT; ++j) sq[i][j] = Acc[i][j];
      tria_inplace<R, NOUT, NOUT>(sq);
#pragma unroll
      for (int i = 0; i < NOUT; ++i)
#pragma unroll
        for (int j = 0; j < NOUT; ++j) Acc[i][j] = sq[i][j];
    }
  } else {
    set_zero(Phi);
  }
  // sigma point p: deviation d = f(chi_p) - mf, weight wc_p; accumulates Psi
  auto point = [&](int p, R (&d)[NOUT], R& w) {
    R chi[NIN], dx[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      R s = m[i];
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
      chi[i] = s;
      dx[i] = s - m[i];
    }
    R f[NOUT];
    ModelView<R, S> mp = md;
    mp.pt = p;
    model_value<R, NIN, NOUT>(mp, chi, f);
    w = sp.pwc(p);
#pragma unroll
    for (int i = 0; i < NOUT; ++i) d[i] = f[i] - mf[i];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      const R wx = w * dx[i];
#pragma unroll
      for (int j = 0; j < NOUT; ++j) Psi[i][j] = fma(wx, d[j], Psi[i][j]);
    }
  };
  bool batched = false;
  if constexpr (SQRT) {
    if (md.kind != PSM_MODEL_CONDITIONAL) {
      // the columns sqrt(wc_p) d_p join the factor kSigmaBatch at a time: one
      // triangularisation of [factor | columns] per batch instead of one per
      // point (the reference triangularises all columns at once); a batch
      // past the last point is padded with zero columns
      batched = true;
      for (int p0 = 0; p0 < sp.npts; p0 += kSigmaBatch) {
        R Ab[NOUT][NOUT + kSigmaBatch];
#pragma unroll
        for (int i = 0; i < NOUT; ++i)
#pragma unroll
          for (int j = 0; j < NOUT; ++j) Ab[i][j] = Acc[i][j];
#pragma unroll
        for (int k = 0; k < kSigmaBatch; ++k) {
          R d[NOUT], w = R(0);
          set_zero(d);
          const bool live = p0 + k < sp.npts;
          if (live) point(p0 + k, d, w);
          const R sw = live ? sqrt_(w) : R(0);
#pragma unroll
          for (int i = 0; i < NOUT; ++i) Ab[i][NOUT + k] = sw * d[i];
        }
        tria_inplace<R, NOUT, NOUT + kSigmaBatch>(Ab);
#pragma unroll
        for (int i = 0; i < NOUT; ++i)
#pragma unroll
          for (int j = 0; j < NOUT; ++j) Acc[i][j] = Ab[i][j];
      }
    }
  }
  for (int p = 0; p < sp.npts && !batched; ++p) {
    R d[NOUT], w;
    point(p, d, w);
    ModelView<R, S> mp = md;
    mp.pt = p;
    if constexpr (SQRT) {
      const R sw = sqrt_(w);
#pragma unroll
      for (int i = 0; i < NOUT; ++i) Acc[i][NOUT] = sw * d[i];
      tria_inplace<R, NOUT, NOUT + 1>(Acc);
      if (md.kind == PSM_MODEL_CONDITIONAL) {
        // conditional moments (parsmooth/linearization/_sigma_points.py:25-50):
        // the columns sqrt(wc_p) chol(x_p) join the factor
        R V[NOUT][NOUT], Ac[NOUT][2 * NOUT];
        model_cond_factor<R, NIN, NOUT>(mp, V);
#pragma unroll
        for (int i = 0; i < NOUT; ++i)
#pragma unroll
          for (int j = 0; j < NOUT; ++j) {
            Ac[i][j] = Acc[i][j];
            Ac[i][NOUT + j] = sw * V[i][j];
          }
        tria_inplace<R, NOUT, 2 * NOUT>(Ac);
#pragma unroll
        for (int i = 0; i < NOUT; ++i)
#pragma unroll
          for (int j = 0; j < NOUT; ++j) Acc[i][j] = Ac[i][j];
      }
    } else {
#pragma unroll
      for (int i = 0; i < NOUT; ++i)
#pragma unroll
        for (int j = 0; j < NOUT; ++j) Phi[i][j] = fma(w * d[i], d[j], Phi[i][j]);
      if (md.kind == PSM_MODEL_CONDITIONAL) {
        // (parsmooth/linearization/_sigma_points.py:52-60): + sum_p wc_p V(x_p)
        R V[NOUT][NOUT];
        model_cond_factor<R, NIN, NOUT>(mp, V);
#pragma unroll
        for (int i = 0; i < NOUT; ++i)
#pragma unroll
          for (int j = 0; j < NOUT; ++j) Phi[i][j] = fma(w, V[i][j], Phi[i][j]);
      }
    }
  }
  // F = (L^-T L^-1 Psi)^T
  R X[NIN][NOUT], Y[NIN][NOUT];
  solve_lower<R, NIN, NOUT>(L, Psi, X);
  solve_lower_T<R, NIN, NOUT>(L, X, Y);
#pragma unroll
  for (int i = 0; i < NOUT; ++i)
#pragma unroll
    for (int j = 0; j < NIN; ++j) F[i][j] = Y[j][i];
#pragma unroll
  for (int i = 0; i < NOUT; ++i) {
    R s = mf[i];
#pragma unroll
    for (int j = 0; j < NIN; ++j) s = fma(-F[i][j], m[j], s);
    b[i] = s + md.qm(i);
  }
  R FL[NOUT][NIN];
  mm<R, NOUT, NIN, NIN>(F, L, FL);
  if constexpr (SQRT) {
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) C[i][j] = Acc[i][j];
#pragma unroll
    for (int k = 0; k < NIN; ++k) {
      R v[NOUT];
#pragma unroll
      for (int i = 0; i < NOUT; ++i) v[i] = FL[i][k];
      cholesky_update<R, NOUT>(C, v, R(-1));
    }
  } else {
    // Phi - (F L)(F L)^T + Q, symmetrised
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) {
        R s = Phi[i][j] + md.qc(i * NOUT + j);
#pragma unroll
        for (int k = 0; k < NIN; ++k) s = fma(-FL[i][k], FL[j][k], s);
        C[i][j] = s;
      }
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < i; ++j) {
        const R s = R(0.5) * (C[i][j] + C[j][i]);
        C[i][j] = s;
        C[j][i] = s;
      }
  }
}

}  // namespace psm
