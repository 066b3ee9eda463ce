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

// Single-precision arithmetic types (FP32 values, FP32 tangents): their
// linearizations use the cancellation-free formulas below (the coordinated
// turn's terms are also used by every tangent type).
template <typename R> struct IsSingle : std::false_type {};
template <> struct IsSingle<float> : std::true_type {};
template <> struct IsSingle<DualT<float>> : std::true_type {};

// The turn-rate terms of the coordinated turn, x = w dt:
//   a = sin(x)/w = dt sinc(x),        bb = (cos(x) - 1)/w = -dt g(x),
//   da/dw = dt^2 sinc'(x),            db/dw = -dt^2 g'(x),
// with sinc(x) = sin(x)/x and g(x) = (1 - cos x)/x.  The textbook quotients
// cancel in FP32 at small x (dt cos(x) - sin(x)/w is a difference of two
// numbers ~dt that agree to O(x^2)), so |x| < 0.5 uses the Taylor series of
// sinc, sinc', g, g' (through x^9: truncation < 3e-11) and larger |x| the closed
// forms with 1 - cos x = 2 sin^2(x/2).  |w| < 1e-6 keeps the reference's branch
// (tests/bearings/bearings_utils.py:30-37): a = dt, bb = 0, zero w-derivatives.
template <typename R>
PSM_DEV void ct_terms_single(R w, R dt, R& a, R& bb, R& da, R& db, R& c, R& s) {
  const R x = w * dt;
  c = cos_(x);
  s = sin_(x);
  if (abs_(w) < R(1e-6)) {
    a = dt;
    bb = R(0);
    da = R(0);
    db = R(0);
    return;
  }
  if (abs_(x) < R(0.5)) {
    const R x2 = x * x;
    const R sinc = R(1) + x2 * (R(-1.0 / 6) + x2 * (R(1.0 / 120) + x2 * (R(-1.0 / 5040) + x2 * R(1.0 / 362880))));
    const R dsinc = x * (R(-1.0 / 3) + x2 * (R(1.0 / 30) + x2 * (R(-1.0 / 840) +
                                                               x2 * (R(1.0 / 45360) + x2 * R(-1.0 / 3991680)))));
    const R g = x * (R(0.5) + x2 * (R(-1.0 / 24) + x2 * (R(1.0 / 720) + x2 * (R(-1.0 / 40320) +
                                                                             x2 * R(1.0 / 3628800)))));
    const R dg = R(0.5) + x2 * (R(-1.0 / 8) + x2 * (R(1.0 / 144) + x2 * (R(-1.0 / 5760) + x2 * R(1.0 / 403200))));
    a = dt * sinc;
    bb = -dt * g;
    da = dt * dt * dsinc;
    db = -dt * dt * dg;
  } else {
    const R sh = sin_(R(0.5) * x);
    const R omc = R(2) * sh * sh;   // 1 - cos x
    const R w2 = w * w;
    a = s / w;
    bb = -omc / w;
    da = (x * c - s) / w2;
    db = -(x * s - omc) / w2;
  }
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
        R c, s, a, bb, da, db;
        if constexpr (IsSingle<R>::value || is_dual<R>::value) {
          // FP32, and every tangent type: the tangents of the textbook
          // quotients (second derivatives) cancel even in FP64
          ct_terms_single(w, dt, a, bb, da, db, c, s);
        } else {
          c = cos_(w * dt);
          s = sin_(w * dt);
          const bool small = abs_(w) < R(1e-6);
          a = small ? dt : s / w;
          bb = small ? R(0) : (c - R(1)) / w;
          da = small ? R(0) : (dt * c - a) / w;
          db = small ? R(0) : (-dt * s - bb) / w;
        }
        y[0] = x[0] + a * x[2] - bb * x[3];
        y[1] = x[1] + bb * x[2] + a * x[3];
        y[2] = c * x[2] + s * x[3];
        y[3] = -s * x[2] + c * x[3];
        y[4] = x[4];
        if (JAC) {
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
  if constexpr (IsSingle<R>::value) {
    // f(m) - F m in closed form where the model gives one (in FP32 the
    // difference above cancels at large |m|): 0 for x -> A x; for the
    // coordinated turn, f is linear in the position and velocity given w, so
    // f(m) - F m = -F[:, 4] w (and 0 in the w row)
    if (md.kind == PSM_MODEL_LINEAR) {
#pragma unroll
      for (int i = 0; i < NOUT; ++i) b[i] = md.qm(i);
    } else if (md.kind == PSM_MODEL_COORDINATED_TURN) {
      if constexpr (NIN == 5 && NOUT == 5) {
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = md.qm(i) - F[i][4] * m[4];
        b[4] = md.qm(4);
      }
    }
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

// ----------------------------------------------------------------------------
// Increment f(m + d) - f(m) of a model and its value at the mean, for the
// single-precision regression below.  For x -> A x and the coordinated turn
// (linear in position and velocity given w) the increment is formed without
// m + d: A d, and F(w') d + (F(w') - F(w)) m with w' = w + d_w, whose terms
// are all of the size of d or of the velocity times a difference of turn-rate
// terms.  In FP32 a large position (|m| ~ 1e3) rounds m + d and f(m + d) by
// ~1e-4, above the process noise's chol Q (~6e-5 on the bearings-only model),
// so forming f(m + d) - f(m) from the two values would leave that rounding in
// the regression's noise factor.  Other models: f(m + d) - f(m) as
// evaluated; a tabulated model's values are relative to its first point.
// ----------------------------------------------------------------------------
template <typename R, int NIN, int NOUT, typename S>
PSM_DEV void model_at_mean(const ModelView<R, S>& md, const R (&m)[NIN], R (&f0)[NOUT]) {
  ModelView<R, S> m0 = md;
  m0.pt = 0;   // tabulated rows: the first point is the reference
  model_value<R, NIN, NOUT>(m0, m, f0);
}

template <typename R, int NIN, int NOUT, typename S>
PSM_DEV void model_increment(const ModelView<R, S>& md, const R (&m)[NIN], const R (&d)[NIN],
                             const R (&f0)[NOUT], R (&df)[NOUT]) {
  if (md.kind == PSM_MODEL_LINEAR) {
#pragma unroll
    for (int i = 0; i < NOUT; ++i) {
      R acc = R(0);
#pragma unroll
      for (int j = 0; j < NIN; ++j) acc = fma(md.par(i * NIN + j), d[j], acc);
      df[i] = acc;
    }
    return;
  }
  if constexpr (NIN == 5 && NOUT == 5) {
    if (md.kind == PSM_MODEL_COORDINATED_TURN) {
      const R dt = md.par(0);
      R a0, b0, c0, s0, a1, b1, c1, s1, u, v;
      ct_terms_single(m[4], dt, a0, b0, u, v, c0, s0);
      ct_terms_single(m[4] + d[4], dt, a1, b1, u, v, c1, s1);
      const R da = a1 - a0, db = b1 - b0, dc = c1 - c0, ds = s1 - s0;
      df[0] = d[0] + a1 * d[2] - b1 * d[3] + (da * m[2] - db * m[3]);
      df[1] = d[1] + b1 * d[2] + a1 * d[3] + (db * m[2] + da * m[3]);
      df[2] = c1 * d[2] + s1 * d[3] + (dc * m[2] + ds * m[3]);
      df[3] = -s1 * d[2] + c1 * d[3] + (-ds * m[2] + dc * m[3]);
      df[4] = d[4];
      return;
    }
  }
  R x[NIN], f[NOUT];
#pragma unroll
  for (int i = 0; i < NIN; ++i) x[i] = m[i] + d[i];
  model_value<R, NIN, NOUT>(md, x, f);
#pragma unroll
  for (int i = 0; i < NOUT; ++i) df[i] = f[i] - f0[i];
}

// ----------------------------------------------------------------------------
// Statistical linear regression from centred residuals (single precision).
// Same F and b as linearize_sigma; the noise factor from the residuals
// r_p = (f_p - m_f) - F (chi_p - m) of the regression: for a rule whose
// weights wc are non-negative and whose points reproduce the mean and
// covariance of x (sum wc (chi - m) = 0, sum wc (chi - m)(chi - m)^T = L L^T:
// cubature, Gauss-Hermite, unscented), sum wc r r^T = Phi - F P F^T, so
//   C = tria([sqrt(wc_p) r_p ... | chol Q])            (square-root form),
//   C = sum wc r r^T + Q, symmetrised                   (covariance form),
// equal to the reference's chol(Phi + Q - F P F^T) and Phi - F P F^T + Q
// (parsmooth/linearization/_sigma_points.py:63-85) without forming the
// difference Phi - F P F^T, which cancels when Q is small against the
// sigma-point spread.  The points enter as displacements d_p = L xi_p and
// increments f(m + d_p) - f(m) (model_increment), so neither the difference
// Phi - F P F^T nor the rounding of a large mean reaches the noise factor.
// Two passes over the points: the mean and the cross-covariance, then the
// residual columns, which read the increments back from the caller's per-lane
// scratch (pc) when it holds them, else evaluate them again.
// Returns false, with nothing written, when some wc is negative.
// ----------------------------------------------------------------------------
// Per-lane scratch for the increments of one regression (value i of point p
// at base[(p * NOUT + i) * stride]), used when there are at most cap points.
template <typename R> struct PointCache {
  R* base;
  int stride;
  int cap;
};

template <typename R, int NIN, int NOUT, bool SQRT, typename S>
PSM_DEV bool linearize_sigma_centred(const ModelView<R, S>& md, const SigmaView<R>& sp, const R (&m)[NIN],
                                     const R (&L)[NIN][NIN], R (&F)[NOUT][NIN], R (&C)[NOUT][NOUT],
                                     R (&b)[NOUT], const PointCache<R>& pc = PointCache<R>{nullptr, 0, 0}) {
  for (int p = 0; p < sp.npts; ++p)
    if (sp.wc[p] < 0.0) return false;
  R f0[NOUT];
  model_at_mean<R, NIN, NOUT>(md, m, f0);
  // point p: displacement d = L xi_p and increment f = f(m + d) - f(m)
  auto point = [&](int p, R (&d)[NIN], R (&f)[NOUT]) {
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      R s = R(0);
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
      d[i] = s;
    }
    ModelView<R, S> mp = md;
    mp.pt = p;
    model_increment<R, NIN, NOUT>(mp, m, d, f0, f);
  };
  // one pass for the weighted mean of the increments and their
  // cross-covariance with the displacements:
  //   Psi = sum wc d (f - mf)^T = sum wc d f^T - (sum wc d) mf^T
  // (sum wc d vanishes for symmetric rules; its term is kept for the others)
  R mf[NOUT], Psi[NIN][NOUT], sd[NIN];
  set_zero(mf);
  set_zero(Psi);
  set_zero(sd);
  const bool keep = pc.base != nullptr && sp.npts <= pc.cap;
  for (int p = 0; p < sp.npts; ++p) {
    R d[NIN], f[NOUT];
    point(p, d, f);
    if (keep) {
#pragma unroll
      for (int i = 0; i < NOUT; ++i) pc.base[(p * NOUT + i) * pc.stride] = f[i];
    }
    const R wm = sp.pwm(p), wc = sp.pwc(p);
#pragma unroll
    for (int i = 0; i < NOUT; ++i) mf[i] = fma(wm, f[i], mf[i]);
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      const R wx = wc * d[i];
      sd[i] += wx;
#pragma unroll
      for (int j = 0; j < NOUT; ++j) Psi[i][j] = fma(wx, f[j], Psi[i][j]);
    }
  }
#pragma unroll
  for (int i = 0; i < NIN; ++i)
#pragma unroll
    for (int j = 0; j < NOUT; ++j) Psi[i][j] = fma(-sd[i], mf[j], Psi[i][j]);
  // F = (L^-T L^-1 Psi)^T
  {
    R X[NIN][NOUT], Y[NIN][NOUT];
    solve_lower<R, NIN, NOUT>(L, Psi, X);
    solve_lower_T<R, NIN, NOUT>(L, X, Y);
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NIN; ++j) F[i][j] = Y[j][i];
  }
#pragma unroll
  for (int i = 0; i < NOUT; ++i) {
    R s = f0[i];
#pragma unroll
    for (int j = 0; j < NIN; ++j) s = fma(-F[i][j], m[j], s);
    b[i] = (s + mf[i]) + md.qm(i);
  }
  // residual of point p
  auto residual = [&](int p, R (&r)[NOUT]) {
    R d[NIN], f[NOUT];
    if (keep) {
#pragma unroll
      for (int i = 0; i < NIN; ++i) {
        R s = R(0);
#pragma unroll
        for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
        d[i] = s;
      }
#pragma unroll
      for (int i = 0; i < NOUT; ++i) f[i] = pc.base[(p * NOUT + i) * pc.stride];
    } else {
      point(p, d, f);
    }
#pragma unroll
    for (int i = 0; i < NOUT; ++i) {
      R s = f[i] - mf[i];
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(-F[i][j], d[j], s);
      r[i] = s;
    }
  };
  if constexpr (SQRT) {
    R Acc[NOUT][NOUT];
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) Acc[i][j] = md.qc(i * NOUT + j);
    tria_inplace<R, NOUT, NOUT>(Acc);
    for (int p0 = 0; p0 < sp.npts; p0 += kSigmaBatch) {
      R Ab[NOUT][NOUT + kSigmaBatch];
#pragma unroll
      for (int i = 0; i < NOUT; ++i)
#pragma unroll
        for (int j = 0; j < NOUT; ++j) Ab[i][j] = Acc[i][j];
#pragma unroll
      for (int k = 0; k < kSigmaBatch; ++k) {
        R r[NOUT];
        set_zero(r);
        const bool live = p0 + k < sp.npts;
        if (live) residual(p0 + k, r);
        const R sw = live ? sqrt_(sp.pwc(p0 + k)) : R(0);
#pragma unroll
        for (int i = 0; i < NOUT; ++i) Ab[i][NOUT + k] = sw * r[i];
      }
      tria_inplace<R, NOUT, NOUT + kSigmaBatch>(Ab);
#pragma unroll
      for (int i = 0; i < NOUT; ++i)
#pragma unroll
        for (int j = 0; j < NOUT; ++j) Acc[i][j] = Ab[i][j];
    }
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) C[i][j] = Acc[i][j];
  } else {
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) C[i][j] = md.qc(i * NOUT + j);
    for (int p = 0; p < sp.npts; ++p) {
      R r[NOUT];
      residual(p, r);
      const R w = sp.pwc(p);
#pragma unroll
      for (int i = 0; i < NOUT; ++i)
#pragma unroll
        for (int j = 0; j < NOUT; ++j) C[i][j] = fma(w * r[i], r[j], C[i][j]);
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
  return true;
}

}  // namespace psm
