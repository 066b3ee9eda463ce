// Reverse-mode adjoint (vector-Jacobian products) of one filter + smoother
// pass and of its log-likelihood, with respect to the nominal trajectory the
// pass linearizes at and to the model parameters.  These are the two VJPs the
// implicit gradient of iterated smoothing needs (parsmooth/_utils.py:118-133:
// x_bar = d ell / d x*, u <- x_bar + J_x^T u, theta_bar = (d f / d theta)^T u +
// d ell / d theta): one adjoint pass evaluates J_x^T u and (d f / d theta)^T u
// together, whatever the number of parameters.
//
// The pass is a function of the affine-Gaussian model it linearizes to,
//   x_k = F_k x_{k-1} + b_k + N(0, Q_k),   y_k = H_k x_k + c_k + N(0, R_k),
//   x_0 ~ N(m0, P0),
// whose smoothed means solve the block-tridiagonal system Lambda m = eta
// (posterior precision Lambda, information vector eta) and whose smoothed
// covariances are the diagonal blocks of Sigma = Lambda^{-1}.  For output
// cotangents (m_bar, P_bar) (P_bar from the chol factors' cotangents):
//   d/d eta = lambda = Lambda^{-1} m_bar,
//   d/d Lambda = -lambda m^T - Sigma P_bar Sigma   (block-tridiagonal part),
// and the chain rule through Lambda and eta gives the cotangents of
// (F, b, Q, H, c, R, m0, P0).  lambda comes from the filter / smoother mean
// recursions driven by m_bar instead of the data:
//   phi_k = Abar_k phi_{k-1} + Pf_k m_bar_k      (Abar = (I - K H) F, forward),
//   lambda_k = G_k lambda_{k+1} + (I - G_k F_{k+1}) phi_k   (backward),
// and the tridiagonal blocks of Sigma P_bar Sigma from two congruence
// recursions with the smoother gains G:
//   B_k = G_k B_{k+1} G_k^T + Ps_k P_bar_k Ps_k   (backward),
//   Psi_k = G_{k-1}^T Psi_{k-1} G_{k-1} + P_bar_k  (forward),
//   D_kk = B_k + Ps_k (Psi_k - P_bar_k) Ps_k,  D_k,k+1 = G_k B_{k+1} + Ps_k Psi_k G_k Ps_{k+1}.
// Each recursion is an affine map per state, composed by a chunked scan
// (lane per chunk, recursive levels) like the forward passes.  The
// log-likelihood's gradient with respect to the same model is Fisher's
// identity E[d log p(x, y) / d theta | y] over the smoothed distribution.
// Last, the cotangents of (F, b, Q) / (H, c, R) at each state go through the
// linearization at that state's nominal (extended: the Jacobian's derivative
// by forward tangents; sigma points: the statistical linear regression in
// reverse) to the nominal mean and chol and to the model parameters.
//
// All adjoint arithmetic is FP64 (the primal checkpoint may be FP32): the
// chain through Lambda uses Q^{-1} and R^{-1}, which on the bearings-only
// model (Q ~ 3e-9 .. 1e-3) FP32 cannot carry.  Square-root form; models the
// kernels evaluate (linear, coordinated turn, bearings, elementwise); the
// extended method or sigma-point rules with non-negative weights.
#pragma once
#include "psm_kernels.hpp"

namespace psm {
namespace adj {

using D = double;

// ---------------------------------------------------------------- small FP64 helpers
template <int N>
PSM_DEV void gram(const D (&L)[N][N], D (&P)[N][N]) {   // P = L L^T
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      D s = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) s = fma(L[i][k], L[j][k], s);
      P[i][j] = s;
    }
}

template <int N>
PSM_DEV void symmetrize(D (&A)[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) {
      const D s = 0.5 * (A[i][j] + A[j][i]);
      A[i][j] = s;
      A[j][i] = s;
    }
}

// Cotangent of P = L L^T for the cotangent Lb of its lower Cholesky factor L:
// sym(L^-T Phi(L^T Lb) L^-1), Phi = lower triangle with the diagonal halved.
template <int N>
PSM_DEV void chol_vjp(const D (&L)[N][N], const D (&Lb)[N][N], D (&Pb)[N][N]) {
  D S[N][N], Y[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      D s = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) s = fma(L[k][i], (k >= j ? Lb[k][j] : 0.0), s);
      S[i][j] = j > i ? 0.0 : (j == i ? 0.5 * s : s);
    }
  solve_right_lower<D, N, N>(L, S, Y);      // Y = S L^-1
  solve_lower_T<D, N, N>(L, Y, Pb);         // Pb = L^-T Y
  symmetrize(Pb);
}

// X = L^-1 S L^-T for lower L (S symmetric)
template <int N>
PSM_DEV void whiten(const D (&L)[N][N], const D (&S)[N][N], D (&X)[N][N]) {
  D Y[N][N];
  solve_lower<D, N, N>(L, S, Y);
  solve_right_lower_T<D, N, N>(L, Y, X);
}

// cotangent of a chol factor C from the cotangent Qb of Q = C C^T: 2 Qb C
template <int N>
PSM_DEV void gram_vjp(const D (&Qb)[N][N], const D (&C)[N][N], D (&Cb)[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      D s = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) s = fma(Qb[i][k], C[k][j], s);
      Cb[i][j] = j > i ? 0.0 : 2.0 * s;
    }
}

// AoS loads of the primal's (or the caller's) arrays into FP64 registers
template <typename S, int N>
PSM_DEV void ld_cv(const S* __restrict__ p, int64_t i, D (&v)[N]) {
#pragma unroll
  for (int a = 0; a < N; ++a) v[a] = (D)p[i * N + a];
}
template <typename S, int N1, int N2>
PSM_DEV void ld_cv(const S* __restrict__ p, int64_t i, D (&v)[N1][N2]) {
#pragma unroll
  for (int a = 0; a < N1; ++a)
#pragma unroll
    for (int c = 0; c < N2; ++c) v[a][c] = (D)p[i * N1 * N2 + a * N2 + c];
}

// ---------------------------------------------------------------- per-state buffers
// FP64 component-major arrays over the states, positions interleaved by
// StateMap{C} (C = kChunk) so that lane-per-chunk scans read consecutive words.
constexpr int kChunk = 32;

template <int N>
PSM_DEV void ldv(const D* __restrict__ b, int64_t st, int64_t p, int q0, D (&v)[N]) {
#pragma unroll
  for (int a = 0; a < N; ++a) v[a] = b[(int64_t)(q0 + a) * st + p];
}
template <int N>
PSM_DEV void ldm(const D* __restrict__ b, int64_t st, int64_t p, int q0, D (&v)[N][N]) {
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int c = 0; c < N; ++c) v[a][c] = b[(int64_t)(q0 + a * N + c) * st + p];
}
template <int N1, int N2>
PSM_DEV void ldm2(const D* __restrict__ b, int64_t st, int64_t p, int q0, D (&v)[N1][N2]) {
#pragma unroll
  for (int a = 0; a < N1; ++a)
#pragma unroll
    for (int c = 0; c < N2; ++c) v[a][c] = b[(int64_t)(q0 + a * N2 + c) * st + p];
}
template <int N1, int N2>
PSM_DEV void stm2(D* __restrict__ b, int64_t st, int64_t p, int q0, const D (&v)[N1][N2]) {
#pragma unroll
  for (int a = 0; a < N1; ++a)
#pragma unroll
    for (int c = 0; c < N2; ++c) b[(int64_t)(q0 + a * N2 + c) * st + p] = v[a][c];
}
template <int N>
PSM_DEV void stv(D* __restrict__ b, int64_t st, int64_t p, int q0, const D (&v)[N]) {
#pragma unroll
  for (int a = 0; a < N; ++a) b[(int64_t)(q0 + a) * st + p] = v[a];
}
template <int N>
PSM_DEV void stm(D* __restrict__ b, int64_t st, int64_t p, int q0, const D (&v)[N][N]) {
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int c = 0; c < N; ++c) b[(int64_t)(q0 + a * N + c) * st + p] = v[a][c];
}

// Parameter cotangents per state: [trans params (NX*NX) | trans q_mean (NX) |
// trans q_chol (NX*NX) | obs params (NY*NX) | obs q_mean (NY) | obs q_chol
// (NY*NY) | m0 (NX) | L0 (NX*NX)]; model kinds use a prefix of their params slot.
template <int NX, int NY> struct Th {
  static constexpr int TP = 0, TQM = NX * NX, TQC = NX * NX + NX, OP = 2 * NX * NX + NX,
                       OQM = 2 * NX * NX + NX + NY * NX, OQC = OQM + NY, M0 = OQC + NY * NY,
                       L0 = M0 + NX, N = L0 + NX * NX;
};

PSM_DEV int64_t pos_of(int64_t i) { return StateMap{kChunk}(i); }
// state of interleaved position t (inverse of StateMap{kChunk})
PSM_DEV int64_t state_of(int64_t t) {
  const int64_t g = (int64_t)kChunk << 6;
  const int64_t blk = t / g;
  const int64_t r = t - blk * g;
  return ((blk << 6) + (r & 63)) * kChunk + (r >> 6);
}

// ---------------------------------------------------------------- linearization (forward, FP64)
template <int NIN, int NOUT, typename S>
PSM_DEV void lin_fwd(const ModelView<D, S>& md, const SigmaView<D>& sp, int use_sigma, const D (&m)[NIN],
                     const D (&L)[NIN][NIN], D (&F)[NOUT][NIN], D (&C)[NOUT][NOUT], D (&b)[NOUT]) {
  if (use_sigma) linearize_sigma<D, NIN, NOUT, true>(md, sp, m, L, F, C, b);
  else linearize_extended<D, NIN, NOUT>(md, m, F, C, b);
}

// Number of model parameters a kind reads (the slot prefix its cotangent uses).
template <int NIN, int NOUT>
PSM_DEV int n_params(int kind) {
  switch (kind) {
    case PSM_MODEL_LINEAR: return NIN * NOUT;
    case PSM_MODEL_COORDINATED_TURN: return 1;
    case PSM_MODEL_BEARINGS: return 4;
    default: return 0;
  }
}

// f and its Jacobian at x with a tangent on input j (j < NIN) or on model
// parameter p (j = NIN + p): returns the tangents of f and (WJ) of J.
template <int NIN, int NOUT, bool WJ = true, typename S>
PSM_DEV void model_tangent(const ModelView<D, S>& md, const D (&x)[NIN], int j, D (&df)[NOUT],
                           D (&dJ)[NOUT][NIN]) {
  using T = DualT<D>;
  T xt[NIN], yt[NOUT], Jt[NOUT][NIN];
#pragma unroll
  for (int a = 0; a < NIN; ++a) xt[a] = T(x[a], a == j ? 1.0 : 0.0);
  // parameters as tangent-carrying values: copy them into a local table
  constexpr int MAXP = (NIN * NOUT > 4 ? NIN * NOUT : 4);
  T par[MAXP];
  const int np = n_params<NIN, NOUT>(md.kind);
  for (int a = 0; a < MAXP; ++a) par[a] = T(a < np ? (D)md.params[a] : 0.0, (NIN + a == j) ? 1.0 : 0.0);
  ModelView<T, T> mt{md.kind, par, nullptr, nullptr, md.row, md.pt, md.npts};
  model_eval<T, NIN, NOUT, WJ>(mt, xt, yt, Jt);
#pragma unroll
  for (int a = 0; a < NOUT; ++a) {
    df[a] = yt[a].d;
#pragma unroll
    for (int c = 0; c < NIN; ++c) dJ[a][c] = WJ ? Jt[a][c].d : 0.0;
  }
}

// Reverse of the linearization at (m, L) for cotangents (Fb, bb, Cb) of
// (F, b, C): accumulates the nominal's cotangents into mb, Lb and the model's
// into pb (params), qmb (noise mean), qcb (noise chol).
template <int NIN, int NOUT, typename S>
PSM_DEV void lin_vjp(const ModelView<D, S>& md, const SigmaView<D>& sp, int use_sigma, const D (&m)[NIN],
                     const D (&L)[NIN][NIN], const D (&C)[NOUT][NOUT], const D (&Fb)[NOUT][NIN],
                     const D (&bb)[NOUT], const D (&Cb)[NOUT][NOUT], D (&mb)[NIN], D (&Lb)[NIN][NIN],
                     D* __restrict__ pb, D (&qmb)[NOUT], D (&qcb)[NOUT][NOUT]) {
  const int np = n_params<NIN, NOUT>(md.kind);
#pragma unroll
  for (int i = 0; i < NOUT; ++i) qmb[i] += bb[i];
  if (!use_sigma) {
    // F = J(m), b = f(m) - J(m) m + q_mean, C = q_chol:  with W = Fb - bb m^T,
    // m_bar_j = <W, dJ/dm_j>,  p_bar = <W, dJ/dp> + <bb, df/dp>
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) qcb[i][j] += Cb[i][j];
    D W[NOUT][NIN];
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NIN; ++j) W[i][j] = Fb[i][j] - bb[i] * m[j];
    for (int j = 0; j < NIN + np; ++j) {
      D df[NOUT], dJ[NOUT][NIN];
      model_tangent<NIN, NOUT>(md, m, j, df, dJ);
      D s = 0.0;
#pragma unroll
      for (int a = 0; a < NOUT; ++a) {
#pragma unroll
        for (int c = 0; c < NIN; ++c) s = fma(W[a][c], dJ[a][c], s);
        if (j >= NIN) s = fma(bb[a], df[a], s);
      }
      if (j < NIN) mb[j] += s;
      else pb[j - NIN] += s;
    }
    return;
  }
  // Sigma points, in the centred form of the regression (same F, b, C):
  //   dx_p = L xi_p, f_p = f(m + dx_p), m_f = sum wm f_p, df_p = f_p - m_f,
  //   Psi = sum wc dx_p df_p^T, F = Psi^T P^-1 (P = L L^T), r_p = df_p - F dx_p,
  //   C C^T = sum wc r_p r_p^T + Q, b = m_f - F m + q_mean.
  const int npts = sp.npts;
  D P[NIN][NIN], Pc[NIN][NIN];
  gram(L, P);
  cholesky(P, Pc);
  D mf[NOUT];
  set_zero(mf);
  for (int p = 0; p < npts; ++p) {
    D dx[NIN], x[NIN], f[NOUT];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      D s = 0.0;
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
      dx[i] = s;
      x[i] = m[i] + s;
    }
    ModelView<D, S> mp = md;
    mp.pt = p;
    model_value<D, NIN, NOUT>(mp, x, f);
#pragma unroll
    for (int i = 0; i < NOUT; ++i) mf[i] = fma(sp.pwm(p), f[i], mf[i]);
  }
  D Psi[NIN][NOUT];
  set_zero(Psi);
  for (int p = 0; p < npts; ++p) {
    D dx[NIN], x[NIN], f[NOUT];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      D s = 0.0;
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
      dx[i] = s;
      x[i] = m[i] + s;
    }
    ModelView<D, S> mp = md;
    mp.pt = p;
    model_value<D, NIN, NOUT>(mp, x, f);
    const D w = sp.pwc(p);
#pragma unroll
    for (int i = 0; i < NIN; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) Psi[i][j] = fma(w * dx[i], f[j] - mf[j], Psi[i][j]);
  }
  D F[NOUT][NIN];
  {
    D X[NIN][NOUT], Y[NIN][NOUT];
    solve_lower<D, NIN, NOUT>(Pc, Psi, X);
    solve_lower_T<D, NIN, NOUT>(Pc, X, Y);
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NIN; ++j) F[i][j] = Y[j][i];
  }
  // Omega_bar from C_bar; Q_bar = Omega_bar
  D Ob[NOUT][NOUT];
  chol_vjp(C, Cb, Ob);
  {
    D qc[NOUT][NOUT], t[NOUT][NOUT];
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) qc[i][j] = md.qc(i * NOUT + j);
    gram_vjp(Ob, qc, t);
#pragma unroll
    for (int i = 0; i < NOUT; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) qcb[i][j] += t[i][j];
  }
  // F_bar total: Fb - bb m^T - sum_p rb_p dx_p^T ; the residual part needs a
  // pass over the points, the rest is known now
  D Fbt[NOUT][NIN];
#pragma unroll
  for (int i = 0; i < NOUT; ++i)
#pragma unroll
    for (int j = 0; j < NIN; ++j) Fbt[i][j] = Fb[i][j] - bb[i] * m[j];
#pragma unroll
  for (int j = 0; j < NIN; ++j) {
    D s = 0.0;
#pragma unroll
    for (int i = 0; i < NOUT; ++i) s = fma(-F[i][j], bb[i], s);
    mb[j] += s;
  }
  // pass A: the residual cotangents rb_p = 2 wc_p Ob r_p add -rb_p dx_p^T to F_bar
  for (int p = 0; p < npts; ++p) {
    D dx[NIN], x[NIN], f[NOUT], r[NOUT];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      D s = 0.0;
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
      dx[i] = s;
      x[i] = m[i] + s;
    }
    ModelView<D, S> mp = md;
    mp.pt = p;
    model_value<D, NIN, NOUT>(mp, x, f);
#pragma unroll
    for (int i = 0; i < NOUT; ++i) {
      D s = f[i] - mf[i];
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(-F[i][j], dx[j], s);
      r[i] = s;
    }
    const D w2 = 2.0 * sp.pwc(p);
#pragma unroll
    for (int i = 0; i < NOUT; ++i) {
      D rb = 0.0;
#pragma unroll
      for (int k = 0; k < NOUT; ++k) rb = fma(Ob[i][k], r[k], rb);
      rb *= w2;
#pragma unroll
      for (int j = 0; j < NIN; ++j) Fbt[i][j] = fma(-rb, dx[j], Fbt[i][j]);
    }
  }
  // F = Psi^T P^-1:  Psi_bar = P^-1 Fbt^T,  P_bar = -P^-1 Psi Fbt P^-1 (symmetrised)
  D Psib[NIN][NOUT];
  {
    D Pb[NIN][NIN];
    D FbT[NIN][NOUT], X[NIN][NOUT];
#pragma unroll
    for (int i = 0; i < NIN; ++i)
#pragma unroll
      for (int j = 0; j < NOUT; ++j) FbT[i][j] = Fbt[j][i];
    solve_lower<D, NIN, NOUT>(Pc, FbT, X);
    solve_lower_T<D, NIN, NOUT>(Pc, X, Psib);
    // P^-1 Psi Fbt P^-1 = (P^-1 Psi) (P^-1 Fbt^T)^T = F^T Psib^T
#pragma unroll
    for (int i = 0; i < NIN; ++i)
#pragma unroll
      for (int j = 0; j < NIN; ++j) {
        D s = 0.0;
#pragma unroll
        for (int k = 0; k < NOUT; ++k) s = fma(F[k][i], Psib[j][k], s);
        Pb[i][j] = -s;
      }
    symmetrize(Pb);
    // P = L L^T: L_bar += 2 P_bar L (lower part)
#pragma unroll
    for (int i = 0; i < NIN; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        D s = 0.0;
#pragma unroll
        for (int k = 0; k < NIN; ++k) s = fma(Pb[i][k], L[k][j], s);
        Lb[i][j] = fma(2.0, s, Lb[i][j]);
      }
  }
  // pass B: f_bar_p = df_bar_p + wm_p m_f_bar with
  //   df_bar_p = rb_p + wc_p Psib^T dx_p, m_f_bar = bb - sum_p df_bar_p,
  //   dx_bar_p = -F^T rb_p + wc_p Psib df_p;  chi_bar_p = J(chi_p)^T f_bar_p
  // (m_f_bar needs every df_bar first: its sum is accumulated here and the
  // wm_p m_f_bar part of f_bar_p is applied in pass C)
  D sdfb[NOUT];
  set_zero(sdfb);
  for (int p = 0; p < npts; ++p) {
    D dx[NIN], x[NIN], f[NOUT], r[NOUT], J[NOUT][NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      D s = 0.0;
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
      dx[i] = s;
      x[i] = m[i] + s;
    }
    ModelView<D, S> mp = md;
    mp.pt = p;
    model_eval<D, NIN, NOUT, true>(mp, x, f, J);
#pragma unroll
    for (int i = 0; i < NOUT; ++i) {
      D s = f[i] - mf[i];
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(-F[i][j], dx[j], s);
      r[i] = s;
    }
    const D wc = sp.pwc(p);
    D rb[NOUT], dfb[NOUT];
#pragma unroll
    for (int i = 0; i < NOUT; ++i) {
      D s = 0.0;
#pragma unroll
      for (int k = 0; k < NOUT; ++k) s = fma(Ob[i][k], r[k], s);
      rb[i] = 2.0 * wc * s;
      D t = rb[i];
#pragma unroll
      for (int j = 0; j < NIN; ++j) t = fma(wc * Psib[j][i], dx[j], t);
      dfb[i] = t;
      sdfb[i] += t;
    }
    // dx_bar_p and the chi_bar_p of the df_bar part of f_bar_p
    D dxb[NIN], xb[NIN];
#pragma unroll
    for (int j = 0; j < NIN; ++j) {
      D s = 0.0;
#pragma unroll
      for (int i = 0; i < NOUT; ++i) s = fma(-F[i][j], rb[i], s);
#pragma unroll
      for (int i = 0; i < NOUT; ++i) s = fma(wc * Psib[j][i], f[i] - mf[i], s);
      dxb[j] = s;
      D c = 0.0;
#pragma unroll
      for (int i = 0; i < NOUT; ++i) c = fma(J[i][j], dfb[i], c);
      xb[j] = c;
    }
    // chi = m + dx, dx = L xi: m_bar += chi_bar, L_bar += (dx_bar + chi_bar) xi^T
#pragma unroll
    for (int j = 0; j < NIN; ++j) {
      mb[j] += xb[j];
      const D g = dxb[j] + xb[j];
#pragma unroll
      for (int c = 0; c <= j; ++c) Lb[j][c] = fma(g, sp.pxi(p * NIN + c), Lb[j][c]);
    }
  }
  // pass C: the wm_p m_f_bar part of f_bar_p, m_f_bar = bb - sum_p df_bar_p
  D mfb[NOUT];
#pragma unroll
  for (int i = 0; i < NOUT; ++i) mfb[i] = bb[i] - sdfb[i];
  for (int p = 0; p < npts; ++p) {
    D x[NIN], f[NOUT], J[NOUT][NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) {
      D s = m[i];
#pragma unroll
      for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
      x[i] = s;
    }
    ModelView<D, S> mp = md;
    mp.pt = p;
    model_eval<D, NIN, NOUT, true>(mp, x, f, J);
    const D wm = sp.pwm(p);
#pragma unroll
    for (int j = 0; j < NIN; ++j) {
      D c = 0.0;
#pragma unroll
      for (int i = 0; i < NOUT; ++i) c = fma(J[i][j], mfb[i], c);
      c *= wm;
      mb[j] += c;
#pragma unroll
      for (int k = 0; k <= j; ++k) Lb[j][k] = fma(c, sp.pxi(p * NIN + k), Lb[j][k]);
    }
  }
  // pass D: the model parameters, <f_bar_p, df(chi_p)/dp> with the whole
  // f_bar_p = df_bar_p + wm_p m_f_bar (a pass of its own: the tangent
  // evaluations need registers that passes B and C hold)
  if (np > 0) {
    for (int p = 0; p < npts; ++p) {
      D dx[NIN], x[NIN], f[NOUT], r[NOUT];
#pragma unroll
      for (int i = 0; i < NIN; ++i) {
        D s = 0.0;
#pragma unroll
        for (int j = 0; j < NIN; ++j) s = fma(L[i][j], sp.pxi(p * NIN + j), s);
        dx[i] = s;
        x[i] = m[i] + s;
      }
      ModelView<D, S> mp = md;
      mp.pt = p;
      model_value<D, NIN, NOUT>(mp, x, f);
#pragma unroll
      for (int i = 0; i < NOUT; ++i) {
        D s = f[i] - mf[i];
#pragma unroll
        for (int j = 0; j < NIN; ++j) s = fma(-F[i][j], dx[j], s);
        r[i] = s;
      }
      const D wc = sp.pwc(p), wm = sp.pwm(p);
      D fb[NOUT];
#pragma unroll
      for (int i = 0; i < NOUT; ++i) {
        D s = 0.0;
#pragma unroll
        for (int k = 0; k < NOUT; ++k) s = fma(Ob[i][k], r[k], s);
        D t = 2.0 * wc * s;
#pragma unroll
        for (int j = 0; j < NIN; ++j) t = fma(wc * Psib[j][i], dx[j], t);
        fb[i] = fma(wm, mfb[i], t);
      }
      for (int q = 0; q < np; ++q) {
        D dfp[NOUT], dJp[NOUT][NIN];
        model_tangent<NIN, NOUT, false>(mp, x, NIN + q, dfp, dJp);
        D s = 0.0;
#pragma unroll
        for (int i = 0; i < NOUT; ++i) s = fma(fb[i], dfp[i], s);
        pb[q] += s;
      }
    }
  }
}

// Checkpoint components per state, each nx x nx: Abar_k = (I - K_k H_k) F_k
// of step k (zero at a sequence's first state), G_k the smoother gain (zero at
// its last state), GP_k = G_{k-1}^T (zero at its first state), Pf_k, Ps_k the
// filtered and smoothed covariances, Ft_k the transition linearized at
// nominal k (used by step k + 1); then MF_k, the filtered mean (nx).
template <int NX> struct Ckp {
  static constexpr int S = NX * NX;
  static constexpr int A = 0, G = S, GP = 2 * S, PF = 3 * S, PS = 4 * S, FT = 5 * S, MF = 6 * S,
                       N = 6 * S + NX;
};
// Per-pass components: phi, lambda (vectors), Pbar (P_bar of the smoothed
// covariance), B, Psi (congruence outputs), WV / WM (scan offsets).
template <int NX> struct Pas {
  static constexpr int S = NX * NX;
  static constexpr int PHI = 0, LAM = NX, PB = 2 * NX, B = 2 * NX + S, PSI = 2 * NX + 2 * S,
                       WV = 2 * NX + 3 * S, WM = 3 * NX + 3 * S, N = 3 * NX + 4 * S;
};
// Cotangents of the linearized model per state: transition (used by step
// k + 1) F, b, C and observation H, c, C.  The linearized model itself, the
// same for every VJP at one fixed point, is kept in the same layout (computed
// once by k_adj_prep).
template <int NX, int NY> struct Lco {
  static constexpr int F = 0, B = NX * NX, CQ = NX * NX + NX, H = 2 * NX * NX + NX,
                       C = 2 * NX * NX + NX + NY * NX, CR = C + NY, N = CR + NY * NY;
};

// ---------------------------------------------------------------- checkpoint
// One thread per interleaved position.  nom / filt / smooth are the primal's
// AoS trajectories (type R); ck the FP64 checkpoint.
template <typename R, int NX, int NY>
__global__ __launch_bounds__(64) void k_adj_prep(int64_t M, int64_t Mp, const int64_t* __restrict__ soff,
                                                 int nseq, ModelView<D, R> tr, ModelView<D, R> ob,
                                                 SigmaView<D> sp, int use_sigma, const R* __restrict__ nom_m,
                                                 const R* __restrict__ nom_L, const R* __restrict__ fm,
                                                 const R* __restrict__ fL, const R* __restrict__ sL,
                                                 D* __restrict__ ck, D* __restrict__ ln) {
  using K = Ckp<NX>;
  using LC = Lco<NX, NY>;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mp) return;
  const int64_t i = state_of(t);
  if (i >= M) return;
  const int s = seq_of(soff, nseq, i);
  const bool first = i == soff[s], last = i == soff[s + 1] - 1;
  D L[NX][NX], Pf[NX][NX];
  ld_cv(fL, i, L);
  gram(L, Pf);
  stm(ck, Mp, t, K::PF, Pf);
  {
    D mf[NX];
    ld_cv(fm, i, mf);
    stv(ck, Mp, t, K::MF, mf);
  }
  ld_cv(sL, i, L);
  {
    D Ps[NX][NX];
    gram(L, Ps);
    stm(ck, Mp, t, K::PS, Ps);
  }
  D zero[NX][NX];
  set_zero(zero);
  // transition linearized at nominal i (step i + 1) and the smoother gain
  {
    D m[NX], Ln[NX][NX], F[NX][NX], C[NX][NX], b[NX];
    ld_cv(nom_m, i, m);
    ld_cv(nom_L, i, Ln);
    ModelView<D, R> mt = tr;
    mt.row = i;
    lin_fwd<NX, NX>(mt, sp, use_sigma, m, Ln, F, C, b);
    stm(ck, Mp, t, K::FT, F);
    stm(ln, Mp, t, LC::F, F);
    stv(ln, Mp, t, LC::B, b);
    stm(ln, Mp, t, LC::CQ, C);
    if (last) {
      stm(ck, Mp, t, K::G, zero);
    } else {
      // G = Pf F^T Pp^-1, Pp = F Pf F^T + C C^T
      D FP[NX][NX], Pp[NX][NX], Lp[NX][NX], X[NX][NX], G[NX][NX];
      mm(F, Pf, FP);
#pragma unroll
      for (int a = 0; a < NX; ++a)
#pragma unroll
        for (int c = 0; c < NX; ++c) {
          D acc = 0.0;
#pragma unroll
          for (int k = 0; k < NX; ++k) acc = fma(FP[a][k], F[c][k], acc);
#pragma unroll
          for (int k = 0; k < NX; ++k) acc = fma(C[a][k], C[c][k], acc);
          Pp[a][c] = acc;
        }
      cholesky(Pp, Lp);
      // G^T = Pp^-1 F Pf
      solve_lower<D, NX, NX>(Lp, FP, X);
      solve_lower_T<D, NX, NX>(Lp, X, G);
#pragma unroll
      for (int a = 0; a < NX; ++a)
#pragma unroll
        for (int c = 0; c < NX; ++c) X[a][c] = G[c][a];
      stm(ck, Mp, t, K::G, X);
    }
  }
  if (first) {
    stm(ck, Mp, t, K::A, zero);
    stm(ck, Mp, t, K::GP, zero);
    return;
  }
  // step i: transition at nominal i - 1, observation at nominal i
  D F[NX][NX], C[NX][NX], b[NX], Pp[NX][NX];
  {
    D m[NX], Ln[NX][NX], Pf0[NX][NX];
    ld_cv(nom_m, i - 1, m);
    ld_cv(nom_L, i - 1, Ln);
    ModelView<D, R> mt = tr;
    mt.row = i - 1;
    lin_fwd<NX, NX>(mt, sp, use_sigma, m, Ln, F, C, b);
    ld_cv(fL, i - 1, L);
    gram(L, Pf0);
    D FP[NX][NX];
    mm(F, Pf0, FP);
#pragma unroll
    for (int a = 0; a < NX; ++a)
#pragma unroll
      for (int c = 0; c < NX; ++c) {
        D acc = 0.0;
#pragma unroll
        for (int k = 0; k < NX; ++k) acc = fma(FP[a][k], F[c][k], acc);
#pragma unroll
        for (int k = 0; k < NX; ++k) acc = fma(C[a][k], C[c][k], acc);
        Pp[a][c] = acc;
      }
    // G_{i-1}^T: recomputed here from the same quantities as the gain of state i - 1
    D Lp[NX][NX], X[NX][NX], Gt[NX][NX];
    cholesky(Pp, Lp);
    solve_lower<D, NX, NX>(Lp, FP, X);
    solve_lower_T<D, NX, NX>(Lp, X, Gt);
    stm(ck, Mp, t, K::GP, Gt);
  }
  {
    D m[NX], Ln[NX][NX], H[NY][NX], Cr[NY][NY], c[NY];
    ld_cv(nom_m, i, m);
    ld_cv(nom_L, i, Ln);
    ModelView<D, R> mo = ob;
    mo.row = i;
    lin_fwd<NX, NY>(mo, sp, use_sigma, m, Ln, H, Cr, c);
    stm2(ln, Mp, t, LC::H, H);
    stv(ln, Mp, t, LC::C, c);
    stm(ln, Mp, t, LC::CR, Cr);
    // K = Pp H^T S^-1, S = H Pp H^T + Cr Cr^T;  Abar = (I - K H) F
    D PH[NX][NY], Sm[NY][NY], Ls[NY][NY], KT[NY][NX], X[NY][NX];
#pragma unroll
    for (int a = 0; a < NX; ++a)
#pragma unroll
      for (int r = 0; r < NY; ++r) {
        D acc = 0.0;
#pragma unroll
        for (int k = 0; k < NX; ++k) acc = fma(Pp[a][k], H[r][k], acc);
        PH[a][r] = acc;
      }
#pragma unroll
    for (int r = 0; r < NY; ++r)
#pragma unroll
      for (int q = 0; q < NY; ++q) {
        D acc = 0.0;
#pragma unroll
        for (int k = 0; k < NX; ++k) acc = fma(H[r][k], PH[k][q], acc);
#pragma unroll
        for (int k = 0; k < NY; ++k) acc = fma(Cr[r][k], Cr[q][k], acc);
        Sm[r][q] = acc;
      }
    cholesky(Sm, Ls);
    D HP[NY][NX];
#pragma unroll
    for (int r = 0; r < NY; ++r)
#pragma unroll
      for (int a = 0; a < NX; ++a) HP[r][a] = PH[a][r];
    solve_lower<D, NY, NX>(Ls, HP, X);
    solve_lower_T<D, NY, NX>(Ls, X, KT);   // K^T = S^-1 H Pp
    D IKH[NX][NX], A[NX][NX];
#pragma unroll
    for (int a = 0; a < NX; ++a)
#pragma unroll
      for (int c = 0; c < NX; ++c) {
        D acc = a == c ? 1.0 : 0.0;
#pragma unroll
        for (int r = 0; r < NY; ++r) acc = fma(-KT[r][a], H[r][c], acc);
        IKH[a][c] = acc;
      }
    mm(IKH, F, A);
    stm(ck, Mp, t, K::A, A);
  }
}

// ---------------------------------------------------------------- chunked affine scans
// z_j = M_j z_{j-1} + w_j (vectors) or z_j = M_j z_{j-1} M_j^T + w_j
// (matrices) over scan positions j, z_{-1} = 0; a sequence head has M_j = 0.
// Level 0 reads its items from the checkpoint / pass buffers (SRC), the levels
// above from aggregate buffers: lane c folds positions [cC, cC + C).
enum { SRC_FWD_VEC = 0, SRC_BWD_VEC = 1, SRC_BWD_CONG = 2, SRC_FWD_CONG = 3, SRC_LEVEL = 4 };

template <int NX, bool CONG> struct Sc {
  static constexpr int NW = CONG ? NX * NX : NX;
  static constexpr int QM = 0, QW = NX * NX, QZ = NX * NX + NW, NQ = NX * NX + 2 * NW;
};

template <int NX, bool CONG, int SRC>
struct ScanIO {
  using S = Sc<NX, CONG>;
  const D* ck;     // checkpoint (level 0)
  D* pa;           // pass buffer (level 0)
  D* lv;           // aggregate buffer (levels >= 1): items of this level
  int64_t Mp;      // component stride of ck / pa
  int64_t lst;     // component stride of lv
  int64_t M;       // states (level 0)
  PSM_DEV int64_t state(int64_t j) const {
    return (SRC == SRC_BWD_VEC || SRC == SRC_BWD_CONG) ? M - 1 - j : j;
  }
  PSM_DEV void load(int64_t j, D (&Mj)[NX][NX], D (&w)[S::NW]) const {
    if constexpr (SRC == SRC_LEVEL) {
      const int64_t p = StateMap{kChunk}(j);
      ldm(lv, lst, p, S::QM, Mj);
      ldv(lv, lst, p, S::QW, w);
    } else {
      using K = Ckp<NX>;
      using P = Pas<NX>;
      const int64_t p = pos_of(state(j));
      if constexpr (SRC == SRC_FWD_VEC) {
        ldm(ck, Mp, p, K::A, Mj);
        ldv(pa, Mp, p, P::WV, w);
      } else if constexpr (SRC == SRC_BWD_VEC) {
        ldm(ck, Mp, p, K::G, Mj);
        ldv(pa, Mp, p, P::WV, w);
      } else if constexpr (SRC == SRC_BWD_CONG) {
        ldm(ck, Mp, p, K::G, Mj);
        ldv(pa, Mp, p, P::WM, w);
      } else {
        ldm(ck, Mp, p, K::GP, Mj);
        ldv(pa, Mp, p, P::PB, w);
      }
    }
  }
  PSM_DEV void store(int64_t j, const D (&z)[S::NW]) const {
    if constexpr (SRC == SRC_LEVEL) {
      stv(lv, lst, StateMap{kChunk}(j), S::QZ, z);
    } else {
      using P = Pas<NX>;
      const int64_t p = pos_of(state(j));
      constexpr int q = SRC == SRC_FWD_VEC ? P::PHI : SRC == SRC_BWD_VEC ? P::LAM
                        : SRC == SRC_BWD_CONG ? P::B : P::PSI;
      stv(pa, Mp, p, q, z);
    }
  }
};

template <int NX, bool CONG>
PSM_DEV void step_z(const D (&Mj)[NX][NX], const D (&w)[Sc<NX, CONG>::NW], D (&z)[Sc<NX, CONG>::NW]) {
  if constexpr (CONG) {
    D T[NX][NX];
#pragma unroll
    for (int a = 0; a < NX; ++a)
#pragma unroll
      for (int c = 0; c < NX; ++c) {
        D acc = 0.0;
#pragma unroll
        for (int k = 0; k < NX; ++k) acc = fma(Mj[a][k], z[k * NX + c], acc);
        T[a][c] = acc;
      }
#pragma unroll
    for (int a = 0; a < NX; ++a)
#pragma unroll
      for (int c = 0; c < NX; ++c) {
        D acc = w[a * NX + c];
#pragma unroll
        for (int k = 0; k < NX; ++k) acc = fma(T[a][k], Mj[c][k], acc);
        z[a * NX + c] = acc;
      }
  } else {
    D y[NX];
#pragma unroll
    for (int a = 0; a < NX; ++a) {
      D acc = w[a];
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(Mj[a][k], z[k], acc);
      y[a] = acc;
    }
#pragma unroll
    for (int a = 0; a < NX; ++a) z[a] = y[a];
  }
}

// aggregates of chunks of C positions: (prod M, offset of the chunk applied to 0)
template <int NX, bool CONG, int SRC>
__global__ __launch_bounds__(64) void k_ascan_reduce(int64_t n, ScanIO<NX, CONG, SRC> io, D* __restrict__ up,
                                                     int64_t ust) {
  using S = Sc<NX, CONG>;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j0 = c * kChunk;
  if (j0 >= n) return;
  const int64_t j1 = min(j0 + (int64_t)kChunk, n);
  D Ma[NX][NX], z[S::NW];
  set_eye(Ma);
  set_zero(z);
  for (int64_t j = j0; j < j1; ++j) {
    D Mj[NX][NX], w[S::NW], T[NX][NX];
    io.load(j, Mj, w);
    step_z<NX, CONG>(Mj, w, z);
    mm(Mj, Ma, T);
    copy(T, Ma);
  }
  const int64_t p = StateMap{kChunk}(c);
  stm(up, ust, p, S::QM, Ma);
  stv(up, ust, p, S::QW, z);
}

// outputs of every position from the incoming z of its chunk (the output of
// the previous chunk, from the level above; 0 for chunk 0 or a one-chunk level)
template <int NX, bool CONG, int SRC>
__global__ __launch_bounds__(64) void k_ascan_apply(int64_t n, ScanIO<NX, CONG, SRC> io,
                                                    const D* __restrict__ up, int64_t ust) {
  using S = Sc<NX, CONG>;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j0 = c * kChunk;
  if (j0 >= n) return;
  const int64_t j1 = min(j0 + (int64_t)kChunk, n);
  D z[S::NW];
  if (c == 0 || up == nullptr) set_zero(z);
  else ldv(up, ust, StateMap{kChunk}(c - 1), S::QZ, z);
  for (int64_t j = j0; j < j1; ++j) {
    D Mj[NX][NX], w[S::NW];
    io.load(j, Mj, w);
    step_z<NX, CONG>(Mj, w, z);
    io.store(j, z);
  }
}

// ---------------------------------------------------------------- pass items
// cot_m (M, NX), cot_L (M, NX, NX): FP64 cotangents of the smoothed means and
// chol factors.  Writes Pbar, WV = Pf cot_m, WM = Ps Pbar Ps.
template <typename R, int NX>
__global__ __launch_bounds__(64) void k_adj_items(int64_t M, int64_t Mp, const R* __restrict__ sL,
                                                  const D* __restrict__ cot_m, const D* __restrict__ cot_L,
                                                  const D* __restrict__ ck, D* __restrict__ pa) {
  using K = Ckp<NX>;
  using P = Pas<NX>;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mp) return;
  const int64_t i = state_of(t);
  if (i >= M) return;
  D L[NX][NX], Lb[NX][NX], Pb[NX][NX];
  ld_cv(sL, i, L);
  ld_cv(cot_L, i, Lb);
  chol_vjp(L, Lb, Pb);
  stm(pa, Mp, t, P::PB, Pb);
  D Pf[NX][NX], v[NX], w[NX];
  ldm(ck, Mp, t, K::PF, Pf);
  ld_cv(cot_m, i, v);
  mv(Pf, v, w);
  stv(pa, Mp, t, P::WV, w);
  D Ps[NX][NX], T[NX][NX], W[NX][NX];
  ldm(ck, Mp, t, K::PS, Ps);
  mm(Ps, Pb, T);
  mm(T, Ps, W);
  stm(pa, Mp, t, P::WM, W);
}

// WV = phi - G Ft phi (offsets of the backward mean recursion)
template <int NX>
__global__ __launch_bounds__(64) void k_adj_items2(int64_t M, int64_t Mp, const D* __restrict__ ck,
                                                   D* __restrict__ pa) {
  using K = Ckp<NX>;
  using P = Pas<NX>;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mp) return;
  if (state_of(t) >= M) return;
  D G[NX][NX], F[NX][NX], phi[NX], a[NX], w[NX];
  ldm(ck, Mp, t, K::G, G);
  ldm(ck, Mp, t, K::FT, F);
  ldv(pa, Mp, t, P::PHI, phi);
  mv(F, phi, a);
#pragma unroll
  for (int r = 0; r < NX; ++r) {
    D acc = phi[r];
#pragma unroll
    for (int k = 0; k < NX; ++k) acc = fma(-G[r][k], a[k], acc);
    w[r] = acc;
  }
  stv(pa, Mp, t, P::WV, w);
}

// ---------------------------------------------------------------- cotangents of the linearized model
// MODE_OUTPUT: from the smoothed outputs' cotangents (the pass buffers);
// MODE_FISHER: of the log-likelihood (Fisher's identity).
enum { MODE_OUTPUT = 0, MODE_FISHER = 1 };

// Q^-1 X = C^-T C^-1 X for lower C
template <int N, int K2>
PSM_DEV void qinv_mul(const D (&C)[N][N], const D (&X)[N][K2], D (&Y)[N][K2]) {
  D T[N][K2];
  solve_lower<D, N, K2>(C, X, T);
  solve_lower_T<D, N, K2>(C, T, Y);
}

// Pp^-1 x, Pp^-1 X and Pp^-1 S Pp^-1 (S symmetric) for Pp = Lp Lp^T
template <int N>
PSM_DEV void pinv_vec(const D (&Lp)[N][N], const D (&x)[N], D (&y)[N]) {
  D X[N][1], Y[N][1];
#pragma unroll
  for (int a = 0; a < N; ++a) X[a][0] = x[a];
  qinv_mul<N, 1>(Lp, X, Y);
#pragma unroll
  for (int a = 0; a < N; ++a) y[a] = Y[a][0];
}
template <int N, int K2>
PSM_DEV void pinv_mul(const D (&Lp)[N][N], const D (&X)[N][K2], D (&Y)[N][K2]) {
  qinv_mul<N, K2>(Lp, X, Y);
}
template <int N>
PSM_DEV void pinv_sandwich(const D (&Lp)[N][N], const D (&S)[N][N], D (&Y)[N][N]) {
  D X[N][N], Z[N][N];
  whiten(Lp, S, X);                      // Lp^-1 S Lp^-T
  solve_lower_T<D, N, N>(Lp, X, Z);       // Lp^-T X
  D Zt[N][N];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int c = 0; c < N; ++c) Zt[a][c] = Z[c][a];
  solve_lower_T<D, N, N>(Lp, Zt, Y);      // Lp^-T (Lp^-T X)^T = Pp^-1 S Pp^-1 (symmetric)
  symmetrize(Y);
}

// cotangent of the chol C of Q from S, where Q_bar = sgn Q^-1 S Q^-1:
// 2 Q_bar C = 2 sgn C^-T (C^-1 S C^-T)
template <int N>
PSM_DEV void chol_from_inner(const D (&C)[N][N], const D (&S)[N][N], D sgn, D (&Cb)[N][N]) {
  D X[N][N], Y[N][N];
  whiten(C, S, X);
  solve_lower_T<D, N, N>(C, X, Y);
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int c = 0; c < N; ++c) Cb[a][c] = c > a ? 0.0 : 2.0 * sgn * Y[a][c];
}

// D_kk = B_k + Ps_k (Psi_k - Pbar_k) Ps_k and Lambda_bar_kk = -sym(lam m^T) - D_kk
template <int NX>
PSM_DEV void lambda_bar_diag(const D* __restrict__ ck, const D* __restrict__ pa, int64_t Mp, int64_t p,
                             const D (&lam)[NX], const D (&m)[NX], D (&Lb)[NX][NX]) {
  using K = Ckp<NX>;
  using P = Pas<NX>;
  D Ps[NX][NX], Psi[NX][NX], Pb[NX][NX], T[NX][NX];
  ldm(ck, Mp, p, K::PS, Ps);
  ldm(pa, Mp, p, P::PSI, Psi);
  ldm(pa, Mp, p, P::PB, Pb);
#pragma unroll
  for (int a = 0; a < NX; ++a)
#pragma unroll
    for (int c = 0; c < NX; ++c) Psi[a][c] -= Pb[a][c];
  mm(Ps, Psi, T);
  mm(T, Ps, Pb);   // Ps Phi Ps
  ldm(pa, Mp, p, P::B, Psi);
#pragma unroll
  for (int a = 0; a < NX; ++a)
#pragma unroll
    for (int c = 0; c < NX; ++c)
      Lb[a][c] = -0.5 * (lam[a] * m[c] + m[a] * lam[c]) - (Psi[a][c] + Pb[a][c]);
  symmetrize(Lb);
}

template <typename R, int NX, int NY, int MODE>
__global__ __launch_bounds__(64) void k_adj_chain(int64_t M, int64_t Mp, const int64_t* __restrict__ soff,
                                                  int nseq, ModelView<D, R> tr, ModelView<D, R> ob,
                                                  SigmaView<D> sp, int use_sigma, const R* __restrict__ ys,
                                                  const R* __restrict__ m0, const R* __restrict__ L0,
                                                  const R* __restrict__ nom_m, const R* __restrict__ nom_L,
                                                  const R* __restrict__ sm, const D* __restrict__ ck,
                                                  const D* __restrict__ ln, const D* __restrict__ pa,
                                                  D* __restrict__ lc, D* __restrict__ th, int64_t tst) {
  using K = Ckp<NX>;
  using P = Pas<NX>;
  using LC = Lco<NX, NY>;
  using TH = Th<NX, NY>;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mp) return;
  const int64_t i = state_of(t);
  if (i >= M) return;
  const int s = seq_of(soff, nseq, i);
  const bool first = i == soff[s], last = i == soff[s + 1] - 1;
  D mi[NX], lam_i[NX];
  ld_cv(sm, i, mi);
  if constexpr (MODE == MODE_OUTPUT) ldv(pa, Mp, t, P::LAM, lam_i);
  // ---- prior (first state of a sequence): cotangents of m0 and L0
  if (first) {
    D mz[NX], Lz[NX][NX], P0[NX][NX], dm[NX], mb[NX], Lb0[NX][NX];
    ld_cv(m0, s, mz);
    ld_cv(L0, s, Lz);
    if constexpr (MODE == MODE_OUTPUT) {
      D Lbd[NX][NX], S0[NX][NX];
      lambda_bar_diag<NX>(ck, pa, Mp, t, lam_i, mi, Lbd);
#pragma unroll
      for (int a = 0; a < NX; ++a)
#pragma unroll
        for (int c = 0; c < NX; ++c) S0[a][c] = Lbd[a][c] + lam_i[a] * mz[c];
      symmetrize(S0);
      D lv[NX][1], xv[NX][1];
#pragma unroll
      for (int a = 0; a < NX; ++a) lv[a][0] = lam_i[a];
      qinv_mul<NX, 1>(Lz, lv, xv);
#pragma unroll
      for (int a = 0; a < NX; ++a) mb[a] = xv[a][0];
      chol_from_inner<NX>(Lz, S0, -1.0, Lb0);
    } else {
      D Ps[NX][NX], S0[NX][NX];
      ldm(ck, Mp, t, K::PS, Ps);
      gram(Lz, P0);
#pragma unroll
      for (int a = 0; a < NX; ++a) dm[a] = mi[a] - mz[a];
#pragma unroll
      for (int a = 0; a < NX; ++a)
#pragma unroll
        for (int c = 0; c < NX; ++c) S0[a][c] = Ps[a][c] + dm[a] * dm[c] - P0[a][c];
      D dv[NX][1], xv[NX][1];
#pragma unroll
      for (int a = 0; a < NX; ++a) dv[a][0] = dm[a];
      qinv_mul<NX, 1>(Lz, dv, xv);
#pragma unroll
      for (int a = 0; a < NX; ++a) mb[a] = xv[a][0];
      chol_from_inner<NX>(Lz, S0, 0.5, Lb0);
    }
    stv(th, tst, t, TH::M0, mb);
    stm(th, tst, t, TH::L0, Lb0);
  } else {
    D zv[NX], zm[NX][NX];
    set_zero(zv);
    set_zero(zm);
    stv(th, tst, t, TH::M0, zv);
    stm(th, tst, t, TH::L0, zm);
  }
  D nm[NX], nL[NX][NX];
  ld_cv(nom_m, i, nm);
  ld_cv(nom_L, i, nL);
  // ---- transition of step i + 1 (linearized at nominal i)
  {
    D Fb[NX][NX], bb[NX], Cb[NX][NX];
    set_zero(Fb);
    set_zero(bb);
    set_zero(Cb);
    if (!last) {
      D F[NX][NX], C[NX][NX], b[NX];
      ldm(ln, Mp, t, LC::F, F);
      ldv(ln, Mp, t, LC::B, b);
      ldm(ln, Mp, t, LC::CQ, C);
      const int64_t tn = pos_of(i + 1);
      D mn[NX];
      ld_cv(sm, i + 1, mn);
      // Predicted moments of step i + 1 (Pp = F Pf_i F^T + C C^T = Lp Lp^T, mp =
      // F mf_i + b).  Every Q^-1 (.) of the precision form is rewritten through
      // I - F G_i = Q Pp^-1: Q^-1 (m_n - F m_i - b) = Pp^-1 (m_n - mp), and the
      // transition rows of Sigma P_bar Sigma are Q Pp^-1 (.), so no difference of
      // O(P)-sized terms is ever divided by Q (on the bearings-only model Q is
      // 1e-6 of P: the precision form loses six digits there, FP32 all of them).
      D Pp[NX][NX], Lp[NX][NX], delta[NX], kap[NX], Dn[NX][NX], G[NX][NX];
      {
        D Pf[NX][NX], FP[NX][NX], mf[NX], mp[NX];
        ldm(ck, Mp, t, K::PF, Pf);
        ldv(ck, Mp, t, K::MF, mf);
        mm(F, Pf, FP);
#pragma unroll
        for (int a = 0; a < NX; ++a)
#pragma unroll
          for (int c = 0; c < NX; ++c) {
            D acc = 0.0;
#pragma unroll
            for (int k = 0; k < NX; ++k) acc = fma(FP[a][k], F[c][k], acc);
#pragma unroll
            for (int k = 0; k < NX; ++k) acc = fma(C[a][k], C[c][k], acc);
            Pp[a][c] = acc;
          }
        cholesky(Pp, Lp);
        mv(F, mf, mp);
#pragma unroll
        for (int a = 0; a < NX; ++a) delta[a] = mn[a] - mp[a] - b[a];
        pinv_vec<NX>(Lp, delta, kap);
        D Psn[NX][NX];
        ldm(ck, Mp, tn, K::PS, Psn);
#pragma unroll
        for (int a = 0; a < NX; ++a)
#pragma unroll
          for (int c = 0; c < NX; ++c) Dn[a][c] = Psn[a][c] - Pp[a][c];
        ldm(ck, Mp, t, K::G, G);
      }
      D DG[NX][NX];   // (Ps_n - Pp) G_i^T
#pragma unroll
      for (int a = 0; a < NX; ++a)
#pragma unroll
        for (int c = 0; c < NX; ++c) {
          D acc = 0.0;
#pragma unroll
          for (int k = 0; k < NX; ++k) acc = fma(Dn[a][k], G[c][k], acc);
          DG[a][c] = acc;
        }
      D Qb[NX][NX];
      if constexpr (MODE == MODE_OUTPUT) {
        // rho = Q^-1 (lam_n - F lam_i) = Pp^-1 (lam_n - F phi_i)
        D lam_n[NX], phi[NX], Fp[NX], r[NX], rho[NX];
        ldv(pa, Mp, tn, P::LAM, lam_n);
        ldv(pa, Mp, t, P::PHI, phi);
        mv(F, phi, Fp);
#pragma unroll
        for (int a = 0; a < NX; ++a) r[a] = lam_n[a] - Fp[a];
        pinv_vec<NX>(Lp, r, rho);
        // Y = DG Psi_i;  X = B_n G_i^T + Y Ps_i;  Om = B_n + Y DG^T
        D Bn[NX][NX], Y[NX][NX], X[NX][NX], Om[NX][NX];
        {
          D Psi[NX][NX], Ps[NX][NX];
          ldm(pa, Mp, tn, P::B, Bn);
          ldm(pa, Mp, t, P::PSI, Psi);
          ldm(ck, Mp, t, K::PS, Ps);
          mm(DG, Psi, Y);
          mm(Y, Ps, X);
        }
#pragma unroll
        for (int a = 0; a < NX; ++a)
#pragma unroll
          for (int c = 0; c < NX; ++c) {
            D acc = X[a][c], acc2 = Bn[a][c];
#pragma unroll
            for (int k = 0; k < NX; ++k) {
              acc = fma(Bn[a][k], G[c][k], acc);
              acc2 = fma(Y[a][k], DG[c][k], acc2);
            }
            X[a][c] = 2.0 * acc;
            Om[a][c] = acc2;
          }
        // F_bar = kap lam_i^T + rho m_i^T + 2 Pp^-1 X;  b_bar = rho
        pinv_mul<NX, NX>(Lp, X, Fb);
#pragma unroll
        for (int a = 0; a < NX; ++a) {
          bb[a] = rho[a];
#pragma unroll
          for (int c = 0; c < NX; ++c) Fb[a][c] += kap[a] * lam_i[c] + rho[a] * mi[c];
        }
        // Q_bar = sym(rho kap^T) + Pp^-1 Om Pp^-1
        symmetrize(Om);
        pinv_sandwich<NX>(Lp, Om, Qb);
#pragma unroll
        for (int a = 0; a < NX; ++a)
#pragma unroll
          for (int c = 0; c < NX; ++c) Qb[a][c] += 0.5 * (rho[a] * kap[c] + kap[a] * rho[c]);
      } else {
        // Fisher: E[e x_i^T] and E[e e^T] - Q of e = x_n - F x_i - b over the
        // smoothed distribution, through the same identity:
        //   F_bar = Pp^-1 (delta m_i^T + (Ps_n - Pp) G_i^T),  b_bar = Pp^-1 delta,
        //   Q_bar = 1/2 Pp^-1 (delta delta^T + Ps_n - Pp) Pp^-1.
        D X[NX][NX];
#pragma unroll
        for (int a = 0; a < NX; ++a)
#pragma unroll
          for (int c = 0; c < NX; ++c) X[a][c] = delta[a] * mi[c] + DG[a][c];
        pinv_mul<NX, NX>(Lp, X, Fb);
#pragma unroll
        for (int a = 0; a < NX; ++a) bb[a] = kap[a];
#pragma unroll
        for (int a = 0; a < NX; ++a)
#pragma unroll
          for (int c = 0; c < NX; ++c) X[a][c] = 0.5 * (delta[a] * delta[c] + Dn[a][c]);
        symmetrize(X);
        pinv_sandwich<NX>(Lp, X, Qb);
      }
      gram_vjp<NX>(Qb, C, Cb);
    }
    stm(lc, Mp, t, LC::F, Fb);
    stv(lc, Mp, t, LC::B, bb);
    stm(lc, Mp, t, LC::CQ, Cb);
  }
  // ---- observation at state i
  {
    D Hb[NY][NX], cb[NY], Rb[NY][NY];
    set_zero(Hb);
    set_zero(cb);
    set_zero(Rb);
    if (!first) {
      D H[NY][NX], Cr[NY][NY], c[NY], y[NY], d[NY];
      ldm2(ln, Mp, t, LC::H, H);
      ldv(ln, Mp, t, LC::C, c);
      ldm(ln, Mp, t, LC::CR, Cr);
      ld_cv(ys, i - s - 1, y);
#pragma unroll
      for (int r = 0; r < NY; ++r) d[r] = y[r] - c[r];
      if constexpr (MODE == MODE_OUTPUT) {
        D Cd[NX][NX], X[NY][NX], HC[NY][NX];
        lambda_bar_diag<NX>(ck, pa, Mp, t, lam_i, mi, Cd);
        mm(H, Cd, HC);
#pragma unroll
        for (int r = 0; r < NY; ++r)
#pragma unroll
          for (int a = 0; a < NX; ++a) X[r][a] = 2.0 * HC[r][a] + d[r] * lam_i[a];
        qinv_mul<NY, NX>(Cr, X, Hb);
        D Hl[NY], v[NY][1], vb[NY][1];
        mv(H, lam_i, Hl);
#pragma unroll
        for (int r = 0; r < NY; ++r) v[r][0] = -Hl[r];
        qinv_mul<NY, 1>(Cr, v, vb);
#pragma unroll
        for (int r = 0; r < NY; ++r) cb[r] = vb[r][0];
        D Sr[NY][NY];
#pragma unroll
        for (int r = 0; r < NY; ++r)
#pragma unroll
          for (int q = 0; q < NY; ++q) {
            D acc = Hl[r] * d[q];
#pragma unroll
            for (int k = 0; k < NX; ++k) acc = fma(HC[r][k], H[q][k], acc);
            Sr[r][q] = acc;
          }
        symmetrize(Sr);
        chol_from_inner<NY>(Cr, Sr, -1.0, Rb);
      } else {
        D Ps[NX][NX], HP[NY][NX], eps[NY], Hm[NY];
        ldm(ck, Mp, t, K::PS, Ps);
        mm(H, Ps, HP);
        mv(H, mi, Hm);
#pragma unroll
        for (int r = 0; r < NY; ++r) eps[r] = d[r] - Hm[r];
        // H_bar = R^-1 ((y - c) m^T - H (Ps + m m^T)) = R^-1 (eps m^T - H Ps)
        D X[NY][NX];
#pragma unroll
        for (int r = 0; r < NY; ++r)
#pragma unroll
          for (int a = 0; a < NX; ++a) X[r][a] = eps[r] * mi[a] - HP[r][a];
        qinv_mul<NY, NX>(Cr, X, Hb);
        D v[NY][1], vb[NY][1];
#pragma unroll
        for (int r = 0; r < NY; ++r) v[r][0] = eps[r];
        qinv_mul<NY, 1>(Cr, v, vb);
#pragma unroll
        for (int r = 0; r < NY; ++r) cb[r] = vb[r][0];
        D Sr[NY][NY];
#pragma unroll
        for (int r = 0; r < NY; ++r)
#pragma unroll
          for (int q = 0; q < NY; ++q) {
            D acc = eps[r] * eps[q];
#pragma unroll
            for (int k = 0; k < NX; ++k) acc = fma(HP[r][k], H[q][k], acc);
#pragma unroll
            for (int k = 0; k < NY; ++k) acc = fma(-Cr[r][k], Cr[q][k], acc);
            Sr[r][q] = acc;
          }
        symmetrize(Sr);
        chol_from_inner<NY>(Cr, Sr, 0.5, Rb);
      }
    }
    stm2(lc, Mp, t, LC::H, Hb);
    stv(lc, Mp, t, LC::C, cb);
    stm(lc, Mp, t, LC::CR, Rb);
  }
}

// ---------------------------------------------------------------- through the linearizations
// Nominal cotangents out_m / out_L (FP64 AoS; plus add_m / add_L when given)
// and per-state model-parameter cotangents th.
// Two launches per VJP, one per model (PART 0: the transition of step i + 1,
// linearized at nominal i; PART 1: the observation at state i), each with the
// extended (US = 0) or the sigma-point (US = 1) path compiled alone: one
// thread holding both models and both paths needed 512 registers and spilled
// 1.7 KB per lane.  PART 0 writes the nominal's cotangents (plus add_m /
// add_L), PART 1 adds its own.
template <typename R, int NX, int NY, int US, int PART>
__global__ __launch_bounds__(64) void k_adj_lin(int64_t M, int64_t Mp, const int64_t* __restrict__ soff,
                                                int nseq, ModelView<D, R> tr, ModelView<D, R> ob, SigmaView<D> sp,
                                                const R* __restrict__ nom_m, const R* __restrict__ nom_L,
                                                const D* __restrict__ ln, const D* __restrict__ lc,
                                                const D* __restrict__ add_m,
                                                const D* __restrict__ add_L, D* __restrict__ out_m,
                                                D* __restrict__ out_L, D* __restrict__ th, int64_t tst) {
  using LC = Lco<NX, NY>;
  using TH = Th<NX, NY>;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mp) return;
  const int64_t i = state_of(t);
  if (i >= M) return;
  const int s = seq_of(soff, nseq, i);
  const bool first = i == soff[s], last = i == soff[s + 1] - 1;
  D m[NX], L[NX][NX], mb[NX], Lb[NX][NX];
  ld_cv(nom_m, i, m);
  ld_cv(nom_L, i, L);
  if constexpr (PART == 0) {
    set_zero(mb);
    set_zero(Lb);
    D pt[NX * NX], qmb[NX], qcb[NX][NX];
#pragma unroll
    for (int a = 0; a < NX * NX; ++a) pt[a] = 0.0;
    set_zero(qmb);
    set_zero(qcb);
    if (!last) {
      D C[NX][NX], Fb[NX][NX], bb[NX], Cb[NX][NX];
      ModelView<D, R> mt = tr;
      mt.row = i;
      ldm(ln, Mp, t, LC::CQ, C);
      ldm(lc, Mp, t, LC::F, Fb);
      ldv(lc, Mp, t, LC::B, bb);
      ldm(lc, Mp, t, LC::CQ, Cb);
      lin_vjp<NX, NX>(mt, sp, US, m, L, C, Fb, bb, Cb, mb, Lb, pt, qmb, qcb);
    }
#pragma unroll
    for (int a = 0; a < NX * NX; ++a) th[(int64_t)(TH::TP + a) * tst + t] = pt[a];
    stv(th, tst, t, TH::TQM, qmb);
    stm(th, tst, t, TH::TQC, qcb);
    if (add_m != nullptr) {
      D am[NX], aL[NX][NX];
      ld_cv(add_m, i, am);
      ld_cv(add_L, i, aL);
#pragma unroll
      for (int a = 0; a < NX; ++a) {
        mb[a] += am[a];
#pragma unroll
        for (int c = 0; c < NX; ++c) Lb[a][c] += aL[a][c];
      }
    }
  } else {
    ld_cv(out_m, i, mb);
    ld_cv(out_L, i, Lb);
    D po[NY * NX], qmb[NY], qcb[NY][NY];
#pragma unroll
    for (int a = 0; a < NY * NX; ++a) po[a] = 0.0;
    set_zero(qmb);
    set_zero(qcb);
    if (!first) {
      D C[NY][NY], Hb[NY][NX], cb[NY], Cb[NY][NY];
      ModelView<D, R> mo = ob;
      mo.row = i;
      ldm(ln, Mp, t, LC::CR, C);
      ldm2(lc, Mp, t, LC::H, Hb);
      ldv(lc, Mp, t, LC::C, cb);
      ldm(lc, Mp, t, LC::CR, Cb);
      lin_vjp<NX, NY>(mo, sp, US, m, L, C, Hb, cb, Cb, mb, Lb, po, qmb, qcb);
    }
#pragma unroll
    for (int a = 0; a < NY * NX; ++a) th[(int64_t)(TH::OP + a) * tst + t] = po[a];
    stv(th, tst, t, TH::OQM, qmb);
    stm(th, tst, t, TH::OQC, qcb);
  }
  st_aos(out_m, i, mb);
  st_aos(out_L, i, Lb);
}

// per-sequence sums of the per-state parameter cotangents: out (nseq, NTH)
template <int NX, int NY>
__global__ __launch_bounds__(256) void k_adj_theta(int64_t Mp, const int64_t* __restrict__ soff,
                                                   const D* __restrict__ th, int64_t tst, D* __restrict__ out) {
  constexpr int NTH = Th<NX, NY>::N;
  __shared__ D red[256];
  const int s = blockIdx.x;
  const int q = blockIdx.y;
  const int64_t a = soff[s], b = soff[s + 1];
  D acc = 0.0;
  for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) acc += th[(int64_t)q * tst + pos_of(i)];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(int64_t)s * NTH + q] = red[0];
}

}  // namespace adj
}  // namespace psm
