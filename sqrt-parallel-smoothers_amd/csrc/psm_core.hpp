// The filter / smoother algebra of the parallel square-root (and covariance
// form) Kalman smoother, per lane, in registers.
//
// Filter aggregate (A, b, U, eta, Z) over a run of steps i..j represents
//   p(x_j | x_{i-1}, y_i..y_j)  = N(A x_{i-1} + b, U U^T)
//   p(y_i..y_j | x_{i-1})       ∝ exp(eta^T x - 1/2 x^T Z Z^T x)
// exactly as the elements of parsmooth/parallel/_filtering.py:115-146 and the
// operator of parsmooth/parallel/_operators.py:43-77.  A run that contains
// the first state of a sequence is a plain distribution (A = 0, flag HEAD);
// any combination with a HEAD on the right returns that right operand, which
// is how several sequences share one scan.
//
// Smoother aggregate (g, E, D): x_i = E x_{j+1} + g + N(0, D D^T)
// (parsmooth/parallel/_smoothing.py:72-85, _operators.py:104-125); the last
// state of a sequence is the HEAD (E = 0).
#pragma once
#include <type_traits>

#include "psm_models.hpp"

namespace psm {

enum { FLAG_NORMAL = 0, FLAG_HEAD = 1, FLAG_IDENT = 2 };


// ----------------------------------------------------------------------------
// linearized step (transition at state k-1, observation at state k)
// ----------------------------------------------------------------------------
template <typename R, int NX, int NY>
struct Step {
  R F[NX][NX];
  R cq[NX][NX];   // chol Q (sqrt form) or Q
  R bt[NX];
  R H[NY][NX];
  R cr[NY][NY];   // chol R (sqrt form) or R
  R c[NY];
  static constexpr int NTRANS = NX * NX + NX * (NX + 1) / 2 + NX;
  static constexpr int NOBS = NY * NX + NY * (NY + 1) / 2 + NY;
  static constexpr int NQ = NTRANS + NOBS;
};

template <typename R, int NX>
struct FAgg {
  R A[NX][NX];
  R b[NX];
  R U[NX][NX];    // U (sqrt) or C (covariance)
  R eta[NX];
  R Z[NX][NX];    // Z (sqrt) or J (covariance)
  int flag;
  static constexpr int NQ = 3 * NX * NX + 2 * NX + 1;
};

template <typename R, int NX>
struct SAgg {
  R g[NX];
  R E[NX][NX];
  R D[NX][NX];    // D (sqrt) or L (covariance)
  int flag;
  static constexpr int NQ = 2 * NX * NX + NX + 1;
};

// ---------------------------------------------------------------- SoA load / store
// Component q of item i lives at base[q * stride + i].
template <typename R, int N>
PSM_DEV void ld(const R* __restrict__ base, int64_t stride, int64_t i, int& q, R (&v)[N]) {
#pragma unroll
  for (int a = 0; a < N; ++a) v[a] = base[(int64_t)(q++) * stride + i];
}
template <typename R, int M, int N>
PSM_DEV void ld(const R* __restrict__ base, int64_t stride, int64_t i, int& q, R (&v)[M][N]) {
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) v[a][b] = base[(int64_t)(q++) * stride + i];
}
template <typename R, int N>
PSM_DEV void st(R* __restrict__ base, int64_t stride, int64_t i, int& q, const R (&v)[N]) {
#pragma unroll
  for (int a = 0; a < N; ++a) base[(int64_t)(q++) * stride + i] = v[a];
}
template <typename R, int M, int N>
PSM_DEV void st(R* __restrict__ base, int64_t stride, int64_t i, int& q, const R (&v)[M][N]) {
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) base[(int64_t)(q++) * stride + i] = v[a][b];
}

// Lower triangle (row by row) of an N x N noise factor or covariance; on load
// the upper triangle is zero (factor) or the mirror image (SYM: covariance).
template <typename R, int N>
PSM_DEV void st_tri(R* __restrict__ base, int64_t stride, int64_t i, int& q, const R (&v)[N][N]) {
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) base[(int64_t)(q++) * stride + i] = v[a][b];
}
template <bool SYM, typename R, int N>
PSM_DEV void ld_tri(const R* __restrict__ base, int64_t stride, int64_t i, int& q, R (&v)[N][N]) {
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) v[a][b] = base[(int64_t)(q++) * stride + i];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = a + 1; b < N; ++b) v[a][b] = SYM ? v[b][a] : R(0);
}

template <bool SYM, typename R, int NX, int NY>
PSM_DEV void load_trans(const R* base, int64_t stride, int64_t i, Step<R, NX, NY>& s) {
  int q = 0;
  ld(base, stride, i, q, s.F);
  ld_tri<SYM>(base, stride, i, q, s.cq);
  ld(base, stride, i, q, s.bt);
}
template <bool SYM, typename R, int NX, int NY>
PSM_DEV void load_obs(const R* base, int64_t stride, int64_t i, Step<R, NX, NY>& s) {
  int q = Step<R, NX, NY>::NTRANS;
  ld(base, stride, i, q, s.H);
  ld_tri<SYM>(base, stride, i, q, s.cr);
  ld(base, stride, i, q, s.c);
}

template <typename R, int NX>
PSM_DEV void load_agg(const R* base, int64_t stride, int64_t i, FAgg<R, NX>& a) {
  int q = 0;
  ld(base, stride, i, q, a.A);
  ld(base, stride, i, q, a.b);
  ld(base, stride, i, q, a.U);
  ld(base, stride, i, q, a.eta);
  ld(base, stride, i, q, a.Z);
  a.flag = (int)base[(int64_t)q * stride + i];
}
template <typename R, int NX>
PSM_DEV void store_agg(R* base, int64_t stride, int64_t i, const FAgg<R, NX>& a) {
  int q = 0;
  st(base, stride, i, q, a.A);
  st(base, stride, i, q, a.b);
  st(base, stride, i, q, a.U);
  st(base, stride, i, q, a.eta);
  st(base, stride, i, q, a.Z);
  base[(int64_t)q * stride + i] = R(a.flag);
}
template <typename R, int NX>
PSM_DEV void load_agg(const R* base, int64_t stride, int64_t i, SAgg<R, NX>& a) {
  int q = 0;
  ld(base, stride, i, q, a.g);
  ld(base, stride, i, q, a.E);
  ld(base, stride, i, q, a.D);
  a.flag = (int)base[(int64_t)q * stride + i];
}
template <typename R, int NX>
PSM_DEV void store_agg(R* base, int64_t stride, int64_t i, const SAgg<R, NX>& a) {
  int q = 0;
  st(base, stride, i, q, a.g);
  st(base, stride, i, q, a.E);
  st(base, stride, i, q, a.D);
  base[(int64_t)q * stride + i] = R(a.flag);
}

template <typename A> struct Agg_nx;
template <typename R, int NX> struct Agg_nx<FAgg<R, NX>> { static constexpr int value = NX; };
template <typename R, int NX> struct Agg_nx<SAgg<R, NX>> { static constexpr int value = NX; };

template <typename R, int NX>
PSM_DEV void set_identity(FAgg<R, NX>& a) {
  set_eye(a.A);
  set_zero(a.b);
  set_zero(a.U);
  set_zero(a.eta);
  set_zero(a.Z);
  a.flag = FLAG_IDENT;
}
template <typename R, int NX>
PSM_DEV void set_identity(SAgg<R, NX>& a) {
  set_zero(a.g);
  set_eye(a.E);
  set_zero(a.D);
  a.flag = FLAG_IDENT;
}
template <typename R, int NX>
PSM_DEV void copy_agg(const FAgg<R, NX>& s, FAgg<R, NX>& d) {
  copy(s.A, d.A); copy(s.b, d.b); copy(s.U, d.U); copy(s.eta, d.eta); copy(s.Z, d.Z);
  d.flag = s.flag;
}
template <typename R, int NX>
PSM_DEV void copy_agg(const SAgg<R, NX>& s, SAgg<R, NX>& d) {
  copy(s.g, d.g); copy(s.E, d.E); copy(s.D, d.D);
  d.flag = s.flag;
}

// ============================================================================
// Square-root form
// ============================================================================

// Predict + update of a distribution (m, L) with one step (sequential sqrt
// Kalman filter, parsmooth/sequential/_filtering.py:80-108).  Returns the
// log-likelihood increment (FP64) of parsmooth/_utils.py:152-160 evaluated at
// the innovation factor, which is what parsmooth/parallel/_filtering.py:149-154
// computes from the filtered state of the previous step.
template <typename R, int NX, int NY>
PSM_DEV typename AccOf<R>::type sqrt_kf_step(const Step<R, NX, NY>& s, const R (&y)[NY], R (&m)[NX], R (&L)[NX][NX]) {
  R m1[NX];
  mv(s.F, m, m1);
#pragma unroll
  for (int i = 0; i < NX; ++i) m1[i] += s.bt[i];
  R M[NX][2 * NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(s.F[i][k], L[k][j], acc);
      M[i][j] = acc;
      M[i][NX + j] = s.cq[i][j];
    }
  tria_inplace<R, NX, 2 * NX>(M);
  R P[NY + NX][NX + NY];
#pragma unroll
  for (int i = 0; i < NY; ++i) {
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(s.H[i][k], M[k][j], acc);
      P[i][j] = acc;
    }
#pragma unroll
    for (int j = 0; j < NY; ++j) P[i][NX + j] = s.cr[i][j];
  }
#pragma unroll
  for (int i = 0; i < NX; ++i) {
#pragma unroll
    for (int j = 0; j < NX; ++j) P[NY + i][j] = M[i][j];
#pragma unroll
    for (int j = 0; j < NY; ++j) P[NY + i][NX + j] = R(0);
  }
  tria_inplace<R, NY + NX, NX + NY>(P);
  // innovation v = y - H m1 - c ; w = P11^{-1} v
  R w[NY];
  using Acc = typename AccOf<R>::type;
  Acc quad = Acc(0.0), logdet = Acc(0.0);
#pragma unroll
  for (int i = 0; i < NY; ++i) {
    R v = y[i] - s.c[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) v = fma(-s.H[i][k], m1[k], v);
#pragma unroll
    for (int j = 0; j < i; ++j) v = fma(-P[i][j], w[j], v);
    w[i] = v / P[i][i];
    quad += to_acc(w[i]) * to_acc(w[i]);
    logdet += log_(abs_(to_acc(P[i][i])));
  }
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R acc = m1[i];
#pragma unroll
    for (int j = 0; j < NY; ++j) acc = fma(P[NY + i][j], w[j], acc);
    m[i] = acc;
#pragma unroll
    for (int j = 0; j < NX; ++j) L[i][j] = P[NY + i][NY + j];
  }
  return Acc(-0.5) * quad - logdet - Acc(0.5 * NY * 1.8378770664093454836);  // log(2 pi)
}

// Absorb one more step into a filter aggregate: agg <- agg (+) element(step).
// Same conditional + likelihood algebra as the operator, specialised to a
// right operand that is a single step, which needs no 2nx x 2nx
// triangularisation.
template <typename R, int NX, int NY>
PSM_DEV void sqrt_absorb(FAgg<R, NX>& a, const Step<R, NX, NY>& s, const R (&y)[NY]) {
  // predict: A1 = F A, b1 = F b + bt, U1 = tria([F U | cq])
  R A1[NX][NX], b1[NX];
  mm(s.F, a.A, A1);
  mv(s.F, a.b, b1);
#pragma unroll
  for (int i = 0; i < NX; ++i) b1[i] += s.bt[i];
  R M[NX][2 * NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(s.F[i][k], a.U[k][j], acc);
      M[i][j] = acc;
      M[i][NX + j] = s.cq[i][j];
    }
  tria_inplace<R, NX, 2 * NX>(M);
  R P[NY + NX][NX + NY];
#pragma unroll
  for (int i = 0; i < NY; ++i) {
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(s.H[i][k], M[k][j], acc);
      P[i][j] = acc;
    }
#pragma unroll
    for (int j = 0; j < NY; ++j) P[i][NX + j] = s.cr[i][j];
  }
#pragma unroll
  for (int i = 0; i < NX; ++i) {
#pragma unroll
    for (int j = 0; j < NX; ++j) P[NY + i][j] = M[i][j];
#pragma unroll
    for (int j = 0; j < NY; ++j) P[NY + i][NX + j] = R(0);
  }
  tria_inplace<R, NY + NX, NX + NY>(P);
  // w = P11^{-1} (y - H b1 - c), G = P11^{-1} H A1
  R w[NY], G[NY][NX];
#pragma unroll
  for (int i = 0; i < NY; ++i) {
    const R inv = rcp_(P[i][i]);
    R v = y[i] - s.c[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) v = fma(-s.H[i][k], b1[k], v);
#pragma unroll
    for (int j = 0; j < i; ++j) v = fma(-P[i][j], w[j], v);
    w[i] = v * inv;
#pragma unroll
    for (int col = 0; col < NX; ++col) {
      R h = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) h = fma(s.H[i][k], A1[k][col], h);
#pragma unroll
      for (int j = 0; j < i; ++j) h = fma(-P[i][j], G[j][col], h);
      G[i][col] = h * inv;
    }
  }
  // A = A1 - P21 G ; b = b1 + P21 w ; U = P22 ; eta += G^T w ; Z = tria([Z | G^T])
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R bb = b1[i];
#pragma unroll
    for (int j = 0; j < NY; ++j) bb = fma(P[NY + i][j], w[j], bb);
    a.b[i] = bb;
#pragma unroll
    for (int col = 0; col < NX; ++col) {
      R acc = A1[i][col];
#pragma unroll
      for (int j = 0; j < NY; ++j) acc = fma(-P[NY + i][j], G[j][col], acc);
      a.A[i][col] = acc;
    }
#pragma unroll
    for (int j = 0; j < NX; ++j) a.U[i][j] = P[NY + i][NY + j];
    R e = a.eta[i];
#pragma unroll
    for (int j = 0; j < NY; ++j) e = fma(G[j][i], w[j], e);
    a.eta[i] = e;
  }
  R Zc[NX][NX + NY];
#pragma unroll
  for (int i = 0; i < NX; ++i) {
#pragma unroll
    for (int j = 0; j < NX; ++j) Zc[i][j] = a.Z[i][j];
#pragma unroll
    for (int j = 0; j < NY; ++j) Zc[i][NX + j] = G[j][i];
  }
  tria_inplace<R, NX, NX + NY>(Zc);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) a.Z[i][j] = Zc[i][j];
}

// Full filtering operator out = e1 (+) e2 (parsmooth/parallel/_operators.py:43-77),
// written with M1 = U1 Xi11^{-T} and P = I - M1 Xi21^T:
//   A = A2 P A1, b = A2 P (b1 + U1 U1^T eta2) + b2, U = tria([A2 M1 | U2]),
//   eta = A1^T P^T (eta2 - Z2 Z2^T b1) + eta1, Z = tria([A1^T Xi22 | Z1]).
// out may alias e2 but not e1.
template <typename R, int NX>
PSM_DEV void sqrt_combine_core(const FAgg<R, NX>& e1, const FAgg<R, NX>& e2, FAgg<R, NX>& out) {
  R X[2 * NX][2 * NX];
  // [[U1^T Z2, I], [Z2, 0]]
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(e1.U[k][i], e2.Z[k][j], acc);
      X[i][j] = acc;
      X[i][NX + j] = (i == j) ? R(1) : R(0);
      X[NX + i][j] = e2.Z[i][j];
      X[NX + i][NX + j] = R(0);
    }
  tria_inplace<R, 2 * NX, 2 * NX>(X);
  // M1 = U1 Xi11^{-T}  (rows of U1 solved against Xi11^T from the right)
  R M1[NX][NX];
#pragma unroll
  for (int j = 0; j < NX; ++j) {
    const R inv = rcp_(X[j][j]);
#pragma unroll
    for (int r = 0; r < NX; ++r) {
      R s = e1.U[r][j];
#pragma unroll
      for (int k = 0; k < j; ++k) s = fma(-M1[r][k], X[j][k], s);
      M1[r][j] = s * inv;
    }
  }
  // P = I - M1 Xi21^T
  R Pm[NX][NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = (i == j) ? R(1) : R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(-M1[i][k], X[NX + j][k], acc);
      Pm[i][j] = acc;
    }
  // v = b1 + U1 (U1^T eta2) ; w = eta2 - Z2 (Z2^T b1)
  R t1[NX], t2[NX], v[NX], w[NX];
  mTv(e1.U, e2.eta, t1);
  mTv(e2.Z, e1.b, t2);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R a1 = e1.b[i], a2 = e2.eta[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      a1 = fma(e1.U[i][k], t1[k], a1);
      a2 = fma(-e2.Z[i][k], t2[k], a2);
    }
    v[i] = a1;
    w[i] = a2;
  }
  // PA1 = P A1, Pv = P v, PTw = P^T w
  R PA1[NX][NX], Pv[NX], PTw[NX];
  mm(Pm, e1.A, PA1);
  mv(Pm, v, Pv);
  mTv(Pm, w, PTw);
  // eta = A1^T PTw + eta1
  R eta[NX];
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R acc = e1.eta[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) acc = fma(e1.A[k][i], PTw[k], acc);
    eta[i] = acc;
  }
  // Z = tria([A1^T Xi22 | Z1])
  R Zc[NX][2 * NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = j; k < NX; ++k) acc = fma(e1.A[k][i], X[NX + k][NX + j], acc);
      Zc[i][j] = acc;
      Zc[i][NX + j] = e1.Z[i][j];
    }
  // U = tria([A2 M1 | U2])
  R Uc[NX][2 * NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(e2.A[i][k], M1[k][j], acc);
      Uc[i][j] = acc;
      Uc[i][NX + j] = e2.U[i][j];
    }
  // A = A2 PA1 ; b = A2 Pv + b2
  R A[NX][NX], b[NX];
  mm(e2.A, PA1, A);
  mv(e2.A, Pv, b);
#pragma unroll
  for (int i = 0; i < NX; ++i) b[i] += e2.b[i];
  tria_inplace<R, NX, 2 * NX>(Uc);
  tria_inplace<R, NX, 2 * NX>(Zc);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    out.b[i] = b[i];
    out.eta[i] = eta[i];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      out.A[i][j] = A[i][j];
      out.U[i][j] = Uc[i][j];
      out.Z[i][j] = Zc[i][j];
    }
  }
}

// Filter element (parsmooth/parallel/_filtering.py:115-146) with a prior
// (m, L) at the previous state.  Z is returned triangularised (the reference
// zero-pads when nx > ny; Z Z^T is the same).
template <typename R, int NX, int NY>
PSM_DEV void sqrt_filter_element(const Step<R, NX, NY>& s, const R (&m)[NX], const R (&L)[NX][NX],
                                 const R (&y)[NY], FAgg<R, NX>& out) {
  R m1[NX];
  mv(s.F, m, m1);
#pragma unroll
  for (int i = 0; i < NX; ++i) m1[i] += s.bt[i];
  R M[NX][2 * NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(s.F[i][k], L[k][j], acc);
      M[i][j] = acc;
      M[i][NX + j] = s.cq[i][j];
    }
  tria_inplace<R, NX, 2 * NX>(M);
  R P[NY + NX][NX + NY];
#pragma unroll
  for (int i = 0; i < NY; ++i) {
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(s.H[i][k], M[k][j], acc);
      P[i][j] = acc;
    }
#pragma unroll
    for (int j = 0; j < NY; ++j) P[i][NX + j] = s.cr[i][j];
  }
#pragma unroll
  for (int i = 0; i < NX; ++i) {
#pragma unroll
    for (int j = 0; j < NX; ++j) P[NY + i][j] = M[i][j];
#pragma unroll
    for (int j = 0; j < NY; ++j) P[NY + i][NX + j] = R(0);
  }
  tria_inplace<R, NY + NX, NX + NY>(P);
  R HF[NY][NX];
  mm(s.H, s.F, HF);
  // G = P11^{-1} HF ; w1 = P11^{-1}(y - H m1 - c) ; w2 = P11^{-1}(y - H bt - c)
  R G[NY][NX], w1[NY], w2[NY];
#pragma unroll
  for (int i = 0; i < NY; ++i) {
    const R inv = rcp_(P[i][i]);
    R v1 = y[i] - s.c[i], v2 = y[i] - s.c[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      v1 = fma(-s.H[i][k], m1[k], v1);
      v2 = fma(-s.H[i][k], s.bt[k], v2);
    }
#pragma unroll
    for (int j = 0; j < i; ++j) {
      v1 = fma(-P[i][j], w1[j], v1);
      v2 = fma(-P[i][j], w2[j], v2);
    }
    w1[i] = v1 * inv;
    w2[i] = v2 * inv;
#pragma unroll
    for (int col = 0; col < NX; ++col) {
      R h = HF[i][col];
#pragma unroll
      for (int j = 0; j < i; ++j) h = fma(-P[i][j], G[j][col], h);
      G[i][col] = h * inv;
    }
  }
  // K = P21 P11^{-1}:  A = F - K HF = F - P21 G ; b = m1 + P21 w1
  // Z0 = G^T ; eta = Z0 P11^{-1} (y - H bt - c) = G^T w2
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R bb = m1[i];
#pragma unroll
    for (int j = 0; j < NY; ++j) bb = fma(P[NY + i][j], w1[j], bb);
    out.b[i] = bb;
#pragma unroll
    for (int col = 0; col < NX; ++col) {
      R acc = s.F[i][col];
#pragma unroll
      for (int j = 0; j < NY; ++j) acc = fma(-P[NY + i][j], G[j][col], acc);
      out.A[i][col] = acc;
    }
#pragma unroll
    for (int j = 0; j < NX; ++j) out.U[i][j] = P[NY + i][NY + j];
    R e = R(0);
#pragma unroll
    for (int j = 0; j < NY; ++j) e = fma(G[j][i], w2[j], e);
    out.eta[i] = e;
  }
  R Zc[NX][NY];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NY; ++j) Zc[i][j] = G[j][i];
  tria_inplace<R, NX, NY>(Zc);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) out.Z[i][j] = (j < NY && j <= i) ? Zc[i][j] : R(0);
  out.flag = FLAG_NORMAL;
}

// Smoother element (parsmooth/parallel/_smoothing.py:72-85):
// Phi = tria([[F L, cq], [L, 0]]), E = Phi21 Phi11^{-1}, g = m - E (F m + bt), D = Phi22.
template <typename R, int NX>
PSM_DEV void sqrt_smoother_element(const R (&F)[NX][NX], const R (&cq)[NX][NX], const R (&bt)[NX],
                                   const R (&m)[NX], const R (&L)[NX][NX], SAgg<R, NX>& out) {
  R X[2 * NX][2 * NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(F[i][k], L[k][j], acc);
      X[i][j] = acc;
      X[i][NX + j] = cq[i][j];
      X[NX + i][j] = L[i][j];
      X[NX + i][NX + j] = R(0);
    }
  tria_inplace<R, 2 * NX, 2 * NX>(X);
  // E = Phi21 Phi11^{-1}: E Phi11 = Phi21, solved column by column from the right
#pragma unroll
  for (int jj = 0; jj < NX; ++jj) {
    const int j = NX - 1 - jj;
    const R inv = rcp_(X[j][j]);
#pragma unroll
    for (int r = 0; r < NX; ++r) {
      R s = X[NX + r][j];
#pragma unroll
      for (int k = j + 1; k < NX; ++k) s = fma(-out.E[r][k], X[k][j], s);
      out.E[r][j] = s * inv;
    }
  }
  R p[NX];
  mv(F, m, p);
#pragma unroll
  for (int i = 0; i < NX; ++i) p[i] += bt[i];
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R acc = m[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) acc = fma(-out.E[i][k], p[k], acc);
    out.g[i] = acc;
#pragma unroll
    for (int j = 0; j < NX; ++j) out.D[i][j] = X[NX + i][NX + j];
  }
  out.flag = FLAG_NORMAL;
}

// Smoothing operator with e1 the later aggregate and e2 the earlier one
// (parsmooth/parallel/_operators.py:104-125): g = E2 g1 + g2, E = E2 E1,
// D = tria([E2 D1 | D2]).  Result written to e2.
template <typename R, int NX>
PSM_DEV void sqrt_smoother_combine_core(const SAgg<R, NX>& e1, SAgg<R, NX>& e2) {
  R g[NX], E[NX][NX], M[NX][2 * NX];
  mv(e2.E, e1.g, g);
  mm(e2.E, e1.E, E);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = j; k < NX; ++k) acc = fma(e2.E[i][k], e1.D[k][j], acc);
      M[i][j] = acc;
      M[i][NX + j] = e2.D[i][j];
    }
  tria_inplace<R, NX, 2 * NX>(M);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    e2.g[i] += g[i];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      e2.E[i][j] = E[i][j];
      e2.D[i][j] = M[i][j];
    }
  }
}

// Apply a smoother element to the smoothed distribution of the next state
// (parsmooth/sequential/_smoothing.py:60-77 in element form):
// m <- g + E m ; Ds <- tria([E Ds | D]).
template <typename R, int NX>
PSM_DEV void sqrt_smoother_apply(const SAgg<R, NX>& e, R (&m)[NX], R (&Ds)[NX][NX]) {
  R g[NX], M[NX][2 * NX];
  mv(e.E, m, g);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(e.E[i][k], Ds[k][j], acc);
      M[i][j] = acc;
      M[i][NX + j] = e.D[i][j];
    }
  tria_inplace<R, NX, 2 * NX>(M);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    m[i] = g[i] + e.g[i];
#pragma unroll
    for (int j = 0; j < NX; ++j) Ds[i][j] = M[i][j];
  }
}

// ============================================================================
// Covariance form (MVNStandard), parsmooth/parallel/_filtering.py:78-97,
// _operators.py:7-40,80-101, parsmooth/parallel/_smoothing.py:60-69
// ============================================================================

// S^{-1} B for symmetric positive definite S via Cholesky.
template <typename R, int N, int K>
PSM_DEV void spd_solve(const R (&S)[N][N], const R (&B)[N][K], R (&X)[N][K]) {
  R Lc[N][N], T[N][K];
  cholesky(S, Lc);
  solve_lower(Lc, B, T);
  solve_lower_T(Lc, T, X);
}

template <typename R, int N>
PSM_DEV void solve_lower_T_v(const R (&L)[N][N], const R (&b)[N], R (&x)[N]) {
#pragma unroll
  for (int ii = 0; ii < N; ++ii) {
    const int i = N - 1 - ii;
    R s = b[i];
#pragma unroll
    for (int j = i + 1; j < N; ++j) s = fma(-L[j][i], x[j], s);
    x[i] = s * rcp_(L[i][i]);
  }
}

// Covariance-form predict + update (parsmooth/sequential/_filtering.py:56-77);
// log-likelihood from the Cholesky factor of S as in
// parsmooth/parallel/_filtering.py:157-163.
template <typename R, int NX, int NY>
PSM_DEV typename AccOf<R>::type std_kf_step(const Step<R, NX, NY>& s, const R (&y)[NY], R (&m)[NX], R (&P)[NX][NX]) {
  R m1[NX], FP[NX][NX], P1[NX][NX];
  mv(s.F, m, m1);
#pragma unroll
  for (int i = 0; i < NX; ++i) m1[i] += s.bt[i];
  mm(s.F, P, FP);
  mmT(FP, s.F, P1);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) P1[i][j] += s.cq[i][j];
  R HP[NY][NX], S[NY][NY];
  mm(s.H, P1, HP);
  mmT(HP, s.H, S);
#pragma unroll
  for (int i = 0; i < NY; ++i)
#pragma unroll
    for (int j = 0; j < NY; ++j) S[i][j] += s.cr[i][j];
  R LS[NY][NY];
  cholesky(S, LS);
  R v[NY], w[NY], u[NY];
#pragma unroll
  for (int i = 0; i < NY; ++i) {
    R a = y[i] - s.c[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) a = fma(-s.H[i][k], m1[k], a);
    v[i] = a;
  }
  solve_lower_v(LS, v, w);
  solve_lower_T_v(LS, w, u);
  using Acc = typename AccOf<R>::type;
  Acc quad = Acc(0.0), logdet = Acc(0.0);
#pragma unroll
  for (int i = 0; i < NY; ++i) {
    quad += to_acc(w[i]) * to_acc(w[i]);
    logdet += log_(abs_(to_acc(LS[i][i])));
  }
  // KT = S^{-1} H P1 ; m = m1 + (H P1)^T S^{-1} v ; P = P1 - (H P1)^T S^{-1} (H P1)
  R T1[NY][NX], KT[NY][NX];
  solve_lower(LS, HP, T1);
  solve_lower_T(LS, T1, KT);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R acc = m1[i];
#pragma unroll
    for (int j = 0; j < NY; ++j) acc = fma(HP[j][i], u[j], acc);
    m[i] = acc;
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R a = P1[i][j];
#pragma unroll
      for (int k = 0; k < NY; ++k) a = fma(-HP[k][i], KT[k][j], a);
      P[i][j] = a;
    }
  }
  return Acc(-0.5) * quad - logdet - Acc(0.5 * NY * 1.8378770664093454836);
}

// Covariance-form absorb of one step into (A, b, C, eta, J).
template <typename R, int NX, int NY>
PSM_DEV void std_absorb(FAgg<R, NX>& a, const Step<R, NX, NY>& s, const R (&y)[NY]) {
  R A1[NX][NX], b1[NX], FC[NX][NX], C1[NX][NX];
  mm(s.F, a.A, A1);
  mv(s.F, a.b, b1);
#pragma unroll
  for (int i = 0; i < NX; ++i) b1[i] += s.bt[i];
  mm(s.F, a.U, FC);
  mmT(FC, s.F, C1);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) C1[i][j] += s.cq[i][j];
  R HC[NY][NX], S[NY][NY], HA[NY][NX];
  mm(s.H, C1, HC);
  mmT(HC, s.H, S);
  mm(s.H, A1, HA);
#pragma unroll
  for (int i = 0; i < NY; ++i)
#pragma unroll
    for (int j = 0; j < NY; ++j) S[i][j] += s.cr[i][j];
  R LS[NY][NY];
  cholesky(S, LS);
  R v[NY], w[NY], u[NY];
#pragma unroll
  for (int i = 0; i < NY; ++i) {
    R t = y[i] - s.c[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) t = fma(-s.H[i][k], b1[k], t);
    v[i] = t;
  }
  solve_lower_v(LS, v, w);
  solve_lower_T_v(LS, w, u);             // u = S^{-1} v
  R T1[NY][NX], KT[NY][NX], G[NY][NX], GA[NY][NX];
  solve_lower(LS, HC, T1);
  solve_lower_T(LS, T1, KT);             // KT = S^{-1} H C1 = K^T
  solve_lower(LS, HA, G);                // G = LS^{-1} H A1
  solve_lower_T(LS, G, GA);              // GA = S^{-1} H A1
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R bb = b1[i];
#pragma unroll
    for (int j = 0; j < NY; ++j) bb = fma(KT[j][i], v[j], bb);
    a.b[i] = bb;
#pragma unroll
    for (int col = 0; col < NX; ++col) {
      R acc = A1[i][col];
#pragma unroll
      for (int j = 0; j < NY; ++j) acc = fma(-KT[j][i], HA[j][col], acc);
      a.A[i][col] = acc;
      R cc = C1[i][col];
#pragma unroll
      for (int j = 0; j < NY; ++j) cc = fma(-KT[j][i], HC[j][col], cc);
      a.U[i][col] = cc;
      R jj = a.Z[i][col];
#pragma unroll
      for (int j = 0; j < NY; ++j) jj = fma(HA[j][i], GA[j][col], jj);
      a.Z[i][col] = jj;
    }
    R e = a.eta[i];
#pragma unroll
    for (int j = 0; j < NY; ++j) e = fma(HA[j][i], u[j], e);
    a.eta[i] = e;
  }
}

// Covariance-form filtering operator (parsmooth/parallel/_operators.py:7-40).
template <typename R, int NX>
PSM_DEV void std_combine_core(const FAgg<R, NX>& e1, const FAgg<R, NX>& e2, FAgg<R, NX>& out) {
  // X1 = A2 (I + C1 J2)^{-1}  via (I + C1 J2)^T X1^T = A2^T
  // X2 = A1^T (I + J2 C1)^{-1} via (I + J2 C1)^T X2^T = A1
  R M1[NX][NX], M2[NX][NX], B1[NX][NX], B2[NX][NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R s1 = (i == j) ? R(1) : R(0), s2 = s1;
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        s1 = fma(e1.U[j][k], e2.Z[k][i], s1);    // (I + C1 J2)^T [i][j]
        s2 = fma(e2.Z[j][k], e1.U[k][i], s2);    // (I + J2 C1)^T [i][j]
      }
      M1[i][j] = s1;
      M2[i][j] = s2;
      B1[i][j] = e2.A[j][i];
      B2[i][j] = e1.A[i][j];
    }
  solve_general<R, NX, NX>(M1, B1);   // B1 = X1^T
  solve_general<R, NX, NX>(M2, B2);   // B2 = X2^T
  R X1[NX][NX], X2[NX][NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      X1[i][j] = B1[j][i];
      X2[i][j] = B2[j][i];
    }
  R A[NX][NX], t[NX], b[NX], CA[NX][NX], C[NX][NX], t2[NX], eta[NX], JA[NX][NX], J[NX][NX];
  mm(X1, e1.A, A);
  mv(e1.U, e2.eta, t);
#pragma unroll
  for (int i = 0; i < NX; ++i) t[i] += e1.b[i];
  mv(X1, t, b);
  mmT(e1.U, e2.A, CA);       // C1 A2^T
  mm(X1, CA, C);
  mv(e2.Z, e1.b, t2);
#pragma unroll
  for (int i = 0; i < NX; ++i) t2[i] = e2.eta[i] - t2[i];
  mv(X2, t2, eta);
  mm(e2.Z, e1.A, JA);
  mm(X2, JA, J);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    out.b[i] = b[i] + e2.b[i];
    out.eta[i] = eta[i] + e1.eta[i];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      out.A[i][j] = A[i][j];
      out.U[i][j] = C[i][j] + e2.U[i][j];
      out.Z[i][j] = J[i][j] + e1.Z[i][j];
    }
  }
}

// Covariance-form filter element (parsmooth/parallel/_filtering.py:78-97).
template <typename R, int NX, int NY>
PSM_DEV void std_filter_element(const Step<R, NX, NY>& s, const R (&m0)[NX], const R (&P0)[NX][NX],
                                const R (&y)[NY], FAgg<R, NX>& out) {
  R m[NX], FP[NX][NX], P[NX][NX];
  mv(s.F, m0, m);
#pragma unroll
  for (int i = 0; i < NX; ++i) m[i] += s.bt[i];
  mm(s.F, P0, FP);
  mmT(FP, s.F, P);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) P[i][j] += s.cq[i][j];
  R HP[NY][NX], S[NY][NY];
  mm(s.H, P, HP);
  mmT(HP, s.H, S);
#pragma unroll
  for (int i = 0; i < NY; ++i)
#pragma unroll
    for (int j = 0; j < NY; ++j) S[i][j] += s.cr[i][j];
  // S_invH = S^{-1} H ; K = (S_invH P)^T
  R SiH[NY][NX], K[NX][NY];
  spd_solve<R, NY, NX>(S, s.H, SiH);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NY; ++j) {
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(SiH[j][k], P[k][i], acc);
      K[i][j] = acc;
    }
  R HF[NY][NX], KHF[NX][NX];
  mm(s.H, s.F, HF);
  mm(K, HF, KHF);
  R v[NY];
#pragma unroll
  for (int i = 0; i < NY; ++i) {
    R a = y[i] - s.c[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) a = fma(-s.H[i][k], m[k], a);
    v[i] = a;
  }
  R KS[NX][NY], KSK[NX][NX];
  mm(K, S, KS);
  mmT(KS, K, KSK);
  // temp = (S_invH F)^T (nx x ny)
  R SiHF[NY][NX];
  mm(SiH, s.F, SiHF);
  R w[NY];
#pragma unroll
  for (int i = 0; i < NY; ++i) {
    R a = y[i] - s.c[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) a = fma(-s.H[i][k], s.bt[k], a);
    w[i] = a;
  }
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R bb = m[i];
#pragma unroll
    for (int j = 0; j < NY; ++j) bb = fma(K[i][j], v[j], bb);
    out.b[i] = bb;
    R e = R(0);
#pragma unroll
    for (int j = 0; j < NY; ++j) e = fma(SiHF[j][i], w[j], e);
    out.eta[i] = e;
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      out.A[i][j] = s.F[i][j] - KHF[i][j];
      out.U[i][j] = P[i][j] - KSK[i][j];
      R jj = R(0);
#pragma unroll
      for (int k = 0; k < NY; ++k) jj = fma(SiHF[k][i], HF[k][j], jj);
      out.Z[i][j] = jj;
    }
  }
  out.flag = FLAG_NORMAL;
}

// Covariance-form smoother element (parsmooth/parallel/_smoothing.py:60-69):
// Pp = F P F^T + Q, E = (Pp^{-1} F P)^T, g = m - E (F m + bt), L = P - E Pp E^T.
template <typename R, int NX>
PSM_DEV void std_smoother_element(const R (&F)[NX][NX], const R (&Q)[NX][NX], const R (&bt)[NX],
                                  const R (&m)[NX], const R (&P)[NX][NX], SAgg<R, NX>& out) {
  R FP[NX][NX], Pp[NX][NX], ET[NX][NX];
  mm(F, P, FP);
  mmT(FP, F, Pp);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) Pp[i][j] += Q[i][j];
  spd_solve<R, NX, NX>(Pp, FP, ET);
  R p[NX], EPp[NX][NX];
  mv(F, m, p);
#pragma unroll
  for (int i = 0; i < NX; ++i) p[i] += bt[i];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) out.E[i][j] = ET[j][i];
  mm(out.E, Pp, EPp);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R acc = m[i];
#pragma unroll
    for (int k = 0; k < NX; ++k) acc = fma(-out.E[i][k], p[k], acc);
    out.g[i] = acc;
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      R a = P[i][j];
#pragma unroll
      for (int k = 0; k < NX; ++k) a = fma(-EPp[i][k], out.E[j][k], a);
      out.D[i][j] = a;
    }
  }
  out.flag = FLAG_NORMAL;
}

// Covariance-form smoothing operator (parsmooth/parallel/_operators.py:95-101);
// result written to e2 (e1 = later aggregate).
template <typename R, int NX>
PSM_DEV void std_smoother_combine_core(const SAgg<R, NX>& e1, SAgg<R, NX>& e2) {
  R g[NX], E[NX][NX], EL[NX][NX], L[NX][NX];
  mv(e2.E, e1.g, g);
  mm(e2.E, e1.E, E);
  mm(e2.E, e1.D, EL);
  mmT(EL, e2.E, L);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    e2.g[i] += g[i];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      e2.E[i][j] = E[i][j];
      e2.D[i][j] += L[i][j];
    }
  }
}

template <typename R, int NX>
PSM_DEV void std_smoother_apply(const SAgg<R, NX>& e, R (&m)[NX], R (&Ps)[NX][NX]) {
  R g[NX], EP[NX][NX], P[NX][NX];
  mv(e.E, m, g);
  mm(e.E, Ps, EP);
  mmT(EP, e.E, P);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    m[i] = g[i] + e.g[i];
#pragma unroll
    for (int j = 0; j < NX; ++j) Ps[i][j] = P[i][j] + e.D[i][j];
  }
}

// ============================================================================
// form dispatch + segmented (flagged) combines
// ============================================================================
template <typename R, int NX, int NY, bool SQRT>
PSM_DEV typename AccOf<R>::type kf_step(const Step<R, NX, NY>& s, const R (&y)[NY], R (&m)[NX], R (&L)[NX][NX]) {
  if constexpr (SQRT) return sqrt_kf_step<R, NX, NY>(s, y, m, L);
  else return std_kf_step<R, NX, NY>(s, y, m, L);
}

template <typename R, int NX, int NY, bool SQRT>
PSM_DEV void absorb(FAgg<R, NX>& a, const Step<R, NX, NY>& s, const R (&y)[NY]) {
  if constexpr (SQRT) sqrt_absorb<R, NX, NY>(a, s, y);
  else std_absorb<R, NX, NY>(a, s, y);
  if (a.flag == FLAG_IDENT) a.flag = FLAG_NORMAL;
}

// e2 <- e1 (+) e2 (e1 precedes e2 in scan order)
template <typename R, int NX, bool SQRT>
PSM_DEV void combine_into(const FAgg<R, NX>& e1, FAgg<R, NX>& e2) {
  if (e2.flag == FLAG_HEAD || e1.flag == FLAG_IDENT) return;
  if (e2.flag == FLAG_IDENT) { copy_agg(e1, e2); return; }
  if constexpr (SQRT) sqrt_combine_core<R, NX>(e1, e2, e2);
  else std_combine_core<R, NX>(e1, e2, e2);
  e2.flag = e1.flag;
}

template <typename R, int NX, bool SQRT>
PSM_DEV void combine_into(const SAgg<R, NX>& e1, SAgg<R, NX>& e2) {
  if (e2.flag == FLAG_HEAD || e1.flag == FLAG_IDENT) return;
  if (e2.flag == FLAG_IDENT) { copy_agg(e1, e2); return; }
  if constexpr (SQRT) sqrt_smoother_combine_core<R, NX>(e1, e2);
  else std_smoother_combine_core<R, NX>(e1, e2);
  e2.flag = e1.flag;
}

template <typename R, int NX, bool SQRT>
PSM_DEV void smoother_element(const R (&F)[NX][NX], const R (&cq)[NX][NX], const R (&bt)[NX],
                              const R (&m)[NX], const R (&L)[NX][NX], SAgg<R, NX>& out) {
  if constexpr (SQRT) sqrt_smoother_element<R, NX>(F, cq, bt, m, L, out);
  else std_smoother_element<R, NX>(F, cq, bt, m, L, out);
}

template <typename R, int NX, bool SQRT>
PSM_DEV void smoother_apply(const SAgg<R, NX>& e, R (&m)[NX], R (&Ds)[NX][NX]) {
  if constexpr (SQRT) sqrt_smoother_apply<R, NX>(e, m, Ds);
  else std_smoother_apply<R, NX>(e, m, Ds);
}

template <typename R, int NX, int NY, bool SQRT>
PSM_DEV void filter_element(const Step<R, NX, NY>& s, const R (&m)[NX], const R (&L)[NX][NX],
                            const R (&y)[NY], FAgg<R, NX>& out) {
  if constexpr (SQRT) sqrt_filter_element<R, NX, NY>(s, m, L, y, out);
  else std_filter_element<R, NX, NY>(s, m, L, y, out);
}

// ============================================================================
// FP32 square-root form: FP64 re-triangularisation fallback
// ============================================================================
// Two ways for an FP32 square-root kernel to meet an ill-conditioned result.
// Fallback recomputes the operation in FP64 in place (the FP64 code is part of
// the kernel and sets its register budget).  FallbackFlag keeps the FP32
// result and raises its workgroup's flag word; the launch sequence then runs
// the same kernel on the same grid again with a Fallback whose guard is that
// array of words: a workgroup whose flag is down returns at once, one whose
// flag is up clears it and recomputes everything it wrote in the first launch
// (every kernel reads only buffers it does not write, and workgroups write
// disjoint outputs, so the second run replaces the first run's outputs where
// it repeats them).  The FP32 kernels that run every pass thus carry no FP64
// code, and the FP64 work is proportional to the workgroups that need it.
// skip() must be reached by every thread of the workgroup (it synchronises).
struct Fallback {
  float thr;                          // min|diag| / max|diag| threshold (0: off)
  unsigned long long* cnt;            // incremented per FP64 recomputation (may be null)
  unsigned* guard = nullptr;          // non-null: per-workgroup flags of the first launch
  static constexpr bool kRedo = true;
  PSM_DEV bool skip() const {
    if (guard == nullptr) return false;
    const unsigned f = guard[blockIdx.x];
    __syncthreads();
    if (f != 0u && threadIdx.x == 0) guard[blockIdx.x] = 0u;
    return f == 0u;
  }
};

struct FallbackFlag {
  float thr;
  unsigned* flag;                     // flag[blockIdx.x] <- 1 on an ill-conditioned result
  static constexpr bool kRedo = false;
  PSM_DEV bool skip() const { return false; }
  PSM_DEV void raise() const { flag[blockIdx.x] = 1u; }
};

template <typename R, bool SQRT>
constexpr bool kFallback = SQRT && std::is_same<R, float>::value;

// true when min|diag| < thr * max|diag| (or a diagonal entry is not finite)
template <typename R, int N>
PSM_DEV bool ill_conditioned(const R (&L)[N][N], float thr) {
  R mn = abs_(L[0][0]), mx = mn;
#pragma unroll
  for (int i = 1; i < N; ++i) {
    const R d = abs_(L[i][i]);
    mn = d < mn ? d : mn;
    mx = d > mx ? d : mx;
  }
  return !(mn >= R(thr) * mx);
}

PSM_DEV void count_fallback(const Fallback& fb) {
  if (fb.cnt) atomicAdd(fb.cnt, 1ull);
}

template <int N> PSM_DEV void cvt(const float (&a)[N], double (&b)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) b[i] = a[i];
}
template <int N> PSM_DEV void cvt(const double (&a)[N], float (&b)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) b[i] = (float)a[i];
}
template <int M, int N> PSM_DEV void cvt(const float (&a)[M][N], double (&b)[M][N]) {
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) b[i][j] = a[i][j];
}
template <int M, int N> PSM_DEV void cvt(const double (&a)[M][N], float (&b)[M][N]) {
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) b[i][j] = (float)a[i][j];
}
template <int NX, int NY>
PSM_DEV void cvt_step(const Step<float, NX, NY>& a, Step<double, NX, NY>& b) {
  cvt(a.F, b.F); cvt(a.cq, b.cq); cvt(a.bt, b.bt); cvt(a.H, b.H); cvt(a.cr, b.cr); cvt(a.c, b.c);
}
template <int NX>
PSM_DEV void cvt_agg(const FAgg<float, NX>& a, FAgg<double, NX>& b) {
  cvt(a.A, b.A); cvt(a.b, b.b); cvt(a.U, b.U); cvt(a.eta, b.eta); cvt(a.Z, b.Z); b.flag = a.flag;
}
template <int NX>
PSM_DEV void cvt_agg(const FAgg<double, NX>& a, FAgg<float, NX>& b) {
  cvt(a.A, b.A); cvt(a.b, b.b); cvt(a.U, b.U); cvt(a.eta, b.eta); cvt(a.Z, b.Z); b.flag = a.flag;
}
template <int NX>
PSM_DEV void cvt_agg(const SAgg<float, NX>& a, SAgg<double, NX>& b) {
  cvt(a.g, b.g); cvt(a.E, b.E); cvt(a.D, b.D); b.flag = a.flag;
}
template <int NX>
PSM_DEV void cvt_agg(const SAgg<double, NX>& a, SAgg<float, NX>& b) {
  cvt(a.g, b.g); cvt(a.E, b.E); cvt(a.D, b.D); b.flag = a.flag;
}

template <typename R, int NX, int NY, bool SQRT, typename FB>
PSM_DEV typename AccOf<R>::type kf_step_fb(const Step<R, NX, NY>& s, const R (&y)[NY], R (&m)[NX], R (&L)[NX][NX],
                          const FB& fb) {
  if constexpr (kFallback<R, SQRT> && !FB::kRedo) {
    const double e = sqrt_kf_step<float, NX, NY>(s, y, m, L);
    if (fb.thr > 0.f && ill_conditioned(L, fb.thr)) fb.raise();
    return e;
  } else if constexpr (kFallback<R, SQRT>) {
    float m0[NX], L0[NX][NX];
    copy(m, m0);
    copy(L, L0);
    double e = sqrt_kf_step<float, NX, NY>(s, y, m, L);
    if (fb.thr > 0.f && ill_conditioned(L, fb.thr)) {
      Step<double, NX, NY> sd;
      cvt_step(s, sd);
      double yd[NY], md[NX], Ld[NX][NX];
      cvt(y, yd); cvt(m0, md); cvt(L0, Ld);
      e = sqrt_kf_step<double, NX, NY>(sd, yd, md, Ld);
      cvt(md, m);
      cvt(Ld, L);
      count_fallback(fb);
    }
    return e;
  } else {
    return kf_step<R, NX, NY, SQRT>(s, y, m, L);
  }
}

template <typename R, int NX, int NY, bool SQRT, typename FB>
PSM_DEV void absorb_fb(FAgg<R, NX>& a, const Step<R, NX, NY>& s, const R (&y)[NY], const FB& fb) {
  if constexpr (kFallback<R, SQRT> && !FB::kRedo) {
    absorb<float, NX, NY, true>(a, s, y);
    if (fb.thr > 0.f && ill_conditioned(a.U, fb.thr)) fb.raise();
  } else if constexpr (kFallback<R, SQRT>) {
    FAgg<float, NX> a0;
    copy_agg(a, a0);
    absorb<float, NX, NY, true>(a, s, y);
    if (fb.thr > 0.f && ill_conditioned(a.U, fb.thr)) {
      FAgg<double, NX> ad;
      cvt_agg(a0, ad);
      Step<double, NX, NY> sd;
      cvt_step(s, sd);
      double yd[NY];
      cvt(y, yd);
      absorb<double, NX, NY, true>(ad, sd, yd);
      cvt_agg(ad, a);
      count_fallback(fb);
    }
  } else {
    absorb<R, NX, NY, SQRT>(a, s, y);
  }
}

template <typename R, int NX, bool SQRT, typename FB>
PSM_DEV void smoother_element_fb(const R (&F)[NX][NX], const R (&cq)[NX][NX], const R (&bt)[NX],
                                 const R (&m)[NX], const R (&L)[NX][NX], SAgg<R, NX>& out,
                                 const FB& fb) {
  smoother_element<R, NX, SQRT>(F, cq, bt, m, L, out);
  if constexpr (kFallback<R, SQRT> && !FB::kRedo) {
    if (fb.thr > 0.f && ill_conditioned(out.D, fb.thr)) fb.raise();
  } else if constexpr (kFallback<R, SQRT>) {
    if (fb.thr > 0.f && ill_conditioned(out.D, fb.thr)) {
      double Fd[NX][NX], cqd[NX][NX], btd[NX], md[NX], Ld[NX][NX];
      cvt(F, Fd); cvt(cq, cqd); cvt(bt, btd); cvt(m, md); cvt(L, Ld);
      SAgg<double, NX> od;
      sqrt_smoother_element<double, NX>(Fd, cqd, btd, md, Ld, od);
      cvt_agg(od, out);
      count_fallback(fb);
    }
  }
}

template <typename R, int NX, bool SQRT, typename FB>
PSM_DEV void smoother_apply_fb(const SAgg<R, NX>& e, R (&m)[NX], R (&Ds)[NX][NX], const FB& fb) {
  if constexpr (kFallback<R, SQRT> && !FB::kRedo) {
    sqrt_smoother_apply<float, NX>(e, m, Ds);
    if (fb.thr > 0.f && ill_conditioned(Ds, fb.thr)) fb.raise();
  } else if constexpr (kFallback<R, SQRT>) {
    float m0[NX], D0[NX][NX];
    copy(m, m0);
    copy(Ds, D0);
    sqrt_smoother_apply<float, NX>(e, m, Ds);
    if (fb.thr > 0.f && ill_conditioned(Ds, fb.thr)) {
      SAgg<double, NX> ed;
      cvt_agg(e, ed);
      double md[NX], Dd[NX][NX];
      cvt(m0, md); cvt(D0, Dd);
      sqrt_smoother_apply<double, NX>(ed, md, Dd);
      cvt(md, m);
      cvt(Dd, Ds);
      count_fallback(fb);
    }
  } else {
    smoother_apply<R, NX, SQRT>(e, m, Ds);
  }
}

// e2 <- e1 (+) e2 for smoother aggregates held in registers, with the fallback
template <typename R, int NX, bool SQRT, typename FB>
PSM_DEV void combine_into_fb(const SAgg<R, NX>& e1, SAgg<R, NX>& e2, const FB& fb) {
  if constexpr (kFallback<R, SQRT> && !FB::kRedo) {
    const bool plain = e2.flag == FLAG_HEAD || e1.flag == FLAG_IDENT || e2.flag == FLAG_IDENT;
    combine_into<R, NX, SQRT>(e1, e2);
    if (!plain && fb.thr > 0.f && ill_conditioned(e2.D, fb.thr)) fb.raise();
  } else if constexpr (kFallback<R, SQRT>) {
    if (e2.flag == FLAG_HEAD || e1.flag == FLAG_IDENT || e2.flag == FLAG_IDENT) {
      combine_into<R, NX, SQRT>(e1, e2);
      return;
    }
    SAgg<float, NX> b0;
    copy_agg(e2, b0);
    combine_into<R, NX, SQRT>(e1, e2);
    if (fb.thr > 0.f && ill_conditioned(e2.D, fb.thr)) {
      SAgg<double, NX> ad, bd;
      cvt_agg(e1, ad);
      cvt_agg(b0, bd);
      combine_into<double, NX, true>(ad, bd);
      cvt_agg(bd, e2);
      count_fallback(fb);
    }
  } else {
    combine_into<R, NX, SQRT>(e1, e2);
  }
}

}  // namespace psm

