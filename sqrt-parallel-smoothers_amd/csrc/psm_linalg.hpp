// Register-resident small dense linear algebra for the square-root smoother
// kernels (gfx950).  Every matrix is a fixed-size C array held in VGPRs; all
// loops have compile-time trip counts and are fully unrolled, so one lane owns
// one whole problem instance (a time step, a chunk aggregate, ...).
//
// Conventions
//  * Row-major R M[rows][cols].
//  * tria(M) returns the lower-triangular L with L L^T = M M^T and a
//    non-negative diagonal (Householder reflections applied from the right,
//    row by row, with the sign choice that maps a row tail onto +||x|| e_0).
//    This is the operation parsmooth/_utils.py:22-24 computes through a QR of
//    M^T; the reference's diagonal signs come from LAPACK and are arbitrary, the
//    sign choice here makes the factor unique.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PSM_DEV __device__ __forceinline__

namespace psm {

template <typename R> PSM_DEV R rsqrt_(R x);
template <> PSM_DEV double rsqrt_<double>(double x) { return 1.0 / sqrt(x); }
template <> PSM_DEV float rsqrt_<float>(float x) { return 1.0f / sqrtf(x); }
// sqrt for the factorisations' non-negative, finite arguments: the hardware
// reciprocal square-root estimate, one Goldschmidt step for sqrt(x) and
// 1 / (2 sqrt(x)), and a final residual correction; it matched the correctly
// rounded square root on a million random normal values in 1e-30..1e30 (and
// at 0), without the IEEE sequence's rescaling of tiny arguments.
PSM_DEV double sqrt_(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  g = fma(d, h, g);
  return x == 0.0 ? x : g;
}
// 1 / x for the normal, finite divisors of the factorisations: in FP64 the
// hardware estimate (about 2^-24 relative) refined by two Newton steps, which
// on a million random normal values in 1e-30..1e30 gave the correctly rounded
// reciprocal every time, in five instructions instead of the IEEE division's
// special-case sequence; in FP32 the division, which the FP32 build already
// lowers to the hardware reciprocal.
PSM_DEV double rcp_(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}
PSM_DEV float rcp_(float x) { return 1.0f / x; }
PSM_DEV float sqrt_(float x) { return sqrtf(x); }
PSM_DEV double abs_(double x) { return fabs(x); }
PSM_DEV float abs_(float x) { return fabsf(x); }
PSM_DEV double log_(double x) { return log(x); }
PSM_DEV float log_(float x) { return logf(x); }

// ----------------------------------------------------------------------------
// basic products
// ----------------------------------------------------------------------------
template <typename R, int M, int K, int N>
PSM_DEV void mm(const R (&a)[M][K], const R (&b)[K][N], R (&c)[M][N]) {
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      R s = a[i][0] * b[0][j];
#pragma unroll
      for (int k = 1; k < K; ++k) s = fma(a[i][k], b[k][j], s);
      c[i][j] = s;
    }
}

// c = a * b^T
template <typename R, int M, int K, int N>
PSM_DEV void mmT(const R (&a)[M][K], const R (&b)[N][K], R (&c)[M][N]) {
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      R s = a[i][0] * b[j][0];
#pragma unroll
      for (int k = 1; k < K; ++k) s = fma(a[i][k], b[j][k], s);
      c[i][j] = s;
    }
}

// c = a^T * b
template <typename R, int K, int M, int N>
PSM_DEV void mTm(const R (&a)[K][M], const R (&b)[K][N], R (&c)[M][N]) {
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      R s = a[0][i] * b[0][j];
#pragma unroll
      for (int k = 1; k < K; ++k) s = fma(a[k][i], b[k][j], s);
      c[i][j] = s;
    }
}

template <typename R, int M, int N>
PSM_DEV void mv(const R (&a)[M][N], const R (&x)[N], R (&y)[M]) {
#pragma unroll
  for (int i = 0; i < M; ++i) {
    R s = a[i][0] * x[0];
#pragma unroll
    for (int k = 1; k < N; ++k) s = fma(a[i][k], x[k], s);
    y[i] = s;
  }
}

// y = a^T x
template <typename R, int M, int N>
PSM_DEV void mTv(const R (&a)[M][N], const R (&x)[M], R (&y)[N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    R s = a[0][j] * x[0];
#pragma unroll
    for (int k = 1; k < M; ++k) s = fma(a[k][j], x[k], s);
    y[j] = s;
  }
}

template <typename R, int M, int N>
PSM_DEV void set_zero(R (&a)[M][N]) {
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) a[i][j] = R(0);
}

template <typename R, int N>
PSM_DEV void set_zero(R (&a)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) a[i] = R(0);
}

template <typename R, int N>
PSM_DEV void set_eye(R (&a)[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) a[i][j] = (i == j) ? R(1) : R(0);
}

template <typename R, int M, int N>
PSM_DEV void copy(const R (&a)[M][N], R (&b)[M][N]) {
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) b[i][j] = a[i][j];
}

template <typename R, int N>
PSM_DEV void copy(const R (&a)[N], R (&b)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) b[i] = a[i];
}

// ----------------------------------------------------------------------------
// tria: in-place Householder triangularisation from the right.
// On exit M[:, :min(ROWS,COLS)] holds L (lower, non-negative diagonal) and the
// remaining columns are zero.
// ----------------------------------------------------------------------------
template <typename R, int ROWS, int COLS>
PSM_DEV void tria_inplace(R (&M)[ROWS][COLS]) {
  constexpr int NR = ROWS < COLS ? ROWS : COLS;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const R alpha = M[i][i];
    R sigma = R(0);
#pragma unroll
    for (int j = i + 1; j < COLS; ++j) sigma = fma(M[i][j], M[i][j], sigma);
    const R nrm = sqrt_(fma(alpha, alpha, sigma));
    // v = x - ||x|| e0, computed without cancellation (Parlett's formula).
    const R v0 = (alpha > R(0)) ? -sigma * rcp_(alpha + nrm) : alpha - nrm;
    const R vv = fma(v0, v0, sigma);
    const R f = (vv > R(0)) ? R(2) * rcp_(vv) : R(0);
#pragma unroll
    for (int k = i + 1; k < ROWS; ++k) {
      R dot = v0 * M[k][i];
#pragma unroll
      for (int j = i + 1; j < COLS; ++j) dot = fma(M[i][j], M[k][j], dot);
      const R t = f * dot;
      M[k][i] = fma(-t, v0, M[k][i]);
#pragma unroll
      for (int j = i + 1; j < COLS; ++j) M[k][j] = fma(-t, M[i][j], M[k][j]);
    }
    M[i][i] = nrm;
#pragma unroll
    for (int j = i + 1; j < COLS; ++j) M[i][j] = R(0);
  }
}

// A lower-triangular square root of C C^T: C itself when it is lower
// triangular, else tria(C).  The linearized-model layout keeps only the lower
// triangle of a noise factor, and a caller's factor (the extended
// linearization passes chol Q through) may be any square root.
template <typename R, int N>
PSM_DEV void lower_factor(R (&C)[N][N]) {
  bool upper = false;
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = a + 1; b < N; ++b) upper = upper || (C[a][b] != R(0));
  if (upper) tria_inplace<R, N, N>(C);
}

// L (ROWS x ROWS) = tria(M)
template <typename R, int ROWS, int COLS>
PSM_DEV void tria(R (&M)[ROWS][COLS], R (&L)[ROWS][ROWS]) {
  tria_inplace<R, ROWS, COLS>(M);
#pragma unroll
  for (int i = 0; i < ROWS; ++i)
#pragma unroll
    for (int j = 0; j < ROWS; ++j) L[i][j] = (j < COLS && j <= i) ? M[i][j] : R(0);
}

// ----------------------------------------------------------------------------
// triangular solves (L lower triangular, N x N)
// ----------------------------------------------------------------------------
// X = L^{-1} B
template <typename R, int N, int K>
PSM_DEV void solve_lower(const R (&L)[N][N], const R (&B)[N][K], R (&X)[N][K]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const R inv = rcp_(L[i][i]);
#pragma unroll
    for (int c = 0; c < K; ++c) {
      R s = B[i][c];
#pragma unroll
      for (int j = 0; j < i; ++j) s = fma(-L[i][j], X[j][c], s);
      X[i][c] = s * inv;
    }
  }
}

// X = L^{-T} B
template <typename R, int N, int K>
PSM_DEV void solve_lower_T(const R (&L)[N][N], const R (&B)[N][K], R (&X)[N][K]) {
#pragma unroll
  for (int ii = 0; ii < N; ++ii) {
    const int i = N - 1 - ii;
    const R inv = rcp_(L[i][i]);
#pragma unroll
    for (int c = 0; c < K; ++c) {
      R s = B[i][c];
#pragma unroll
      for (int j = i + 1; j < N; ++j) s = fma(-L[j][i], X[j][c], s);
      X[i][c] = s * inv;
    }
  }
}

// x = L^{-1} b
template <typename R, int N>
PSM_DEV void solve_lower_v(const R (&L)[N][N], const R (&b)[N], R (&x)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    R s = b[i];
#pragma unroll
    for (int j = 0; j < i; ++j) s = fma(-L[i][j], x[j], s);
    x[i] = s * rcp_(L[i][i]);
  }
}

// X = B L^{-1}   (B is M x N):  X L = B
template <typename R, int M, int N>
PSM_DEV void solve_right_lower(const R (&L)[N][N], const R (&B)[M][N], R (&X)[M][N]) {
#pragma unroll
  for (int jj = 0; jj < N; ++jj) {
    const int j = N - 1 - jj;
    const R inv = rcp_(L[j][j]);
#pragma unroll
    for (int r = 0; r < M; ++r) {
      R s = B[r][j];
#pragma unroll
      for (int k = j + 1; k < N; ++k) s = fma(-X[r][k], L[k][j], s);
      X[r][j] = s * inv;
    }
  }
}

// X = B L^{-T}   (B is M x N):  X L^T = B
template <typename R, int M, int N>
PSM_DEV void solve_right_lower_T(const R (&L)[N][N], const R (&B)[M][N], R (&X)[M][N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const R inv = rcp_(L[j][j]);
#pragma unroll
    for (int r = 0; r < M; ++r) {
      R s = B[r][j];
#pragma unroll
      for (int k = 0; k < j; ++k) s = fma(-X[r][k], L[j][k], s);
      X[r][j] = s * inv;
    }
  }
}

// ----------------------------------------------------------------------------
// general square solve with partial pivoting: X = A^{-1} B (covariance form)
// ----------------------------------------------------------------------------
template <typename R, int N, int K>
PSM_DEV void solve_general(R (&A)[N][N], R (&B)[N][K]) {
#pragma unroll
  for (int c = 0; c < N; ++c) {
    // pivot search among rows c..N-1 (swap by predicated selects: the row
    // indices stay compile-time constants so nothing leaves the registers)
#pragma unroll
    for (int r = c + 1; r < N; ++r) {
      const bool sw = abs_(A[r][c]) > abs_(A[c][c]);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const R t = A[c][j];
        A[c][j] = sw ? A[r][j] : t;
        A[r][j] = sw ? t : A[r][j];
      }
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const R t = B[c][j];
        B[c][j] = sw ? B[r][j] : t;
        B[r][j] = sw ? t : B[r][j];
      }
    }
    const R inv = rcp_(A[c][c]);
#pragma unroll
    for (int r = c + 1; r < N; ++r) {
      const R l = A[r][c] * inv;
#pragma unroll
      for (int j = c; j < N; ++j) A[r][j] = fma(-l, A[c][j], A[r][j]);
#pragma unroll
      for (int j = 0; j < K; ++j) B[r][j] = fma(-l, B[c][j], B[r][j]);
    }
  }
#pragma unroll
  for (int ii = 0; ii < N; ++ii) {
    const int i = N - 1 - ii;
    const R inv = rcp_(A[i][i]);
#pragma unroll
    for (int j = 0; j < K; ++j) {
      R s = B[i][j];
#pragma unroll
      for (int k = i + 1; k < N; ++k) s = fma(-A[i][k], B[k][j], s);
      B[i][j] = s * inv;
    }
  }
}

// Cholesky factorisation (lower) of a symmetric positive definite matrix.
template <typename R, int N>
PSM_DEV void cholesky(const R (&A)[N][N], R (&L)[N][N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    R s = A[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) s = fma(-L[j][k], L[j][k], s);
    const R d = sqrt_(s);
    L[j][j] = d;
    const R inv = rcp_(d);
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      R t = A[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) t = fma(-L[i][k], L[j][k], t);
      L[i][j] = t * inv;
    }
#pragma unroll
    for (int i = 0; i < j; ++i) L[i][j] = R(0);
  }
}

// Rank-one update / downdate of a lower factor: L L^T + mult v v^T.
// Column recurrence of parsmooth/_utils.py:39-81 (_cholesky_update): entries
// above the diagonal are dropped and non-finite entries are set to zero.
template <typename R, int N>
PSM_DEV void cholesky_update(R (&L)[N][N], const R (&v)[N], R mult) {
  R omega[N];
#pragma unroll
  for (int i = 0; i < N; ++i) omega[i] = v[i];
  R b = R(1);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const R d = L[j][j];
    const R wj = omega[j];
    const R new_d = sqrt_(d * d + mult / b * wj * wj);
    const R gamma = d * d * b + mult * wj * wj;
    const R r = wj / d;
    const R s = mult * wj / gamma;
#pragma unroll
    for (int i = 0; i < N; ++i) omega[i] = fma(-r, L[i][j], omega[i]);
#pragma unroll
    for (int i = j + 1; i < N; ++i) L[i][j] = new_d * (L[i][j] / d + s * omega[i]);
    L[j][j] = new_d;
    b = b + mult * r * r;
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const R x = L[i][j];
      L[i][j] = (j > i || !isfinite(x)) ? R(0) : x;
    }
}

}  // namespace psm
