/*
 * FP64 C restatement of the square-root iterated extended Kalman smoother
 * (IEKS) on the coordinated-turn / bearings-only model and on linear models.
 *
 * TEST INFRASTRUCTURE ONLY: this is the CPU baseline and a second checker.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  It is written independently of the HIP kernels.
 *
 * It follows the reference's sequential square-root filter and smoother
 * (parsmooth/sequential/_filtering.py:80-108, parsmooth/sequential/_smoothing.py:60-77)
 * run at a fixed nominal trajectory, which yields the same filtered and
 * smoothed trajectories as the parallel formulation at that nominal
 * (parsmooth/parallel/_filtering.py:13-61, _smoothing.py:14-57); the
 * iteration is parsmooth/methods.py:54-76 with a fixed iteration count.
 * Models: tests/bearings/bearings_utils.py:7-69 and x -> A x, linearized
 * either by the extended method (parsmooth/linearization/_extended.py:59-70)
 * or by the sigma-point statistical linear regression
 * (parsmooth/linearization/_sigma_points.py:63-100, with the rank-one
 * downdate of parsmooth/_utils.py:13-81) from a caller-supplied table of unit
 * points and weights (cubature: _cubature.py:36-85).
 *
 * oracle_iterated_smoother_batch runs a ragged batch of independent
 * sequences, one OpenMP thread per sequence at a time (the all-core CPU
 * baseline of the batch workloads).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 16
#define MAXPTS 256

/* L (r x r, row-major, lower, non-negative diagonal) with L L^T = M M^T for
 * M r x c (row-major, destroyed).  Householder reflections from the right. */
static void tria(double* M, int r, int c, double* L) {
  int nr = r < c ? r : c;
  for (int i = 0; i < nr; ++i) {
    double alpha = M[i * c + i], sigma = 0.0;
    for (int j = i + 1; j < c; ++j) sigma += M[i * c + j] * M[i * c + j];
    double nrm = sqrt(alpha * alpha + sigma);
    double v0 = alpha > 0 ? -sigma / (alpha + nrm) : alpha - nrm;
    double vv = v0 * v0 + sigma;
    double f = vv > 0 ? 2.0 / vv : 0.0;
    for (int k = i + 1; k < r; ++k) {
      double dot = v0 * M[k * c + i];
      for (int j = i + 1; j < c; ++j) dot += M[i * c + j] * M[k * c + j];
      double t = f * dot;
      M[k * c + i] -= t * v0;
      for (int j = i + 1; j < c; ++j) M[k * c + j] -= t * M[i * c + j];
    }
    M[i * c + i] = nrm;
    for (int j = i + 1; j < c; ++j) M[i * c + j] = 0.0;
  }
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < r; ++j) L[i * r + j] = (j < c && j <= i) ? M[i * c + j] : 0.0;
}

/* ---------------------------------------------------------------- models */
typedef struct {
  int kind;          /* 0 linear, 1 coordinated turn, 2 bearings */
  int nin, nout;
  const double* A;   /* linear: nout x nin */
  double dt;         /* CT */
  double s[4];       /* bearings: s1x s1y s2x s2y */
} model_t;

/* f(x) and Jacobian J (nout x nin) */
static void model_eval(const model_t* md, const double* x, double* y, double* J) {
  int n = md->nin, m = md->nout;
  if (md->kind == 0) {
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) {
        s += md->A[i * n + j] * x[j];
        J[i * n + j] = md->A[i * n + j];
      }
      y[i] = s;
    }
  } else if (md->kind == 1) {
    double dt = md->dt, w = x[4];
    double c = cos(w * dt), s = sin(w * dt);
    int small = fabs(w) < 1e-6;
    double a = small ? dt : s / w;
    double b = small ? 0.0 : (c - 1.0) / w;
    double da = small ? 0.0 : (dt * c - a) / w;
    double db = small ? 0.0 : (-dt * s - b) / w;
    double dc = -dt * s, ds = dt * c;
    y[0] = x[0] + a * x[2] - b * x[3];
    y[1] = x[1] + b * x[2] + a * x[3];
    y[2] = c * x[2] + s * x[3];
    y[3] = -s * x[2] + c * x[3];
    y[4] = x[4];
    memset(J, 0, sizeof(double) * 25);
    J[0] = 1; J[2] = a; J[3] = -b; J[4] = da * x[2] - db * x[3];
    J[6] = 1; J[7] = b; J[8] = a; J[9] = db * x[2] + da * x[3];
    J[12] = c; J[13] = s; J[14] = dc * x[2] + ds * x[3];
    J[17] = -s; J[18] = c; J[19] = -ds * x[2] + dc * x[3];
    J[24] = 1;
  } else {
    memset(J, 0, sizeof(double) * 2 * n);
    for (int r = 0; r < 2; ++r) {
      double dx = x[0] - md->s[2 * r], dy = x[1] - md->s[2 * r + 1];
      double q = dx * dx + dy * dy;
      y[r] = atan2(dy, dx);
      J[r * n + 0] = -dy / q;
      J[r * n + 1] = dx / q;
    }
  }
}

/* Sigma-point table: unit points xi (npts x n, row-major), weights wm, wc;
 * npts == 0 selects the extended linearization. */
typedef struct {
  int npts;
  const double* xi;
  const double* wm;
  const double* wc;
} sigma_t;

/* Rank-one update / downdate L L^T + mult v v^T of a lower factor L (d x d,
 * row-major), the column recurrence of parsmooth/_utils.py:39-81 (omega and b
 * carried over the columns; the diagonal set afterwards, the upper triangle
 * cleared, non-finite entries set to zero). */
static void chol_update(double* L, const double* v, double mult, int d) {
  double omega[MAXN], out[MAXN * MAXN], nd[MAXN];
  double b = 1.0;
  memcpy(omega, v, sizeof(double) * d);
  for (int j = 0; j < d; ++j) {
    double diag = L[j * d + j], om = omega[j];
    nd[j] = sqrt(diag * diag + mult / b * om * om);
    double gamma = diag * diag * b + mult * om * om;
    for (int i = 0; i < d; ++i) omega[i] -= (om / diag) * L[i * d + j];
    for (int i = 0; i < d; ++i) out[i * d + j] = nd[j] * (L[i * d + j] / diag + (mult * om / gamma) * omega[i]);
    b += mult * (om / diag) * (om / diag);
  }
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      double x = j > i ? 0.0 : (j == i ? nd[j] : out[i * d + j]);
      L[i * d + j] = isfinite(x) ? x : 0.0;
    }
}

/* Linearization of md at the Gaussian (m, L): F (nout x nin), noise factor C
 * (nout x nout) and offset b.  Extended: F = Jacobian at m, C = cq,
 * b = f(m) - F m + qm.  Sigma points (parsmooth/linearization/_sigma_points.py:63-100):
 * points m + L xi_p, m_f = sum wm_p f_p, Psi = sum wc_p (chi_p - m)(f_p - m_f)^T,
 * F = (P^-1 Psi)^T with P = L L^T, C = tria([tria(sqrt(wc) (f - m_f)) | cq])
 * downdated by every column of F L, b = m_f - F m + qm. */
static void linearize(const model_t* md, const sigma_t* sg, const double* m, const double* L,
                      const double* qm, const double* cq, double* F, double* C, double* b) {
  int n = md->nin, d = md->nout;
  if (sg == NULL || sg->npts == 0) {
    double y[MAXN];
    model_eval(md, m, y, F);
    for (int i = 0; i < d; ++i) {
      double s = y[i];
      for (int j = 0; j < n; ++j) s -= F[i * n + j] * m[j];
      b[i] = s + qm[i];
    }
    memcpy(C, cq, sizeof(double) * d * d);
    return;
  }
  int np = sg->npts;
  static __thread double fp[MAXPTS * MAXN], chi[MAXPTS * MAXN];
  double J[MAXN * MAXN], mf[MAXN], Psi[MAXN * MAXN];
  memset(mf, 0, sizeof(mf));
  for (int p = 0; p < np; ++p) {
    for (int i = 0; i < n; ++i) {
      double s = m[i];
      for (int j = 0; j < n; ++j) s += L[i * n + j] * sg->xi[p * n + j];
      chi[p * n + i] = s;
    }
    model_eval(md, chi + p * n, fp + p * d, J);
    for (int i = 0; i < d; ++i) mf[i] += sg->wm[p] * fp[p * d + i];
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < d; ++j) {
      double s = 0.0;
      for (int p = 0; p < np; ++p) s += (chi[p * n + i] - m[i]) * sg->wc[p] * (fp[p * d + j] - mf[j]);
      Psi[i * d + j] = s;
    }
  /* F^T = L^-T L^-1 Psi (cho_solve) */
  double X[MAXN * MAXN];
  for (int c = 0; c < d; ++c) {
    for (int i = 0; i < n; ++i) {
      double s = Psi[i * d + c];
      for (int k = 0; k < i; ++k) s -= L[i * n + k] * X[k * d + c];
      X[i * d + c] = s / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = X[i * d + c];
      for (int k = i + 1; k < n; ++k) s -= L[k * n + i] * X[k * d + c];
      X[i * d + c] = s / L[i * n + i];
    }
  }
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < n; ++j) F[i * n + j] = X[j * d + i];
  /* sqrt_Phi: d x np columns sqrt(wc_p)(f_p - m_f), triangularised when np >= d */
  double Phi[MAXN * MAXN];
  memset(Phi, 0, sizeof(Phi));
  {
    static __thread double S[MAXN * MAXPTS];
    for (int i = 0; i < d; ++i)
      for (int p = 0; p < np; ++p) S[i * np + p] = sqrt(sg->wc[p]) * (fp[p * d + i] - mf[i]);
    if (np >= d) {
      tria(S, d, np, Phi);
    } else {
      for (int i = 0; i < d; ++i)
        for (int p = 0; p < np; ++p) Phi[i * d + p] = S[i * np + p];
    }
  }
  double M2[MAXN * 2 * MAXN];
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      M2[i * 2 * d + j] = Phi[i * d + j];
      M2[i * 2 * d + d + j] = cq[i * d + j];
    }
  tria(M2, d, 2 * d, C);
  /* downdate by the columns of F L */
  for (int k = 0; k < n; ++k) {
    double v[MAXN];
    for (int i = 0; i < d; ++i) {
      double s = 0.0;
      for (int q = 0; q < n; ++q) s += F[i * n + q] * L[q * n + k];
      v[i] = s;
    }
    chol_update(C, v, -1.0, d);
  }
  for (int i = 0; i < d; ++i) {
    double s = mf[i];
    for (int j = 0; j < n; ++j) s -= F[i * n + j] * m[j];
    b[i] = s + qm[i];
  }
}

/* ---------------------------------------------------------------- one filter + smoother pass
 * ys (T x ny); x0 (m0, L0); nominal (T+1) states (means nom_m, chols nom_L;
 * nom_L may be NULL for the extended method); trans / obs models with noise
 * (qm, cq), (rm, cr).  Writes filtered (fm, fL) and smoothed (sm, sL); returns
 * the log-likelihood. */
static double filter_smoother(int T, int nx, int ny, const double* ys, const double* m0,
                              const double* L0, const double* nom_m, const double* nom_L,
                              const model_t* tr, const model_t* ob, const sigma_t* sg,
                              const double* qm, const double* cq, const double* rm,
                              const double* cr, double* fm, double* fL, double* sm, double* sL) {
  double F[MAXN * MAXN], bt[MAXN], H[MAXN * MAXN], c[MAXN], CQ[MAXN * MAXN], CR[MAXN * MAXN];
  double M[MAXN * 2 * MAXN], P[(2 * MAXN) * (2 * MAXN)], Lp[MAXN * MAXN], Ltmp[(2 * MAXN) * (2 * MAXN)];
  double eye[MAXN * MAXN];
  double ell = 0.0;
  for (int i = 0; i < nx * nx; ++i) eye[i] = (i % (nx + 1)) == 0 ? 1.0 : 0.0;
#define NOM_L(k) (nom_L ? nom_L + (size_t)(k) * nx * nx : eye)
  memcpy(fm, m0, sizeof(double) * nx);
  memcpy(fL, L0, sizeof(double) * nx * nx);
  for (int k = 0; k < T; ++k) {
    const double* m = fm + k * nx;
    const double* L = fL + k * nx * nx;
    double* mo = fm + (k + 1) * nx;
    double* Lo = fL + (k + 1) * nx * nx;
    linearize(tr, sg, nom_m + k * nx, NOM_L(k), qm, cq, F, CQ, bt);
    linearize(ob, sg, nom_m + (k + 1) * nx, NOM_L(k + 1), rm, cr, H, CR, c);
    /* predict */
    double mp[MAXN];
    for (int i = 0; i < nx; ++i) {
      double s = bt[i];
      for (int j = 0; j < nx; ++j) s += F[i * nx + j] * m[j];
      mp[i] = s;
    }
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < nx; ++j) {
        double s = 0.0;
        for (int q = 0; q < nx; ++q) s += F[i * nx + q] * L[q * nx + j];
        M[i * 2 * nx + j] = s;
        M[i * 2 * nx + nx + j] = CQ[i * nx + j];
      }
    tria(M, nx, 2 * nx, Lp);
    /* update: tria([[H Lp, cr], [Lp, 0]]) */
    int R = ny + nx;
    for (int i = 0; i < ny; ++i) {
      for (int j = 0; j < nx; ++j) {
        double s = 0.0;
        for (int q = 0; q < nx; ++q) s += H[i * nx + q] * Lp[q * nx + j];
        P[i * R + j] = s;
      }
      for (int j = 0; j < ny; ++j) P[i * R + nx + j] = CR[i * ny + j];
    }
    for (int i = 0; i < nx; ++i) {
      for (int j = 0; j < nx; ++j) P[(ny + i) * R + j] = Lp[i * nx + j];
      for (int j = 0; j < ny; ++j) P[(ny + i) * R + nx + j] = 0.0;
    }
    tria(P, R, R, Ltmp);
    double v[MAXN], w[MAXN];
    for (int i = 0; i < ny; ++i) {
      double s = ys[k * ny + i] - c[i];
      for (int q = 0; q < nx; ++q) s -= H[i * nx + q] * mp[q];
      v[i] = s;
    }
    double quad = 0.0, logdet = 0.0;
    for (int i = 0; i < ny; ++i) {
      double s = v[i];
      for (int j = 0; j < i; ++j) s -= Ltmp[i * R + j] * w[j];
      w[i] = s / Ltmp[i * R + i];
      quad += w[i] * w[i];
      logdet += log(fabs(Ltmp[i * R + i]));
    }
    ell += -0.5 * quad - logdet - 0.5 * ny * 1.8378770664093454836 /* log(2 pi) */;
    for (int i = 0; i < nx; ++i) {
      double s = mp[i];
      for (int j = 0; j < ny; ++j) s += Ltmp[(ny + i) * R + j] * w[j];
      mo[i] = s;
      for (int j = 0; j < nx; ++j) Lo[i * nx + j] = Ltmp[(ny + i) * R + ny + j];
    }
  }
  /* smoother */
  memcpy(sm + T * nx, fm + T * nx, sizeof(double) * nx);
  memcpy(sL + T * nx * nx, fL + T * nx * nx, sizeof(double) * nx * nx);
  for (int k = T - 1; k >= 0; --k) {
    const double* m = fm + k * nx;
    const double* L = fL + k * nx * nx;
    linearize(tr, sg, nom_m + k * nx, NOM_L(k), qm, cq, F, CQ, bt);
    int R = 2 * nx;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < nx; ++j) {
        double s = 0.0;
        for (int q = 0; q < nx; ++q) s += F[i * nx + q] * L[q * nx + j];
        P[i * R + j] = s;
        P[i * R + nx + j] = CQ[i * nx + j];
        P[(nx + i) * R + j] = L[i * nx + j];
        P[(nx + i) * R + nx + j] = 0.0;
      }
    tria(P, R, R, Ltmp);
    /* gain G = Phi21 Phi11^{-1} */
    double G[MAXN * MAXN];
    for (int r = 0; r < nx; ++r)
      for (int jj = nx - 1; jj >= 0; --jj) {
        double s = Ltmp[(nx + r) * R + jj];
        for (int q = jj + 1; q < nx; ++q) s -= G[r * nx + q] * Ltmp[q * R + jj];
        G[r * nx + jj] = s / Ltmp[jj * R + jj];
      }
    double pm[MAXN];
    for (int i = 0; i < nx; ++i) {
      double s = bt[i];
      for (int j = 0; j < nx; ++j) s += F[i * nx + j] * m[j];
      pm[i] = s;
    }
    const double* ms = sm + (k + 1) * nx;
    const double* Ls = sL + (k + 1) * nx * nx;
    for (int i = 0; i < nx; ++i) {
      double s = m[i];
      for (int j = 0; j < nx; ++j) s += G[i * nx + j] * (ms[j] - pm[j]);
      sm[k * nx + i] = s;
    }
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < nx; ++j) {
        M[i * 2 * nx + j] = Ltmp[(nx + i) * R + nx + j];
        double s = 0.0;
        for (int q = 0; q < nx; ++q) s += G[i * nx + q] * Ls[q * nx + j];
        M[i * 2 * nx + nx + j] = s;
      }
    tria(M, nx, 2 * nx, sL + k * nx * nx);
  }
#undef NOM_L
  return ell;
}

/* Iterated smoother of one sequence with a fixed iteration count: n_iter
 * filter + smoother passes, each at the previous pass's smoothed trajectory
 * (means, and chols for the sigma-point methods), starting at (init_m,
 * init_L); init_L may be NULL (identity at every state).  Smoothed (T+1)
 * trajectory in out_m / out_L; returns the log-likelihood of one more filter
 * pass at the final nominal when want_ell != 0 (else 0). */
static double iterate_one(int T, int nx, int ny, const double* ys, const double* m0,
                          const double* L0, const double* init_m, const double* init_L,
                          const model_t* tr, const model_t* ob, const sigma_t* sg,
                          const double* qm, const double* cq, const double* rm, const double* cr,
                          int n_iter, int want_ell, double* out_m, double* out_L) {
  size_t n1 = (size_t)(T + 1);
  double* nom = (double*)malloc(sizeof(double) * n1 * nx);
  double* nomL = (double*)malloc(sizeof(double) * n1 * nx * nx);
  double* fm = (double*)malloc(sizeof(double) * n1 * nx);
  double* fL = (double*)malloc(sizeof(double) * n1 * nx * nx);
  memcpy(nom, init_m, sizeof(double) * n1 * nx);
  if (init_L) {
    memcpy(nomL, init_L, sizeof(double) * n1 * nx * nx);
  } else {
    memset(nomL, 0, sizeof(double) * n1 * nx * nx);
    for (size_t k = 0; k < n1; ++k)
      for (int i = 0; i < nx; ++i) nomL[k * nx * nx + i * nx + i] = 1.0;
  }
  if (n_iter == 0) {
    memcpy(out_m, nom, sizeof(double) * n1 * nx);
    memcpy(out_L, nomL, sizeof(double) * n1 * nx * nx);
  }
  for (int it = 0; it < n_iter; ++it) {
    filter_smoother(T, nx, ny, ys, m0, L0, nom, nomL, tr, ob, sg, qm, cq, rm, cr, fm, fL, out_m, out_L);
    memcpy(nom, out_m, sizeof(double) * n1 * nx);
    memcpy(nomL, out_L, sizeof(double) * n1 * nx * nx);
  }
  double ell = 0.0;
  if (want_ell) {
    double* tm = (double*)malloc(sizeof(double) * n1 * nx);
    double* tL = (double*)malloc(sizeof(double) * n1 * nx * nx);
    ell = filter_smoother(T, nx, ny, ys, m0, L0, nom, nomL, tr, ob, sg, qm, cq, rm, cr, fm, fL, tm, tL);
    free(tm);
    free(tL);
  }
  free(nom);
  free(nomL);
  free(fm);
  free(fL);
  return ell;
}

/* One sequence, extended linearization (init chols are identity).
 * kinds: transition / observation model kinds (0 linear, 1 CT, 2 bearings);
 * tr_A / ob_A: linear matrices (or NULL); dt; s (4 sensor coordinates). */
double oracle_iterated_smoother(int T, int nx, int ny, const double* ys, const double* m0,
                                const double* L0, const double* init_m, int tr_kind,
                                const double* tr_A, int ob_kind, const double* ob_A, double dt,
                                const double* s, const double* qm, const double* cq,
                                const double* rm, const double* cr, int n_iter, int want_ell,
                                double* out_m, double* out_L) {
  model_t tr = {tr_kind, nx, nx, tr_A, dt, {0, 0, 0, 0}};
  model_t ob = {ob_kind, nx, ny, ob_A, dt, {s ? s[0] : 0, s ? s[1] : 0, s ? s[2] : 0, s ? s[3] : 0}};
  return iterate_one(T, nx, ny, ys, m0, L0, init_m, NULL, &tr, &ob, NULL, qm, cq, rm, cr, n_iter,
                     want_ell, out_m, out_L);
}

/* Ragged batch: sequence q has lengths[q] observations (ys rows
 * [sum_{r<q} lengths[r], ...)) and owns states [offs[q], offs[q+1]) with
 * offs[q] = sum_{r<q} (lengths[r] + 1) of init_m / init_L / out_m / out_L;
 * priors m0 (n_seq x nx), L0 (n_seq x nx x nx).  npts == 0: extended, else
 * the sigma-point table (xi, wm, wc).  init_L may be NULL.  ell (n_seq) may be
 * NULL.  threads: OpenMP threads (0: the runtime default).  Returns the number
 * of threads used. */
int oracle_iterated_smoother_batch(int n_seq, const int* lengths, int nx, int ny, const double* ys,
                                   const double* m0, const double* L0, const double* init_m,
                                   const double* init_L, int tr_kind, const double* tr_A,
                                   int ob_kind, const double* ob_A, double dt, const double* s,
                                   int npts, const double* xi, const double* wm, const double* wc,
                                   const double* qm, const double* cq, const double* rm,
                                   const double* cr, int n_iter, double* ell, double* out_m,
                                   double* out_L, int threads) {
  model_t tr = {tr_kind, nx, nx, tr_A, dt, {0, 0, 0, 0}};
  model_t ob = {ob_kind, nx, ny, ob_A, dt, {s ? s[0] : 0, s ? s[1] : 0, s ? s[2] : 0, s ? s[3] : 0}};
  sigma_t sg = {npts, xi, wm, wc};
  long* offs = (long*)malloc(sizeof(long) * (n_seq + 1));
  long* yoff = (long*)malloc(sizeof(long) * (n_seq + 1));
  offs[0] = 0;
  yoff[0] = 0;
  for (int q = 0; q < n_seq; ++q) {
    offs[q + 1] = offs[q] + lengths[q] + 1;
    yoff[q + 1] = yoff[q] + lengths[q];
  }
  int used = 1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 1)
    for (int q = 0; q < n_seq; ++q) {
#else
  {
    for (int q = 0; q < n_seq; ++q) {
#endif
      double e = iterate_one(lengths[q], nx, ny, ys + yoff[q] * ny, m0 + (size_t)q * nx,
                             L0 + (size_t)q * nx * nx, init_m + offs[q] * nx,
                             init_L ? init_L + offs[q] * nx * nx : NULL, &tr, &ob,
                             npts > 0 ? &sg : NULL, qm, cq, rm, cr, n_iter, ell != NULL,
                             out_m + offs[q] * nx, out_L + offs[q] * nx * nx);
      if (ell) ell[q] = e;
    }
  }
  free(offs);
  free(yoff);
  return used;
}
