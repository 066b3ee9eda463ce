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
 * Models: tests/bearings/bearings_utils.py:7-69 (extended linearization,
 * parsmooth/linearization/_extended.py:59-70) and x -> A x.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXN 16

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

/* extended linearization at m: F, b = f(m) - F m + qm */
static void linearize(const model_t* md, const double* m, const double* qm, double* F, double* b) {
  double y[MAXN];
  model_eval(md, m, y, F);
  for (int i = 0; i < md->nout; ++i) {
    double s = y[i];
    for (int j = 0; j < md->nin; ++j) s -= F[i * md->nin + j] * m[j];
    b[i] = s + qm[i];
  }
}

/* ---------------------------------------------------------------- one filter + smoother pass
 * ys (T x ny); x0 (m0, L0); nominal (T+1) states; trans / obs models with
 * noise (qm, cq), (rm, cr).  Writes filtered (fm, fL) and smoothed (sm, sL);
 * returns the log-likelihood. */
static double filter_smoother(int T, int nx, int ny, const double* ys, const double* m0,
                              const double* L0, const double* nom_m, const model_t* tr,
                              const double* qm, const double* cq, const model_t* ob,
                              const double* rm, const double* cr, double* fm, double* fL,
                              double* sm, double* sL) {
  double F[MAXN * MAXN], bt[MAXN], H[MAXN * MAXN], c[MAXN];
  double M[MAXN * 2 * MAXN], P[(2 * MAXN) * (2 * MAXN)], Lp[MAXN * MAXN], Ltmp[(2 * MAXN) * (2 * MAXN)];
  double ell = 0.0;
  memcpy(fm, m0, sizeof(double) * nx);
  memcpy(fL, L0, sizeof(double) * nx * nx);
  for (int k = 0; k < T; ++k) {
    const double* m = fm + k * nx;
    const double* L = fL + k * nx * nx;
    double* mo = fm + (k + 1) * nx;
    double* Lo = fL + (k + 1) * nx * nx;
    linearize(tr, nom_m + k * nx, qm, F, bt);
    linearize(ob, nom_m + (k + 1) * nx, rm, H, c);
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
        M[i * 2 * nx + nx + j] = cq[i * nx + j];
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
      for (int j = 0; j < ny; ++j) P[i * R + nx + j] = cr[i * ny + j];
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
    linearize(tr, nom_m + k * nx, qm, F, bt);
    int R = 2 * nx;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < nx; ++j) {
        double s = 0.0;
        for (int q = 0; q < nx; ++q) s += F[i * nx + q] * L[q * nx + j];
        P[i * R + j] = s;
        P[i * R + nx + j] = cq[i * nx + j];
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
  return ell;
}

/* Iterated smoother with a fixed iteration count on one sequence.
 * kinds: transition / observation model kinds (0 linear, 1 CT, 2 bearings);
 * tr_A / ob_A: linear matrices (or NULL); dt; s (4 sensor coordinates).
 * init_m: (T+1) x nx initial nominal means.  Outputs smoothed (T+1) trajectory
 * in out_m / out_L (chol); returns the log-likelihood of the last filter pass
 * at the final nominal when want_ell != 0 (else 0). */
double oracle_iterated_smoother(int T, int nx, int ny, const double* ys, const double* m0,
                                const double* L0, const double* init_m, int tr_kind,
                                const double* tr_A, int ob_kind, const double* ob_A, double dt,
                                const double* s, const double* qm, const double* cq,
                                const double* rm, const double* cr, int n_iter, int want_ell,
                                double* out_m, double* out_L) {
  model_t tr = {tr_kind, nx, nx, tr_A, dt, {0, 0, 0, 0}};
  model_t ob = {ob_kind, nx, ny, ob_A, dt, {s ? s[0] : 0, s ? s[1] : 0, s ? s[2] : 0, s ? s[3] : 0}};
  size_t n1 = (size_t)(T + 1);
  double* nom = (double*)malloc(sizeof(double) * n1 * nx);
  double* fm = (double*)malloc(sizeof(double) * n1 * nx);
  double* fL = (double*)malloc(sizeof(double) * n1 * nx * nx);
  memcpy(nom, init_m, sizeof(double) * n1 * nx);
  for (int it = 0; it < n_iter; ++it) {
    filter_smoother(T, nx, ny, ys, m0, L0, nom, &tr, qm, cq, &ob, rm, cr, fm, fL, out_m, out_L);
    memcpy(nom, out_m, sizeof(double) * n1 * nx);
  }
  double ell = 0.0;
  if (want_ell) {
    double* tm = (double*)malloc(sizeof(double) * n1 * nx);
    double* tL = (double*)malloc(sizeof(double) * n1 * nx * nx);
    ell = filter_smoother(T, nx, ny, ys, m0, L0, nom, &tr, qm, cq, &ob, rm, cr, fm, fL, tm, tL);
    free(tm);
    free(tL);
  }
  free(nom);
  free(fm);
  free(fL);
  return ell;
}
