// Team-cooperative square-root scan operators.
//
// The per-lane operators of psm_scan_mem.hpp run one whole combination in one
// lane: ~5,800 dependent instructions for the nx=5 filtering operator, which
// is what bounds a scan level whose few items cannot fill the chip.  Here a
// team of TL lanes (TL = next power of two of the number of matrix rows) runs
// one combination: lane r owns row r of every block it works on, matrix
// products are row-times-matrix in each lane, and the Householder
// triangularisation is row-parallel — the pivot lane forms the reflector from
// its row and publishes it through LDS, the other lanes update their rows.
// The critical path of a combination shrinks by the row count (5-10x), at the
// price of idle lanes, which the scan levels have to spare.
//
// Operands are read from memory views (LDS slots or global scan buffers) and
// the result is written at the end, so `out` must not alias `e1` or `e2`; the
// FP32 fallback relies on that to recompute from intact operands.  All lanes
// of a team live in one wavefront and a team's control flow is uniform, so
// the team synchronises with wavefront-scope fences only (never a workgroup
// barrier): teams of one wave may diverge (combine vs copy) safely.
#pragma once
#include "psm_scan_mem.hpp"

namespace psm {

__host__ __device__ constexpr int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Team size and LDS scratch (in words of the compute type) per operator.
template <typename V> struct TeamOf;
template <typename R, int NX> struct TeamOf<FView<R, NX>> {
  static constexpr int TL = next_pow2(2 * NX);
  // Xi (4 NX^2), two reflectors (2 (2 NX + 1)), M1, P, P A1 (3 NX^2), 6 vectors
  static constexpr int TS = 7 * NX * NX + 10 * NX + 2;
  static constexpr int NROW = NX;  // rows holding the triangular factor U
};
template <typename R, int NX> struct TeamOf<SView<R, NX>> {
  static constexpr int TL = next_pow2(NX);
  static constexpr int TS = 3 * NX + 2;  // |diag| (NX) + reflector (2 NX + 1)
  static constexpr int NROW = NX;
};

// Virtual lanes (teams) per scan workgroup, for the filter and the smoother
// alike (the two scans share one level plan).
constexpr int kTeamVL = 16;

// Team scan kernels in use for the square-root form (float / double).
constexpr bool kTeamScan = false;

PSM_DEV void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename V>
PSM_DEV void team_copy(const V& s, const V& d, int r) {
  constexpr int NQ = V::QF + 1;
  constexpr int TL = TeamOf<V>::TL;
  for (int q = r; q < NQ; q += TL) d.at(q) = s.at(q);
}

template <typename V>
PSM_DEV void team_identity(const V& d, int r);
template <typename R, int NX>
PSM_DEV void team_identity(const FView<R, NX>& d, int r) {
  constexpr int NQ = FView<R, NX>::QF + 1;
  for (int q = r; q < NQ; q += TeamOf<FView<R, NX>>::TL) {
    R v = R(0);
    if (q < NX * NX && (q / NX) == (q % NX)) v = R(1);
    if (q == FView<R, NX>::QF) v = R(FLAG_IDENT);
    d.at(q) = v;
  }
}
template <typename R, int NX>
PSM_DEV void team_identity(const SView<R, NX>& d, int r) {
  constexpr int NQ = SView<R, NX>::QF + 1;
  constexpr int QE = SView<R, NX>::QE;
  for (int q = r; q < NQ; q += TeamOf<SView<R, NX>>::TL) {
    R v = R(0);
    if (q >= QE && q < QE + NX * NX && ((q - QE) / NX) == ((q - QE) % NX)) v = R(1);
    if (q == SView<R, NX>::QF) v = R(FLAG_IDENT);
    d.at(q) = v;
  }
}

// Row-parallel Householder triangularisation from the right of an M x NC
// matrix whose row i is held in x by team lane lane0 + i (the same reflector
// as tria_inplace: non-negative diagonal, Parlett's v0).  hv: LDS broadcast
// area of NC + 1 words for this matrix.  Every team lane calls it.
template <typename C, int M, int NC>
PSM_DEV void team_tria(C (&x)[NC], int r, int lane0, C* hv) {
  constexpr int NR = M < NC ? M : NC;
  const int me = r - lane0;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    if (me == i) {
      const C alpha = x[i];
      C sigma = C(0);
#pragma unroll
      for (int j = i + 1; j < NC; ++j) sigma = fma(x[j], x[j], sigma);
      const C nrm = sqrt_(fma(alpha, alpha, sigma));
      const C v0 = (alpha > C(0)) ? -sigma / (alpha + nrm) : alpha - nrm;
      const C vv = fma(v0, v0, sigma);
      hv[i] = v0;
#pragma unroll
      for (int j = i + 1; j < NC; ++j) hv[j] = x[j];
      hv[NC] = (vv > C(0)) ? C(2) / vv : C(0);
      x[i] = nrm;
#pragma unroll
      for (int j = i + 1; j < NC; ++j) x[j] = C(0);
    }
    team_sync();
    if (me > i && me < M) {
      const C v0 = hv[i];
      C dot = v0 * x[i];
#pragma unroll
      for (int j = i + 1; j < NC; ++j) dot = fma(hv[j], x[j], dot);
      const C t = hv[NC] * dot;
      x[i] = fma(-t, v0, x[i]);
#pragma unroll
      for (int j = i + 1; j < NC; ++j) x[j] = fma(-t, hv[j], x[j]);
    }
    team_sync();
  }
}

// ---------------------------------------------------------------- filtering
// out = e1 (+) e2 (e1 earlier), the blocked form of
// parsmooth/parallel/_operators.py:43-77:
//   Xi = tria([[U1^T Z2, I], [Z2, 0]]),  M1 = U1 Xi11^{-T},  P = I - M1 Xi21^T
//   A = A2 P A1,  b = A2 P (b1 + U1 U1^T eta2) + b2,  U = tria([A2 M1 | U2])
//   eta = A1^T P^T (eta2 - Z2 Z2^T b1) + eta1,  Z = tria([A1^T Xi22 | Z1])
// Lanes 0..2nx-1 own the rows of Xi; afterwards lanes 0..nx-1 build Z and
// eta, lanes nx..2nx-1 build A, b and U (the two final trias run side by side).
template <typename R, typename C, int NX>
PSM_DEV void team_combine_core(const FView<R, NX>& e1, const FView<R, NX>& e2,
                               const FView<R, NX>& out, int r, C* sh) {
  constexpr int n = NX, n2 = 2 * NX;
  C* X = sh;
  C* hvA = X + n2 * n2;
  C* hvB = hvA + n2 + 1;
  C* M1 = hvB + n2 + 1;
  C* P = M1 + n * n;
  C* PA = P + n * n;
  C* t1 = PA + n * n;
  C* t2 = t1 + n;
  C* vv = t2 + n;
  C* ww = vv + n;
  C* PTw = ww + n;
  C* Pv = PTw + n;

  // rows of [[U1^T Z2, I], [Z2, 0]]; t1 = U1^T eta2, t2 = Z2^T b1
  C x[n2];
#pragma unroll
  for (int j = 0; j < n2; ++j) x[j] = C(0);
  if (r < n) {
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) acc = fma((C)e1.U(k, r), (C)e2.Z(k, j), acc);
      x[j] = acc;
      x[n + j] = (r == j) ? C(1) : C(0);
    }
    C a1 = C(0), a2 = C(0);
#pragma unroll
    for (int k = 0; k < n; ++k) {
      a1 = fma((C)e1.U(k, r), (C)e2.eta(k), a1);
      a2 = fma((C)e2.Z(k, r), (C)e1.b(k), a2);
    }
    t1[r] = a1;
    t2[r] = a2;
  } else if (r < n2) {
#pragma unroll
    for (int j = 0; j < n; ++j) x[j] = (C)e2.Z(r - n, j);
  }
  team_tria<C, n2, n2>(x, r, 0, hvA);
  if (r < n2) {
#pragma unroll
    for (int c = 0; c < n2; ++c) X[r * n2 + c] = x[c];
  }
  team_sync();
  // v = b1 + U1 t1, w = eta2 - Z2 t2, M1 = U1 Xi11^{-T}, P = I - M1 Xi21^T
  if (r < n) {
    C a1 = (C)e1.b(r), a2 = (C)e2.eta(r);
#pragma unroll
    for (int k = 0; k < n; ++k) {
      a1 = fma((C)e1.U(r, k), t1[k], a1);
      a2 = fma(-(C)e2.Z(r, k), t2[k], a2);
    }
    vv[r] = a1;
    ww[r] = a2;
    C m1[n];
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C s = (C)e1.U(r, j);
#pragma unroll
      for (int k = 0; k < j; ++k) s = fma(-m1[k], X[j * n2 + k], s);
      m1[j] = s / X[j * n2 + j];
      M1[r * n + j] = m1[j];
    }
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = (r == j) ? C(1) : C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) acc = fma(-m1[k], X[(n + j) * n2 + k], acc);
      P[r * n + j] = acc;
    }
  }
  team_sync();
  // P^T w, P v, P A1
  if (r < n) {
    C a = C(0), b = C(0);
#pragma unroll
    for (int k = 0; k < n; ++k) {
      a = fma(P[k * n + r], ww[k], a);
      b = fma(P[r * n + k], vv[k], b);
    }
    PTw[r] = a;
    Pv[r] = b;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) acc = fma(P[r * n + k], (C)e1.A(k, j), acc);
      PA[r * n + j] = acc;
    }
  }
  team_sync();
  // lanes 0..n-1: eta and the rows of [A1^T Xi22 | Z1]
  // lanes n..2n-1: A, b and the rows of [A2 M1 | U2]
  C y[n2];
#pragma unroll
  for (int j = 0; j < n2; ++j) y[j] = C(0);
  C e_or_b = C(0);
  C arow[n];
#pragma unroll
  for (int j = 0; j < n; ++j) arow[j] = C(0);
  if (r < n) {
    C a = (C)e1.eta(r);
#pragma unroll
    for (int k = 0; k < n; ++k) a = fma((C)e1.A(k, r), PTw[k], a);
    e_or_b = a;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = j; k < n; ++k) acc = fma((C)e1.A(k, r), X[(n + k) * n2 + n + j], acc);
      y[j] = acc;
      y[n + j] = (C)e1.Z(r, j);
    }
  } else if (r < n2) {
    const int i = r - n;
    C a2[n];
#pragma unroll
    for (int k = 0; k < n; ++k) a2[k] = (C)e2.A(i, k);
    C bb = (C)e2.b(i);
#pragma unroll
    for (int k = 0; k < n; ++k) bb = fma(a2[k], Pv[k], bb);
    e_or_b = bb;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C s1 = C(0), s2 = C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) {
        s1 = fma(a2[k], M1[k * n + j], s1);
        s2 = fma(a2[k], PA[k * n + j], s2);
      }
      y[j] = s1;
      arow[j] = s2;
      y[n + j] = (C)e2.U(i, j);
    }
  }
  team_tria<C, n, n2>(y, r, r < n ? 0 : n, r < n ? hvA : hvB);
  if (r < n) {
    out.eta(r) = (R)e_or_b;
#pragma unroll
    for (int j = 0; j < n; ++j) out.Z(r, j) = (R)y[j];
  } else if (r < n2) {
    const int i = r - n;
    out.b(i) = (R)e_or_b;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      out.A(i, j) = (R)arow[j];
      out.U(i, j) = (R)y[j];
    }
  }
  // |diag U| for the fallback check (lanes n..2n-1 hold the rows of U)
  if (r >= n && r < n2) sh[r - n] = abs_(y[r - n]);
}

// ---------------------------------------------------------------- smoothing
// out = e1 (+) e2 with e1 the later aggregate
// (parsmooth/parallel/_operators.py:104-125): g = E2 g1 + g2, E = E2 E1,
// D = tria([E2 D1 | D2]); lane r < nx owns row r.
template <typename R, typename C, int NX>
PSM_DEV void team_combine_core(const SView<R, NX>& e1, const SView<R, NX>& e2,
                               const SView<R, NX>& out, int r, C* sh) {
  constexpr int n = NX, n2 = 2 * NX;
  C* hv = sh + n;
  C y[n2];
#pragma unroll
  for (int j = 0; j < n2; ++j) y[j] = C(0);
  C g = C(0);
  C erow[n];
#pragma unroll
  for (int j = 0; j < n; ++j) erow[j] = C(0);
  if (r < n) {
    C e2r[n];
#pragma unroll
    for (int k = 0; k < n; ++k) e2r[k] = (C)e2.E(r, k);
    g = (C)e2.g(r);
#pragma unroll
    for (int k = 0; k < n; ++k) g = fma(e2r[k], (C)e1.g(k), g);
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = C(0), acd = C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) acc = fma(e2r[k], (C)e1.E(k, j), acc);
#pragma unroll
      for (int k = j; k < n; ++k) acd = fma(e2r[k], (C)e1.D(k, j), acd);
      erow[j] = acc;
      y[j] = acd;
      y[n + j] = (C)e2.D(r, j);
    }
  }
  team_tria<C, n, n2>(y, r, 0, hv);
  if (r < n) {
    out.g(r) = (R)g;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      out.E(r, j) = (R)erow[j];
      out.D(r, j) = (R)y[j];
    }
    sh[r] = abs_(y[r]);
  }
}

// Scratch of one team in bytes (room for the FP64 recomputation of an FP32
// operator).
template <typename V, typename R>
constexpr int team_scratch_bytes() {
  return TeamOf<V>::TS * (int)(sizeof(R) < sizeof(double) ? sizeof(double) : sizeof(R));
}

// One combination by a team: flag handling of combine_mem, the cooperative
// operator, and for FP32 the recomputation in FP64 when the combined factor's
// min|diag| / max|diag| falls below the threshold (or is not finite).
template <typename R, int NX, typename V>
PSM_DEV void team_combine(const V& e1, const V& e2, const V& out, int r, void* scratch,
                          const Fallback& fb) {
  team_sync();
  const int f1 = e1.flag(), f2 = e2.flag();
  if (f2 == FLAG_HEAD || f1 == FLAG_IDENT) {
    team_copy(e2, out, r);
  } else if (f2 == FLAG_IDENT) {
    team_copy(e1, out, r);
  } else {
    R* sh = static_cast<R*>(scratch);
    team_combine_core<R, R, NX>(e1, e2, out, r, sh);
    if constexpr (kFallback<R, true>) {
      if (fb.thr > 0.f) {
        team_sync();
        R mn = sh[0], mx = sh[0];
#pragma unroll
        for (int i = 1; i < TeamOf<V>::NROW; ++i) {
          mn = sh[i] < mn ? sh[i] : mn;
          mx = sh[i] > mx ? sh[i] : mx;
        }
        team_sync();
        if (!(mn >= R(fb.thr) * mx)) {
          team_combine_core<R, double, NX>(e1, e2, out, r, static_cast<double*>(scratch));
          if (r == 0) count_fallback(fb);
        }
      }
    }
    if (r == 0) out.set_flag(f1);
  }
  team_sync();
}

}  // namespace psm
