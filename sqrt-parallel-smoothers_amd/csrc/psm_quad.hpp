// Quad-cooperative square-root scan operators for level 0 of the scans.
//
// Level 0 of a scan holds one aggregate per chunk: many items, but not enough
// for one lane per combination to give every SIMD more than one wavefront,
// and each lane then runs a chain of ~6,000 dependent instructions per
// filtering combination (nx = 5).  Here four lanes -- one DPP quad -- run one
// combination.  An M-row matrix being built or triangularised is dealt out
// over the quad: row 4 s + r sits in slot s of lane r for the D = M / 4 whole
// quads of rows, and the RP = M % 4 rows left over are replicated on all four
// lanes (every lane holds them and applies the same updates, so they need no
// broadcast).  A Householder reflector of a dealt-out pivot row is broadcast
// with DPP quad_perm (a register move); every lane then reflects all of its
// rows, up to three independent row updates in flight per lane.
//
// The steps mirror the per-lane operators (psm_scan_mem.hpp, tria_inplace in
// psm_linalg.hpp) -- the same reflector, the same fma chains in the same
// order -- except that A and b are formed as (A2 P) A1 and (A2 P) v rather
// than A2 (P A1) and A2 (P v), which saves broadcasting P A1.  The code is branch-free between the operand reads
// and the stores: a DPP broadcast reads the source lane's register, so the
// four lanes of a quad execute every broadcast together.  Quads of one
// wavefront may take different paths, since quad_perm never crosses quads.
// `out` must not alias `e1` or `e2` (the FP32 recomputation in FP64 rereads
// the operands).
#pragma once
#include <cstdint>
#include <type_traits>

#include "psm_team.hpp"

namespace psm {

constexpr int kQuadLanes = 4;

// Keeps the scheduler from moving instructions across a phase boundary of a
// combination: it would otherwise hoist operand loads and broadcasts of later
// phases far ahead and run out of registers.
#define PSM_QUAD_PHASE() __builtin_amdgcn_sched_barrier(0)

// Value of lane L of this lane's quad (DPP quad_perm [L, L, L, L]).
template <int L>
PSM_DEV int qbcast_i32(int v) {
  static_assert(L >= 0 && L < 4, "quad lane");
  // every lane of the quad is active and written, so no "old" value is needed
  return __builtin_amdgcn_mov_dpp(v, L * 0x55, 0xF, 0xF, false);
}
template <int L>
PSM_DEV float qb(float v) {
  return __builtin_bit_cast(float, qbcast_i32<L>(__builtin_bit_cast(int, v)));
}
template <int L>
PSM_DEV double qb(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)qbcast_i32<L>((int)(uint32_t)u);
  const uint32_t hi = (uint32_t)qbcast_i32<L>((int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Row layout of an M-row matrix over a quad: slot s < D holds row 4 s + r on
// lane r; slot D + e (e < RP) holds row 4 D + e on every lane.
template <int M>
struct QRows {
  static constexpr int D = M / 4;
  static constexpr int RP = M - 4 * D;
  static constexpr int NS = D + RP;
  // row held in slot S by lane r
  template <int S>
  static PSM_DEV int row(int r) {
    if constexpr (S < D) return 4 * S + r;
    else return 4 * D + (S - D);
  }
  // smallest and largest row of slot S over the quad
  template <int S> static constexpr int lo() { return S < D ? 4 * S : 4 * D + (S - D); }
  template <int S> static constexpr int hi() { return S < D ? 4 * S + 3 : 4 * D + (S - D); }
};

// Entry (K, J) of a matrix in quad layout, on every lane of the quad.
template <int M, int K, int J, typename C, int NS, int NC>
PSM_DEV C qget(const C (&x)[NS][NC]) {
  using L = QRows<M>;
  static_assert(K < M && J < NC && NS == L::NS, "quad entry");
  if constexpr (K < 4 * L::D) return qb<K % 4>(x[K / 4][J]);
  else return x[L::D + (K - 4 * L::D)][J];
}

// Householder triangularisation from the right of an M x NC matrix held in
// quad layout: the reflector of tria_inplace (non-negative diagonal, Parlett's
// v0), pivot by pivot.  Every lane forms the reflector of its own row in the
// pivot's slot; the pivot lane's is broadcast.  Unlike tria_inplace, the
// entries right of the diagonal are left as they were rather than zeroed (a
// row above the pivot is never updated, so they are never read again): the
// caller reads the lower triangle only.
template <typename C, int M, int NC>
PSM_DEV void quad_tria(C (&x)[QRows<M>::NS][NC], int r) {
  using L = QRows<M>;
  constexpr int D = L::D, NS = L::NS;
  constexpr int NR = M < NC ? M : NC;
  static_for<0, NR>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr bool dist = i < 4 * D;
    constexpr int ps = dist ? i / 4 : D + (i - 4 * D);
    const C alpha = x[ps][i];
    C sigma = C(0);
#pragma unroll
    for (int j = i + 1; j < NC; ++j) sigma = fma(x[ps][j], x[ps][j], sigma);
    const C nrm = sqrt_(fma(alpha, alpha, sigma));
    const C v0 = (alpha > C(0)) ? -sigma * rcp_(alpha + nrm) : alpha - nrm;
    const C vv = fma(v0, v0, sigma);
    const C f = (vv > C(0)) ? C(2) * rcp_(vv) : C(0);
    C bv0 = v0, bf = f;
    C bx[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) bx[j] = C(0);
    if constexpr (dist) {
      bv0 = qb<i % 4>(v0);
      bf = qb<i % 4>(f);
      static_for<i + 1, NC>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        bx[j] = qb<i % 4>(x[ps][j]);
      });
    } else {
#pragma unroll
      for (int j = i + 1; j < NC; ++j) bx[j] = x[ps][j];
    }
    static_for<0, NS>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      constexpr int lo = L::template lo<s>(), hi = L::template hi<s>();
      if constexpr (lo > i) {  // every row of the slot lies below the pivot
        C dot = bv0 * x[s][i];
#pragma unroll
        for (int j = i + 1; j < NC; ++j) dot = fma(bx[j], x[s][j], dot);
        const C t = bf * dot;
        x[s][i] = fma(-t, bv0, x[s][i]);
#pragma unroll
        for (int j = i + 1; j < NC; ++j) x[s][j] = fma(-t, bx[j], x[s][j]);
      } else if constexpr (hi >= i && s >= D) {  // the replicated pivot row
        x[s][i] = nrm;
      } else if constexpr (hi >= i) {  // the quad of rows holding the pivot
        const int row = L::template row<s>(r);
        C dot = bv0 * x[s][i];
#pragma unroll
        for (int j = i + 1; j < NC; ++j) dot = fma(bx[j], x[s][j], dot);
        // rows at or above the pivot take a zero update
        const C t = (row > i) ? bf * dot : C(0);
        x[s][i] = (row == i) ? nrm : fma(-t, bv0, x[s][i]);
#pragma unroll
        for (int j = i + 1; j < NC; ++j) x[s][j] = fma(-t, bx[j], x[s][j]);
      }
    });
  });
}

template <typename V>
PSM_DEV void quad_copy(const V& s, const V& d, int r) {
  constexpr int NQ = V::QF + 1;
  for (int q = r; q < NQ; q += kQuadLanes) d.at(q) = s.at(q);
}

template <typename R, int NX>
PSM_DEV void quad_identity(const FView<R, NX>& d, int r) {
  constexpr int NQ = FView<R, NX>::QF + 1;
  for (int q = r; q < NQ; q += kQuadLanes) {
    R v = R(0);
    if (q < NX * NX && (q / NX) == (q % NX)) v = R(1);
    if (q == FView<R, NX>::QF) v = R(FLAG_IDENT);
    d.at(q) = v;
  }
}
template <typename R, int NX>
PSM_DEV void quad_identity(const SView<R, NX>& d, int r) {
  constexpr int NQ = SView<R, NX>::QF + 1;
  constexpr int QE = SView<R, NX>::QE;
  for (int q = r; q < NQ; q += kQuadLanes) {
    R v = R(0);
    if (q >= QE && q < QE + NX * NX && ((q - QE) / NX) == ((q - QE) % NX)) v = R(1);
    if (q == SView<R, NX>::QF) v = R(FLAG_IDENT);
    d.at(q) = v;
  }
}

// true when min|diag| < thr * max|diag| of the NX x NX factor y (the columns
// 0..NX-1 of a matrix in quad layout) or a diagonal entry is not finite: the
// rule of ill_conditioned, evaluated identically on every lane.
template <typename C, int NX, int NC>
PSM_DEV bool quad_ill(const C (&y)[QRows<NX>::NS][NC], float thr) {
  C mn = C(0), mx = C(0);
  static_for<0, NX>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const C d = abs_(qget<NX, i, i>(y));
    mn = (i == 0 || d < mn) ? d : mn;
    mx = (i == 0 || d > mx) ? d : mx;
  });
  return !(mn >= C(thr) * mx);
}

// ---------------------------------------------------------------- filtering
// out = e1 (+) e2 (e1 earlier): sqrt_combine_mem in quad layout.
//   Xi = tria([[U1^T Z2, I], [Z2, 0]]),  M1 = U1 Xi11^{-T},  P = I - M1 Xi21^T
//   v = b1 + U1 U1^T eta2,  w = eta2 - Z2 Z2^T b1
//   eta = A1^T P^T w + eta1,  Z = tria([A1^T Xi22 | Z1])
//   A = (A2 P) A1,  b = (A2 P) v + b2,  U = tria([A2 M1 | U2])
// The 2nx rows of Xi and the nx rows of every nx-row result are in quad
// layout; the nx-vectors t1, t2, v, w and P^T w are formed on every lane.
// Returns (when ILL) whether U is ill-conditioned at thr.
template <typename R, typename C, int NX, bool ILL>
PSM_DEV bool quad_combine_core(const FView<R, NX>& e1, const FView<R, NX>& e2,
                               const FView<R, NX>& out, int r, float thr) {
  constexpr int n = NX, n2 = 2 * NX;
  using LX = QRows<n2>;
  using LN = QRows<n>;
  constexpr int SX = LX::NS, SN = LN::NS;

  // rows of [[U1^T Z2, I], [Z2, 0]]
  C x[SX][n2];
  static_for<0, SX>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int k = LX::template row<s>(r);
    constexpr bool all_top = LX::template hi<s>() < n, all_bot = LX::template lo<s>() >= n;
    const bool top = k < n;
    const int kk = top ? k : k - n;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = C(0);
      if constexpr (!all_bot) {
#pragma unroll
        for (int m = 0; m < n; ++m) acc = fma((C)e1.U(m, kk), (C)e2.Z(m, j), acc);
      }
      if constexpr (all_top) {
        x[s][j] = acc;
      } else if constexpr (all_bot) {
        x[s][j] = (C)e2.Z(kk, j);
      } else {
        const C z = (C)e2.Z(kk, j);
        x[s][j] = top ? acc : z;
      }
      x[s][n + j] = (top && kk == j) ? C(1) : C(0);
    }
  });
  quad_tria<C, n2, n2>(x, r);
  PSM_QUAD_PHASE();

  // Xi11 (lower), Xi21 and Xi22 (lower) on every lane
  C xi11[n][n], xi21[n][n], xi22[n][n];
  static_for<0, n>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    static_for<0, n>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if constexpr (k <= j) {
        xi11[j][k] = qget<n2, j, k>(x);
        xi22[j][k] = qget<n2, n + j, n + k>(x);
      } else {
        xi11[j][k] = C(0);
        xi22[j][k] = C(0);
      }
      xi21[j][k] = qget<n2, n + j, k>(x);
    });
  });

  // rows of M1 = U1 Xi11^{-T} and of P = I - M1 Xi21^T
  C m1[SN][n], p[SN][n];
  static_for<0, n>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    const C inv = rcp_(xi11[j][j]);
    static_for<0, SN>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      const int i = LN::template row<s>(r);
      C acc = (C)e1.U(i, j);
#pragma unroll
      for (int k = 0; k < j; ++k) acc = fma(-m1[s][k], xi11[j][k], acc);
      m1[s][j] = acc * inv;
    });
  });
  static_for<0, SN>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int i = LN::template row<s>(r);
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = (i == j) ? C(1) : C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) acc = fma(-m1[s][k], xi21[j][k], acc);
      p[s][j] = acc;
    }
  });
  PSM_QUAD_PHASE();

  // v = b1 + U1 (U1^T eta2), w = eta2 - Z2 (Z2^T b1) on every lane
  C v[n], w[n];
  {
    C t1[n], t2[n];
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C a1 = C(0), a2 = C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) {
        a1 = fma((C)e1.U(k, j), (C)e2.eta(k), a1);
        a2 = fma((C)e2.Z(k, j), (C)e1.b(k), a2);
      }
      t1[j] = a1;
      t2[j] = a2;
    }
#pragma unroll
    for (int i = 0; i < n; ++i) {
      C a1 = (C)e1.b(i), a2 = (C)e2.eta(i);
#pragma unroll
      for (int k = 0; k < n; ++k) {
        a1 = fma((C)e1.U(i, k), t1[k], a1);
        a2 = fma(-(C)e2.Z(i, k), t2[k], a2);
      }
      v[i] = a1;
      w[i] = a2;
    }
  }
  PSM_QUAD_PHASE();

  // P on every lane; P^T w; eta rows
  C pf[n][n];
  static_for<0, n>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    static_for<0, n>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      pf[k][j] = qget<n, k, j>(p);
    });
  });
  C ptw[n];
#pragma unroll
  for (int j = 0; j < n; ++j) {
    C a = C(0);
#pragma unroll
    for (int k = 0; k < n; ++k) a = fma(pf[k][j], w[k], a);
    ptw[j] = a;
  }
  // stores: dealt-out rows by their lanes, replicated rows by lane 0
  static_for<0, SN>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int i = LN::template row<s>(r);
    C a = (C)e1.eta(i);
#pragma unroll
    for (int k = 0; k < n; ++k) a = fma((C)e1.A(k, i), ptw[k], a);
    if (s < LN::D || r == 0) out.eta(i) = (R)a;
  });
  PSM_QUAD_PHASE();

  // Z = tria([A1^T Xi22 | Z1])
  C zc[SN][n2];
  static_for<0, SN>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int i = LN::template row<s>(r);
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = j; k < n; ++k) acc = fma((C)e1.A(k, i), xi22[k][j], acc);
      zc[s][j] = acc;
      zc[s][n + j] = (C)e1.Z(i, j);
    }
  });
  quad_tria<C, n, n2>(zc, r);
  static_for<0, SN>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int i = LN::template row<s>(r);
    if (s < LN::D || r == 0) {
#pragma unroll
      for (int j = 0; j < n; ++j) out.Z(i, j) = (j <= i) ? (R)zc[s][j] : R(0);
    }
  });
  PSM_QUAD_PHASE();

  // M1 on every lane
  C m1f[n][n];
  static_for<0, n>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    static_for<0, n>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      m1f[k][j] = qget<n, k, j>(m1);
    });
  });
  PSM_QUAD_PHASE();

  // u = A2(i, :) P; A(i, :) = u A1, b(i) = b2(i) + u v, S(i, :) = A2(i, :) M1;
  // U = tria([S | U2])
  C uc[SN][n2];
  static_for<0, SN>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int i = LN::template row<s>(r);
    const bool mine = s < LN::D || r == 0;
    C a2[n], u[n];
#pragma unroll
    for (int k = 0; k < n; ++k) a2[k] = (C)e2.A(i, k);
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C s1 = C(0), s2 = C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) {
        s1 = fma(a2[k], m1f[k][j], s1);
        s2 = fma(a2[k], pf[k][j], s2);
      }
      uc[s][j] = s1;
      u[j] = s2;
      uc[s][n + j] = (C)e2.U(i, j);
    }
    C b = (C)e2.b(i);
#pragma unroll
    for (int k = 0; k < n; ++k) b = fma(u[k], v[k], b);
    if (mine) out.b(i) = (R)b;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) acc = fma(u[k], (C)e1.A(k, j), acc);
      if (mine) out.A(i, j) = (R)acc;
    }
  });
  PSM_QUAD_PHASE();
  quad_tria<C, n, n2>(uc, r);

  bool ill = false;
  if constexpr (ILL) ill = quad_ill<C, n, n2>(uc, thr);
  static_for<0, SN>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int i = LN::template row<s>(r);
    if (s < LN::D || r == 0) {
#pragma unroll
      for (int j = 0; j < n; ++j) out.U(i, j) = (j <= i) ? (R)uc[s][j] : R(0);
    }
  });
  return ill;
}

// ---------------------------------------------------------------- smoothing
// out = e1 (+) e2 with e1 the later aggregate: sqrt_smoother_combine_mem in
// quad layout (g = E2 g1 + g2, E = E2 E1, D = tria([E2 D1 | D2])).
template <typename R, typename C, int NX, bool ILL>
PSM_DEV bool quad_combine_core(const SView<R, NX>& e1, const SView<R, NX>& e2,
                               const SView<R, NX>& out, int r, float thr) {
  constexpr int n = NX, n2 = 2 * NX;
  using LN = QRows<n>;
  constexpr int SN = LN::NS;
  C y[SN][n2], g[SN], erow[SN][n];
  static_for<0, SN>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int i = LN::template row<s>(r);
    C e2r[n];
#pragma unroll
    for (int k = 0; k < n; ++k) e2r[k] = (C)e2.E(i, k);
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = j; k < n; ++k) acc = fma(e2r[k], (C)e1.D(k, j), acc);
      y[s][j] = acc;
      y[s][n + j] = (C)e2.D(i, j);
    }
    C gg = (C)e2.g(i);
#pragma unroll
    for (int k = 0; k < n; ++k) gg = fma(e2r[k], (C)e1.g(k), gg);
    g[s] = gg;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = 0; k < n; ++k) acc = fma(e2r[k], (C)e1.E(k, j), acc);
      erow[s][j] = acc;
    }
  });
  quad_tria<C, n, n2>(y, r);
  bool ill = false;
  if constexpr (ILL) ill = quad_ill<C, n, n2>(y, thr);
  static_for<0, SN>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    const int i = LN::template row<s>(r);
    if (s < LN::D || r == 0) {
      out.g(i) = (R)g[s];
#pragma unroll
      for (int j = 0; j < n; ++j) {
        out.E(i, j) = (R)erow[s][j];
        out.D(i, j) = (j <= i) ? (R)y[s][j] : R(0);
      }
    }
  });
  return ill;
}

// One combination by a quad: the flag cases of combine_mem, the cooperative
// operator, and for FP32 the recomputation in FP64 when the combined factor is
// ill-conditioned (the rule of combine_mem_fb).  The flags and the decision
// are the same on the four lanes, so the quad stays together.
template <typename R, int NX, typename V, typename FB>
PSM_DEV void quad_combine(const V& e1, const V& e2, const V& out, int r, const FB& fb) {
  team_sync();
  const int f1 = e1.flag(), f2 = e2.flag();
  if (f2 == FLAG_HEAD || f1 == FLAG_IDENT) {
    quad_copy(e2, out, r);
  } else if (f2 == FLAG_IDENT) {
    quad_copy(e1, out, r);
  } else {
    constexpr bool FBK = kFallback<R, true>;
    const bool ill = quad_combine_core<R, R, NX, FBK>(e1, e2, out, r, fb.thr);
    if constexpr (FBK && !FB::kRedo) {
      if (fb.thr > 0.f && ill && r == 0) fb.raise();
    } else if constexpr (FBK) {
      if (fb.thr > 0.f && ill) {
        team_sync();
        quad_combine_core<R, double, NX, false>(e1, e2, out, r, 0.f);
        if (r == 0) count_fallback(fb);
      }
    }
    if (r == 0) out.set_flag(f1);
  }
  team_sync();
}

}  // namespace psm
