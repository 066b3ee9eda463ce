// Team-cooperative square-root scan operators.
//
// The per-lane operators of psm_scan_mem.hpp run one whole combination in one
// lane: ~6,000 dependent instructions for the nx=5 filtering operator, which
// is what bounds a scan level whose few items cannot fill the chip.  Here a
// team of 16 lanes -- one DPP row of a wavefront -- runs one combination:
// lane r owns row r of the blocks being triangularised, and every value
// another lane needs is broadcast inside the row with DPP row_newbcast (a
// register-to-register move, no LDS round trip).  The Householder
// triangularisation is row-parallel: the pivot row's reflector is broadcast
// and every lane updates its own row.
//
// The code is branch-free up to the final stores: a DPP broadcast reads the
// source lane's register, so every lane of the team executes every broadcast.
// Teams of one wavefront may take different paths (combine / copy), since a
// DPP row never crosses teams.  Operands are staged in LDS once (one
// overlapped batch of loads), the result is written at the end, so `out` must
// not alias `e1` or `e2`; the FP32 fallback recomputes from the staged copies.
#pragma once
#include <cstdint>
#include <type_traits>

#include "psm_scan_mem.hpp"

namespace psm {

constexpr int kTeamLanes = 16;  // one DPP row
// Virtual lanes (teams) per scan workgroup, for the filter and the smoother
// alike (the two scans share one level plan).
constexpr int kTeamVL = 16;

template <int I, int N, typename F>
PSM_DEV void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// Value of lane L of this lane's row of 16 (DPP row_newbcast:L).
template <int L>
PSM_DEV int bcast_i32(int v) {
  static_assert(L >= 0 && L < 16, "row_newbcast lane");
  return __builtin_amdgcn_update_dpp(0, v, 0x150 + L, 0xF, 0xF, false);
}
template <int L>
PSM_DEV float bcast(float v) {
  return __builtin_bit_cast(float, bcast_i32<L>(__builtin_bit_cast(int, v)));
}
template <int L>
PSM_DEV double bcast(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)bcast_i32<L>((int)(uint32_t)u);
  const uint32_t hi = (uint32_t)bcast_i32<L>((int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

PSM_DEV void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename V>
PSM_DEV void team_copy(const V& s, const V& d, int r) {
  constexpr int NQ = V::QF + 1;
  for (int q = r; q < NQ; q += kTeamLanes) d.at(q) = s.at(q);
}

template <typename R, int NX>
PSM_DEV void team_identity(const FView<R, NX>& d, int r) {
  constexpr int NQ = FView<R, NX>::QF + 1;
  for (int q = r; q < NQ; q += kTeamLanes) {
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
  for (int q = r; q < NQ; q += kTeamLanes) {
    R v = R(0);
    if (q >= QE && q < QE + NX * NX && ((q - QE) / NX) == ((q - QE) % NX)) v = R(1);
    if (q == SView<R, NX>::QF) v = R(FLAG_IDENT);
    d.at(q) = v;
  }
}

// Broadcast from lane G0 + I to the lanes of group 0 and from lane G1 + I to
// the others (two matrices triangularised side by side when G1 != G0).
template <int I, int G0, int G1, typename C>
PSM_DEV C bcast2(C v, bool grp0) {
  const C a = bcast<G0 + I>(v);
  if constexpr (G0 == G1) {
    return a;
  } else {
    const C b = bcast<G1 + I>(v);
    return grp0 ? a : b;
  }
}

// Row-parallel Householder triangularisation from the right of an M x NC
// matrix whose row i is held in x by lane G0 + i -- and, when G1 != G0, of a
// second M x NC matrix held by lanes G1 + i.  The reflector is that of
// tria_inplace (non-negative diagonal, Parlett's v0).  Every team lane calls
// it; `me` is the lane's row in its matrix (>= M: not a row).
template <typename C, int M, int NC, int G0, int G1>
PSM_DEV void team_tria(C (&x)[NC], int me, bool grp0) {
  constexpr int NR = M < NC ? M : NC;
  static_for<0, NR>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    // every lane forms the reflector of its own row; the pivot's is used
    const C alpha = x[i];
    C sigma = C(0);
#pragma unroll
    for (int j = i + 1; j < NC; ++j) sigma = fma(x[j], x[j], sigma);
    const C nrm = sqrt_(fma(alpha, alpha, sigma));
    const C v0 = (alpha > C(0)) ? -sigma * rcp_(alpha + nrm) : alpha - nrm;
    const C vv = fma(v0, v0, sigma);
    const C f = (vv > C(0)) ? C(2) * rcp_(vv) : C(0);
    const C bv0 = bcast2<i, G0, G1>(v0, grp0);
    const C bf = bcast2<i, G0, G1>(f, grp0);
    C bx[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) bx[j] = C(0);
    static_for<i + 1, NC>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      bx[j] = bcast2<i, G0, G1>(x[j], grp0);
    });
    C dot = bv0 * x[i];
#pragma unroll
    for (int j = i + 1; j < NC; ++j) dot = fma(bx[j], x[j], dot);
    // rows below the pivot are reflected; other lanes apply a zero update
    const C t = (me > i && me < M) ? bf * dot : C(0);
    const bool piv = me == i;
    x[i] = piv ? nrm : fma(-t, bv0, x[i]);
#pragma unroll
    for (int j = i + 1; j < NC; ++j) x[j] = piv ? C(0) : fma(-t, bx[j], x[j]);
  });
}

// ---------------------------------------------------------------- filtering
// out = e1 (+) e2 (e1 earlier), the blocked form of
// parsmooth/parallel/_operators.py:43-77:
//   Xi = tria([[U1^T Z2, I], [Z2, 0]]),  M1 = U1 Xi11^{-T},  P = I - M1 Xi21^T
//   A = A2 P A1,  b = A2 P (b1 + U1 U1^T eta2) + b2,  U = tria([A2 M1 | U2])
//   eta = A1^T P^T (eta2 - Z2 Z2^T b1) + eta1,  Z = tria([A1^T Xi22 | Z1])
// Lanes 0..2nx-1 hold the rows of Xi; then lanes 0..nx-1 build Z and eta and
// lanes nx..2nx-1 build A, b and U (the two final trias run side by side).
// Returns min|diag U| / max|diag U| (on every lane; 0 if not finite).
template <typename R, typename C, int NX>
PSM_DEV C team_combine_core(const FView<R, NX>& e1, const FView<R, NX>& e2,
                            const FView<R, NX>& out, int r) {
  constexpr int n = NX, n2 = 2 * NX;
  const bool top = r < n;
  const int ri = top ? r : (r < n2 ? r - n : 0);  // row of this lane in its block

  // rows of [[U1^T Z2, I], [Z2, 0]]; entry ri of t1 = U1^T eta2, t2 = Z2^T b1
  C x[n2];
  C t1r = C(0), t2r = C(0);
#pragma unroll
  for (int j = 0; j < n; ++j) {
    C acc = C(0);
#pragma unroll
    for (int k = 0; k < n; ++k) acc = fma((C)e1.U(k, ri), (C)e2.Z(k, j), acc);
    const C z = (C)e2.Z(ri, j);
    x[j] = top ? acc : z;
    x[n + j] = (top && ri == j) ? C(1) : C(0);
  }
#pragma unroll
  for (int k = 0; k < n; ++k) {
    t1r = fma((C)e1.U(k, ri), (C)e2.eta(k), t1r);
    t2r = fma((C)e2.Z(k, ri), (C)e1.b(k), t2r);
  }
  team_tria<C, n2, n2, 0, 0>(x, r, true);

  // entry ri of v = b1 + U1 t1 and w = eta2 - Z2 t2
  C vr = (C)e1.b(ri), wr = (C)e2.eta(ri);
  static_for<0, n>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    vr = fma((C)e1.U(ri, k), bcast<k>(t1r), vr);
    wr = fma(-(C)e2.Z(ri, k), bcast<k>(t2r), wr);
  });
  // row ri of M1 = U1 Xi11^{-T} and of P = I - M1 Xi21^T
  C m1[n], p[n];
  static_for<0, n>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    C s = (C)e1.U(ri, j);
    static_for<0, j>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      s = fma(-m1[k], bcast<j>(x[k]), s);
    });
    m1[j] = s * rcp_(bcast<j>(x[j]));
  });
  static_for<0, n>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    C acc = (ri == j) ? C(1) : C(0);
    static_for<0, n>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      acc = fma(-m1[k], bcast<n + j>(x[k]), acc);
    });
    p[j] = acc;
  });
  // P^T w, u = A2(ri, :) P and S(ri, :) = A2(ri, :) M1
  C ptw[n], u[n], srow[n];
#pragma unroll
  for (int j = 0; j < n; ++j) {
    ptw[j] = C(0);
    u[j] = C(0);
    srow[j] = C(0);
  }
  static_for<0, n>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const C wk = bcast<k>(wr);
    const C a2k = (C)e2.A(ri, k);
    static_for<0, n>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      const C pkj = bcast<k>(p[j]);
      ptw[j] = fma(pkj, wk, ptw[j]);
      u[j] = fma(a2k, pkj, u[j]);
      srow[j] = fma(a2k, bcast<k>(m1[j]), srow[j]);
    });
  });
  C vfull[n];
  static_for<0, n>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    vfull[k] = bcast<k>(vr);
  });
  // lanes < n: eta(ri) and row ri of [A1^T Xi22 | Z1]
  // lanes n..2n-1: A(ri, :) = u A1, b(ri) = b2(ri) + u v and row ri of [S | U2]
  C eta = (C)e1.eta(ri);
#pragma unroll
  for (int k = 0; k < n; ++k) eta = fma((C)e1.A(k, ri), ptw[k], eta);
  C bb = (C)e2.b(ri);
#pragma unroll
  for (int k = 0; k < n; ++k) bb = fma(u[k], vfull[k], bb);
  C arow[n];
#pragma unroll
  for (int j = 0; j < n; ++j) {
    C acc = C(0);
#pragma unroll
    for (int k = 0; k < n; ++k) acc = fma(u[k], (C)e1.A(k, j), acc);
    arow[j] = acc;
  }
  C y[n2];
  static_for<0, n>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    C zc = C(0);
    static_for<j, n>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      zc = fma((C)e1.A(k, ri), bcast<n + k>(x[n + j]), zc);
    });
    y[j] = top ? zc : srow[j];
    y[n + j] = top ? (C)e1.Z(ri, j) : (C)e2.U(ri, j);
  });
  team_tria<C, n, n2, 0, n>(y, r < n2 ? ri : n, top);

  // |diag U| (the rows of U sit on lanes n..2n-1)
  C mn = C(0), mx = C(0);
  static_for<0, n>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const C d = abs_(bcast<n + i>(y[i]));
    mn = (i == 0 || d < mn) ? d : mn;
    mx = (i == 0 || d > mx) ? d : mx;
  });
  if (top) {
    out.eta(ri) = (R)eta;
#pragma unroll
    for (int j = 0; j < n; ++j) out.Z(ri, j) = (R)y[j];
  } else if (r < n2) {
    out.b(ri) = (R)bb;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      out.A(ri, j) = (R)arow[j];
      out.U(ri, j) = (R)y[j];
    }
  }
  return (mn >= C(0) && mx > C(0)) ? mn / mx : C(0);
}

// ---------------------------------------------------------------- smoothing
// out = e1 (+) e2 with e1 the later aggregate
// (parsmooth/parallel/_operators.py:104-125): g = E2 g1 + g2, E = E2 E1,
// D = tria([E2 D1 | D2]); lane r < nx owns row r.  Returns min|diag D| /
// max|diag D| (0 if not finite).
template <typename R, typename C, int NX>
PSM_DEV C team_combine_core(const SView<R, NX>& e1, const SView<R, NX>& e2,
                            const SView<R, NX>& out, int r) {
  constexpr int n = NX, n2 = 2 * NX;
  const int ri = r < n ? r : 0;
  C e2r[n];
#pragma unroll
  for (int k = 0; k < n; ++k) e2r[k] = (C)e2.E(ri, k);
  C g = (C)e2.g(ri);
#pragma unroll
  for (int k = 0; k < n; ++k) g = fma(e2r[k], (C)e1.g(k), g);
  C erow[n], y[n2];
#pragma unroll
  for (int j = 0; j < n; ++j) {
    C acc = C(0), acd = C(0);
#pragma unroll
    for (int k = 0; k < n; ++k) acc = fma(e2r[k], (C)e1.E(k, j), acc);
#pragma unroll
    for (int k = j; k < n; ++k) acd = fma(e2r[k], (C)e1.D(k, j), acd);
    erow[j] = acc;
    y[j] = acd;
    y[n + j] = (C)e2.D(ri, j);
  }
  team_tria<C, n, n2, 0, 0>(y, r, true);
  C mn = C(0), mx = C(0);
  static_for<0, n>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const C d = abs_(bcast<i>(y[i]));
    mn = (i == 0 || d < mn) ? d : mn;
    mx = (i == 0 || d > mx) ? d : mx;
  });
  if (r < n) {
    out.g(ri) = (R)g;
#pragma unroll
    for (int j = 0; j < n; ++j) {
      out.E(ri, j) = (R)erow[j];
      out.D(ri, j) = (R)y[j];
    }
  }
  return (mn >= C(0) && mx > C(0)) ? mn / mx : C(0);
}

// LDS bytes per team: both operands staged in the storage type.
template <typename V> constexpr int team_scratch_bytes() {
  using R = typename V::real_type;
  return ((2 * (V::QF + 1) * (int)sizeof(R)) + 15) / 16 * 16;
}

// One combination by a team: the flag cases of combine_mem, the cooperative
// operator, and for FP32 the recomputation in FP64 when the combined factor's
// min|diag| / max|diag| falls below the threshold (or is not finite).
template <typename R, int NX, typename V, typename FB>
PSM_DEV void team_combine(const V& e1, const V& e2, const V& out, int r, void* scratch,
                          const FB& fb) {
  team_sync();
  const int f1 = e1.flag(), f2 = e2.flag();
  if (f2 == FLAG_HEAD || f1 == FLAG_IDENT) {
    team_copy(e2, out, r);
  } else if (f2 == FLAG_IDENT) {
    team_copy(e1, out, r);
  } else {
    constexpr int NQ = V::QF + 1;
    R* st = static_cast<R*>(scratch);
    const V s1{st, 1, 0}, s2{st + NQ, 1, 0};
    for (int q = r; q < NQ; q += kTeamLanes) {
      st[q] = e1.at(q);
      st[NQ + q] = e2.at(q);
    }
    team_sync();
    const R ratio = team_combine_core<R, R, NX>(s1, s2, out, r);
    if constexpr (kFallback<R, true> && !FB::kRedo) {
      if (fb.thr > 0.f && !(ratio >= R(fb.thr)) && r == 0) fb.raise();
    } else if constexpr (kFallback<R, true>) {
      if (fb.thr > 0.f && !(ratio >= R(fb.thr))) {
        team_combine_core<R, double, NX>(s1, s2, out, r);
        if (r == 0) count_fallback(fb);
      }
    }
    if (r == 0) out.set_flag(f1);
  }
  team_sync();
}

}  // namespace psm
