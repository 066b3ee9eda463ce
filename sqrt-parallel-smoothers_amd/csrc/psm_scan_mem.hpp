// Memory-operand forms of the filtering / smoothing operators used by the scan
// kernels.  Both operands are read straight from memory (LDS or global,
// struct-of-arrays) when needed and the result is written field by field as
// soon as it is final, so a lane never holds two aggregates plus the
// 2nx x 2nx triangularisation at once.  The output may alias the second
// operand: every field of e2 is read for the last time before the same field
// of the output is written.
#pragma once
#include "psm_core.hpp"

namespace psm {

template <typename R, int NX>
struct FView {
  using real_type = R;
  R* base;
  int64_t stride;
  int64_t idx;
  static constexpr int QA = 0, QB = NX * NX, QU = NX * NX + NX, QE = 2 * NX * NX + NX,
                       QZ = 2 * NX * NX + 2 * NX, QF = 3 * NX * NX + 2 * NX;
  PSM_DEV R& at(int q) const { return base[(int64_t)q * stride + idx]; }
  PSM_DEV R& A(int i, int j) const { return at(QA + i * NX + j); }
  PSM_DEV R& b(int i) const { return at(QB + i); }
  PSM_DEV R& U(int i, int j) const { return at(QU + i * NX + j); }
  PSM_DEV R& eta(int i) const { return at(QE + i); }
  PSM_DEV R& Z(int i, int j) const { return at(QZ + i * NX + j); }
  PSM_DEV int flag() const { return (int)at(QF); }
  PSM_DEV void set_flag(int f) const { at(QF) = R(f); }
};

template <typename R, int NX>
struct SView {
  using real_type = R;
  R* base;
  int64_t stride;
  int64_t idx;
  static constexpr int QG = 0, QE = NX, QD = NX + NX * NX, QF = NX + 2 * NX * NX;
  PSM_DEV R& at(int q) const { return base[(int64_t)q * stride + idx]; }
  PSM_DEV R& g(int i) const { return at(QG + i); }
  PSM_DEV R& E(int i, int j) const { return at(QE + i * NX + j); }
  PSM_DEV R& D(int i, int j) const { return at(QD + i * NX + j); }
  PSM_DEV int flag() const { return (int)at(QF); }
  PSM_DEV void set_flag(int f) const { at(QF) = R(f); }
};

template <typename V>
PSM_DEV void view_copy(const V& s, const V& d) {
  constexpr int NQ = V::QF + 1;
#pragma unroll
  for (int q = 0; q < NQ; ++q) d.at(q) = s.at(q);
}

template <typename R, int NX>
PSM_DEV void view_identity(const FView<R, NX>& d) {
#pragma unroll
  for (int q = 0; q < FView<R, NX>::QF; ++q) d.at(q) = R(0);
#pragma unroll
  for (int i = 0; i < NX; ++i) d.A(i, i) = R(1);
  d.set_flag(FLAG_IDENT);
}

template <typename R, int NX>
PSM_DEV void view_identity(const SView<R, NX>& d) {
#pragma unroll
  for (int q = 0; q < SView<R, NX>::QF; ++q) d.at(q) = R(0);
#pragma unroll
  for (int i = 0; i < NX; ++i) d.E(i, i) = R(1);
  d.set_flag(FLAG_IDENT);
}

// ---------------------------------------------------------------- square-root filter
template <typename R, int NX, typename C = R>
PSM_DEV void sqrt_combine_mem(const FView<R, NX>& e1, const FView<R, NX>& e2,
                              const FView<R, NX>& out) {
  C X[2 * NX][2 * NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma((C)e1.U(k, i), (C)e2.Z(k, j), acc);
      X[i][j] = acc;
      X[i][NX + j] = (i == j) ? C(1) : C(0);
      X[NX + i][j] = (C)e2.Z(i, j);
      X[NX + i][NX + j] = C(0);
    }
  tria_inplace<C, 2 * NX, 2 * NX>(X);
  // M1 = U1 Xi11^{-T}
  C M1[NX][NX];
#pragma unroll
  for (int j = 0; j < NX; ++j) {
    const C inv = rcp_(X[j][j]);
#pragma unroll
    for (int r = 0; r < NX; ++r) {
      C s = (C)e1.U(r, j);
#pragma unroll
      for (int k = 0; k < j; ++k) s = fma(-M1[r][k], X[j][k], s);
      M1[r][j] = s * inv;
    }
  }
  // P = I - M1 Xi21^T
  C Pm[NX][NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      C acc = (i == j) ? C(1) : C(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(-M1[i][k], X[NX + j][k], acc);
      Pm[i][j] = acc;
    }
  // v = b1 + U1 (U1^T eta2) ; w = eta2 - Z2 (Z2^T b1)   (last reads of e2.eta, e2.Z)
  C v[NX], w[NX];
  {
    C t1[NX], t2[NX];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      C a1 = C(0), a2 = C(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        a1 = fma((C)e1.U(k, j), (C)e2.eta(k), a1);
        a2 = fma((C)e2.Z(k, j), (C)e1.b(k), a2);
      }
      t1[j] = a1;
      t2[j] = a2;
    }
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      C a1 = (C)e1.b(i), a2 = (C)e2.eta(i);
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        a1 = fma((C)e1.U(i, k), t1[k], a1);
        a2 = fma(-(C)e2.Z(i, k), t2[k], a2);
      }
      v[i] = a1;
      w[i] = a2;
    }
  }
  // eta = A1^T P^T w + eta1
  {
    C PTw[NX];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      C a = C(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) a = fma(Pm[k][j], w[k], a);
      PTw[j] = a;
    }
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      C a = (C)e1.eta(i);
#pragma unroll
      for (int k = 0; k < NX; ++k) a = fma((C)e1.A(k, i), PTw[k], a);
      out.eta(i) = (R)(a);
    }
  }
  // Z = tria([A1^T Xi22 | Z1])
  {
    C Zc[NX][2 * NX];
#pragma unroll
    for (int i = 0; i < NX; ++i)
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        C acc = C(0);
#pragma unroll
        for (int k = j; k < NX; ++k) acc = fma((C)e1.A(k, i), X[NX + k][NX + j], acc);
        Zc[i][j] = acc;
        Zc[i][NX + j] = (C)e1.Z(i, j);
      }
    tria_inplace<C, NX, 2 * NX>(Zc);
#pragma unroll
    for (int i = 0; i < NX; ++i)
#pragma unroll
      for (int j = 0; j < NX; ++j) out.Z(i, j) = (R)(Zc[i][j]);
  }
  // PA1 = P A1 ; Pv = P v
  C PA1[NX][NX], Pv[NX];
#pragma unroll
  for (int i = 0; i < NX; ++i) {
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(Pm[i][k], (C)e1.A(k, j), acc);
      PA1[i][j] = acc;
    }
    C a = C(0);
#pragma unroll
    for (int k = 0; k < NX; ++k) a = fma(Pm[i][k], v[k], a);
    Pv[i] = a;
  }
  // S = A2 M1 ; A = A2 PA1 ; b = A2 Pv + b2    (last reads of e2.A, e2.b)
  C S[NX][NX];
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    C a2[NX];
#pragma unroll
    for (int k = 0; k < NX; ++k) a2[k] = (C)e2.A(i, k);
    C bb = (C)e2.b(i);
#pragma unroll
    for (int k = 0; k < NX; ++k) bb = fma(a2[k], Pv[k], bb);
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      C s1 = C(0), s2 = C(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        s1 = fma(a2[k], M1[k][j], s1);
        s2 = fma(a2[k], PA1[k][j], s2);
      }
      S[i][j] = s1;
      out.A(i, j) = (R)(s2);
    }
    out.b(i) = (R)(bb);
  }
  // U = tria([S | U2])
  {
    C Uc[NX][2 * NX];
#pragma unroll
    for (int i = 0; i < NX; ++i)
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        Uc[i][j] = S[i][j];
        Uc[i][NX + j] = (C)e2.U(i, j);
      }
    tria_inplace<C, NX, 2 * NX>(Uc);
#pragma unroll
    for (int i = 0; i < NX; ++i)
#pragma unroll
      for (int j = 0; j < NX; ++j) out.U(i, j) = (R)(Uc[i][j]);
  }
}

// ---------------------------------------------------------------- square-root smoother
// out = e1 (+) e2 with e1 the later aggregate: g = E2 g1 + g2, E = E2 E1,
// D = tria([E2 D1 | D2]).
template <typename R, int NX, typename C = R>
PSM_DEV void sqrt_smoother_combine_mem(const SView<R, NX>& e1, const SView<R, NX>& e2,
                                       const SView<R, NX>& out) {
  C E2[NX][NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) E2[i][j] = (C)e2.E(i, j);
  C M[NX][2 * NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = j; k < NX; ++k) acc = fma(E2[i][k], (C)e1.D(k, j), acc);
      M[i][j] = acc;
      M[i][NX + j] = (C)e2.D(i, j);
    }
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    C g = (C)e2.g(i);
#pragma unroll
    for (int k = 0; k < NX; ++k) g = fma(E2[i][k], (C)e1.g(k), g);
    out.g(i) = (R)(g);
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      C acc = C(0);
#pragma unroll
      for (int k = 0; k < NX; ++k) acc = fma(E2[i][k], (C)e1.E(k, j), acc);
      out.E(i, j) = (R)(acc);
    }
  }
  tria_inplace<C, NX, 2 * NX>(M);
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) out.D(i, j) = (R)(M[i][j]);
}

}  // namespace psm

namespace psm {

template <typename R, int NX>
PSM_DEV void view_load(const FView<R, NX>& v, FAgg<R, NX>& a) {
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    a.b[i] = v.b(i);
    a.eta[i] = v.eta(i);
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      a.A[i][j] = v.A(i, j);
      a.U[i][j] = v.U(i, j);
      a.Z[i][j] = v.Z(i, j);
    }
  }
  a.flag = v.flag();
}
template <typename R, int NX>
PSM_DEV void view_store(const FAgg<R, NX>& a, const FView<R, NX>& v) {
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    v.b(i) = a.b[i];
    v.eta(i) = a.eta[i];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      v.A(i, j) = a.A[i][j];
      v.U(i, j) = a.U[i][j];
      v.Z(i, j) = a.Z[i][j];
    }
  }
  v.set_flag(a.flag);
}
template <typename R, int NX>
PSM_DEV void view_load(const SView<R, NX>& v, SAgg<R, NX>& a) {
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    a.g[i] = v.g(i);
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      a.E[i][j] = v.E(i, j);
      a.D[i][j] = v.D(i, j);
    }
  }
  a.flag = v.flag();
}
template <typename R, int NX>
PSM_DEV void view_store(const SAgg<R, NX>& a, const SView<R, NX>& v) {
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    v.g(i) = a.g[i];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      v.E(i, j) = a.E[i][j];
      v.D(i, j) = a.D[i][j];
    }
  }
  v.set_flag(a.flag);
}

// out <- e1 (+) e2 with the HEAD / IDENT rules of combine_into; out may alias e2.
template <typename R, int NX, bool SQRT>
PSM_DEV void combine_mem(const FView<R, NX>& e1, const FView<R, NX>& e2, const FView<R, NX>& out) {
  const int f1 = e1.flag(), f2 = e2.flag();
  if (f2 == FLAG_HEAD || f1 == FLAG_IDENT) { view_copy(e2, out); return; }
  if (f2 == FLAG_IDENT) { view_copy(e1, out); return; }
  if constexpr (SQRT) {
    sqrt_combine_mem<R, NX>(e1, e2, out);
  } else {
    FAgg<R, NX> a, b;
    view_load(e1, a);
    view_load(e2, b);
    std_combine_core<R, NX>(a, b, b);
    view_store(b, out);
  }
  out.set_flag(f1);
}

template <typename R, int NX, bool SQRT>
PSM_DEV void combine_mem(const SView<R, NX>& e1, const SView<R, NX>& e2, const SView<R, NX>& out) {
  const int f1 = e1.flag(), f2 = e2.flag();
  if (f2 == FLAG_HEAD || f1 == FLAG_IDENT) { view_copy(e2, out); return; }
  if (f2 == FLAG_IDENT) { view_copy(e1, out); return; }
  if constexpr (SQRT) {
    sqrt_smoother_combine_mem<R, NX>(e1, e2, out);
  } else {
    SAgg<R, NX> a, b;
    view_load(e1, a);
    view_load(e2, b);
    std_smoother_combine_core<R, NX>(a, b);
    view_store(b, out);
  }
  out.set_flag(f1);
}

template <typename Agg> struct ViewOf;
template <typename R, int NX> struct ViewOf<FAgg<R, NX>> { using type = FView<R, NX>; };
template <typename R, int NX> struct ViewOf<SAgg<R, NX>> { using type = SView<R, NX>; };

}  // namespace psm

namespace psm {

template <typename R, int NX>
PSM_DEV bool view_ill(const FView<R, NX>& v, float thr) {
  R L[NX][NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) L[i][j] = (i == j) ? v.U(i, i) : R(0);
  return ill_conditioned(L, thr);
}
template <typename R, int NX>
PSM_DEV bool view_ill(const SView<R, NX>& v, float thr) {
  R L[NX][NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) L[i][j] = (i == j) ? v.D(i, i) : R(0);
  return ill_conditioned(L, thr);
}

// combine_mem with the FP64 re-triangularisation fallback (FP32 square-root
// form).  out must not alias e2 here, so the operands survive for the redo.
template <typename R, int NX, bool SQRT, typename V, typename FB>
PSM_DEV void combine_mem_fb(const V& e1, const V& e2, const V& out, const FB& fb) {
  const int f1 = e1.flag(), f2 = e2.flag();
  combine_mem<R, NX, SQRT>(e1, e2, out);
  if constexpr (kFallback<R, SQRT> && !FB::kRedo) {
    if (fb.thr > 0.f && f2 != FLAG_HEAD && f1 != FLAG_IDENT && f2 != FLAG_IDENT && view_ill(out, fb.thr))
      fb.raise();
  } else if constexpr (kFallback<R, SQRT>) {
    if (fb.thr > 0.f && f2 != FLAG_HEAD && f1 != FLAG_IDENT && f2 != FLAG_IDENT &&
        view_ill(out, fb.thr)) {
      // the memory-operand operator in FP64 on the FP32 operands: no aggregate
      // is ever whole in registers
      if constexpr (std::is_same<V, FView<R, NX>>::value)
        sqrt_combine_mem<R, NX, double>(e1, e2, out);
      else
        sqrt_smoother_combine_mem<R, NX, double>(e1, e2, out);
      count_fallback(fb);
    }
  }
}

}  // namespace psm
