// Forward-mode (tangent) numbers: a value and one directional derivative.
// Instantiating the filter / smoother kernels on DualT<B> propagates a tangent
// of every input (model parameters, noise factors, prior, nominal trajectory)
// through the exact same arithmetic as the value, which is how the library
// evaluates Jacobian-vector products of the iterated smoother and of its
// log-likelihood.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "psm_linalg.hpp"

namespace psm {

// Plain overloads so that the tangent overloads below do not hide the scalar
// ones inside this namespace.
PSM_DEV double fma(double a, double b, double c) { return ::fma(a, b, c); }
PSM_DEV float fma(float a, float b, float c) { return ::fmaf(a, b, c); }
PSM_DEV bool isfinite(double a) { return ::isfinite(a); }
PSM_DEV bool isfinite(float a) { return ::isfinite(a); }

template <typename B>
struct DualT {
  B v;
  B d;
  DualT() = default;
  PSM_DEV constexpr DualT(double x) : v((B)x), d(B(0)) {}
  PSM_DEV constexpr DualT(float x) : v((B)x), d(B(0)) {}
  PSM_DEV constexpr DualT(int x) : v((B)x), d(B(0)) {}
  PSM_DEV constexpr DualT(B x, B y) : v(x), d(y) {}
  template <typename B2>
  PSM_DEV explicit constexpr DualT(const DualT<B2>& o) : v((B)o.v), d((B)o.d) {}
  PSM_DEV explicit operator int() const { return (int)v; }
  PSM_DEV DualT& operator+=(const DualT& o) { v += o.v; d += o.d; return *this; }
  PSM_DEV DualT& operator-=(const DualT& o) { v -= o.v; d -= o.d; return *this; }
  PSM_DEV DualT& operator*=(const DualT& o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
};

template <typename T> struct is_dual : std::false_type {};
template <typename B> struct is_dual<DualT<B>> : std::true_type {};

template <typename B> PSM_DEV DualT<B> operator+(DualT<B> a, DualT<B> b) { return {a.v + b.v, a.d + b.d}; }
template <typename B> PSM_DEV DualT<B> operator-(DualT<B> a, DualT<B> b) { return {a.v - b.v, a.d - b.d}; }
template <typename B> PSM_DEV DualT<B> operator-(DualT<B> a) { return {-a.v, -a.d}; }
template <typename B> PSM_DEV DualT<B> operator*(DualT<B> a, DualT<B> b) {
  return {a.v * b.v, fma(a.d, b.v, a.v * b.d)};
}
template <typename B> PSM_DEV DualT<B> operator/(DualT<B> a, DualT<B> b) {
  const B q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
template <typename B> PSM_DEV bool operator<(DualT<B> a, DualT<B> b) { return a.v < b.v; }
template <typename B> PSM_DEV bool operator>(DualT<B> a, DualT<B> b) { return a.v > b.v; }
template <typename B> PSM_DEV bool operator<=(DualT<B> a, DualT<B> b) { return a.v <= b.v; }
template <typename B> PSM_DEV bool operator>=(DualT<B> a, DualT<B> b) { return a.v >= b.v; }
template <typename B> PSM_DEV bool operator==(DualT<B> a, DualT<B> b) { return a.v == b.v; }
template <typename B> PSM_DEV bool operator!=(DualT<B> a, DualT<B> b) { return a.v != b.v; }

template <typename B> PSM_DEV DualT<B> fma(DualT<B> a, DualT<B> b, DualT<B> c) {
  return {fma(a.v, b.v, c.v), fma(a.d, b.v, fma(a.v, b.d, c.d))};
}
// derivative of sqrt taken as 0 at 0 (zero rows carry zero tangents)
template <typename B> PSM_DEV DualT<B> sqrt_(DualT<B> a) {
  const B s = sqrt_(a.v);
  return {s, s > B(0) ? a.d / (B(2) * s) : B(0)};
}
template <typename B> PSM_DEV DualT<B> rcp_(DualT<B> a) { return DualT<B>(1.0) / a; }
template <typename B> PSM_DEV DualT<B> abs_(DualT<B> a) { return a.v < B(0) ? -a : a; }
template <typename B> PSM_DEV DualT<B> log_(DualT<B> a) { return {log_(a.v), a.d / a.v}; }
template <typename B> PSM_DEV bool isfinite(DualT<B> a) { return isfinite(a.v) && isfinite(a.d); }

// value type used for FP64 accumulation of log-likelihood terms
template <typename R> struct AccOf { using type = double; };
template <typename B> struct AccOf<DualT<B>> { using type = DualT<double>; };

template <typename R> PSM_DEV typename AccOf<R>::type to_acc(R x) { return (double)x; }
template <typename B> PSM_DEV DualT<double> to_acc(DualT<B> x) { return {(double)x.v, (double)x.d}; }

}  // namespace psm
