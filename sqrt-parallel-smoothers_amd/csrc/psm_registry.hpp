// The compiled (dtype, form, nx, ny) combinations.  nx in 1..8, ny in 1..4.
#pragma once
#define PSM_FOR_NY(X, T, NX) X(T, NX, 1) X(T, NX, 2) X(T, NX, 3) X(T, NX, 4)
#define PSM_FOR_TAG(X, T)                                                                          \
  PSM_FOR_NY(X, T, 1) PSM_FOR_NY(X, T, 2) PSM_FOR_NY(X, T, 3) PSM_FOR_NY(X, T, 4)                  \
  PSM_FOR_NY(X, T, 5) PSM_FOR_NY(X, T, 6) PSM_FOR_NY(X, T, 7) PSM_FOR_NY(X, T, 8)
#define PSM_FOR_ALL(X) \
  PSM_FOR_TAG(X, f32s) PSM_FOR_TAG(X, f32c) PSM_FOR_TAG(X, f64s) PSM_FOR_TAG(X, f64c)
#define PSM_MAX_NX 8
#define PSM_MAX_NY 4

// Tangent (forward-mode) slices, square-root form only.
#define PSM_FOR_DUAL(X) X(d64s, 1, 1) X(d64s, 2, 1) X(d64s, 2, 2) X(d64s, 3, 2) X(d64s, 4, 2) X(d64s, 5, 2) X(d32s, 5, 2)
