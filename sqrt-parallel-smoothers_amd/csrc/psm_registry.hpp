// The compiled (dtype, form, nx, ny) combinations.  nx in 1..5, ny in 1..3.
#pragma once
#define PSM_FOR_NY(X, T, NX) X(T, NX, 1) X(T, NX, 2) X(T, NX, 3)
#define PSM_FOR_TAG(X, T) \
  PSM_FOR_NY(X, T, 1) PSM_FOR_NY(X, T, 2) PSM_FOR_NY(X, T, 3) PSM_FOR_NY(X, T, 4) PSM_FOR_NY(X, T, 5)
#define PSM_FOR_ALL(X) \
  PSM_FOR_TAG(X, f32s) PSM_FOR_TAG(X, f32c) PSM_FOR_TAG(X, f64s) PSM_FOR_TAG(X, f64c)
#define PSM_MAX_NX 5
#define PSM_MAX_NY 3

// Tangent (forward-mode) slices, square-root form only.
#define PSM_FOR_DUAL(X) X(d64s, 1, 1) X(d64s, 2, 1) X(d64s, 2, 2) X(d64s, 3, 2) X(d64s, 4, 2) X(d64s, 5, 2) X(d32s, 5, 2)
