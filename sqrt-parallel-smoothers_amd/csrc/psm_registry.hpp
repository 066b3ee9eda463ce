// The compiled (dtype, form, nx, ny) combinations.  nx in 1..5, ny in 1..3.
#pragma once
#define PSM_FOR_NY(X, T, NX) X(T, NX, 1) X(T, NX, 2) X(T, NX, 3)
#define PSM_FOR_TAG(X, T) \
  PSM_FOR_NY(X, T, 1) PSM_FOR_NY(X, T, 2) PSM_FOR_NY(X, T, 3) PSM_FOR_NY(X, T, 4) PSM_FOR_NY(X, T, 5)
#define PSM_FOR_ALL(X) \
  PSM_FOR_TAG(X, f32s) PSM_FOR_TAG(X, f32c) PSM_FOR_TAG(X, f64s) PSM_FOR_TAG(X, f64c)
#define PSM_MAX_NX 5
#define PSM_MAX_NY 3
