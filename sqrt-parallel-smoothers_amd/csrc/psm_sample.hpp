// Pathwise sampling from the smoothing distribution
// (parsmooth/_pathwise_sampler.py:13-38 and :64-127).
//
// With the smoother elements (g_k, E_k, D_k) of the nominal trajectory, a
// sample path is the backward affine recursion
//     x_k = E_k x_{k+1} + g_k + D_k eps_k          (D_k: chol; in covariance
//                                                   form its Cholesky factor)
// whose last state of every sequence is its own head element (E = 0:
// x_T = m_T + L_T eps_T), so a ragged batch needs no segment bookkeeping.  The
// reference runs the recursion as an associative scan of the affine maps
// (G2 G1, G2 e1 + e2); here it is a two-level scan over chunks of Cs states:
//   1. k_sample_reduce  lane (chunk, sample): the chunk's composite map
//                       P_c = E_i0 ... E_i1-1 and offset o_c
//   2. k_sample_bounds  lane per sample: the states at the chunk starts,
//                       backwards over the chunks
//   3. k_sample_apply   lane (chunk, sample): the recursion through the chunk
// with depth ~ 3 sqrt(n_states) instead of n_states.  The standard normals
// eps (n_states, n_samples, nx) are drawn by the caller.
#pragma once
#include "psm_kernels.hpp"

namespace psm {

inline int64_t sample_chunk(int64_t M) {
  int64_t c = 16;
  while (c * c < M && c < 1024) c <<= 1;
  return c;
}

template <typename R, int NX, bool SQRT>
PSM_DEV void sample_element(const R* __restrict__ elems, int64_t Mp, int64_t pos,
                            const R* __restrict__ eps, R (&E)[NX][NX], R (&e)[NX]) {
  SAgg<R, NX> a;
  load_agg(elems, Mp, pos, a);
  copy(a.E, E);
  R L[NX][NX];
  if constexpr (SQRT) {
    copy(a.D, L);
  } else {
    cholesky(a.D, L);
  }
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    R s = a.g[i];
#pragma unroll
    for (int j = 0; j <= i; ++j) s = fma(L[i][j], eps[j], s);
    e[i] = s;
  }
}

template <typename R, int NX, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_sample_reduce(
    int64_t M, int64_t Mp, int chunk, int64_t Cs, int64_t nch, int S, const R* __restrict__ elems,
    const R* __restrict__ eps, R* __restrict__ P, R* __restrict__ o) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nch * S) return;
  const int64_t c = t / S;
  const int s = (int)(t - c * S);
  const StateMap sm{chunk};
  const int64_t i0 = c * Cs;
  const int64_t i1 = min(i0 + Cs, M);
  R Pm[NX][NX], ov[NX];
#pragma unroll
  for (int i = 0; i < NX; ++i)
#pragma unroll
    for (int j = 0; j < NX; ++j) Pm[i][j] = (i == j) ? R(1) : R(0);
  set_zero(ov);
  for (int64_t k = i1 - 1; k >= i0; --k) {
    R E[NX][NX], e[NX];
    sample_element<R, NX, SQRT>(elems, Mp, sm(k), eps + (k * S + s) * NX, E, e);
    R Pn[NX][NX], on[NX];
    mm(E, Pm, Pn);
    mv(E, ov, on);
#pragma unroll
    for (int i = 0; i < NX; ++i) on[i] += e[i];
    copy(Pn, Pm);
    copy(on, ov);
  }
  st_aos(o, t, ov);
  if (s == 0) st_aos(P, c, Pm);
}

template <typename R, int NX>
__global__ __launch_bounds__(kLaneWG) void k_sample_bounds(int64_t nch, int S,
                                                           const R* __restrict__ P,
                                                           const R* __restrict__ o,
                                                           R* __restrict__ xb) {
  const int s = (int)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= S) return;
  R x[NX];
  set_zero(x);
  for (int64_t c = nch - 1; c >= 0; --c) {
    R Pm[NX][NX], ov[NX], xn[NX];
    ld_aos(P, c, Pm);
    ld_aos(o, c * S + s, ov);
    mv(Pm, x, xn);
#pragma unroll
    for (int i = 0; i < NX; ++i) x[i] = xn[i] + ov[i];
    st_aos(xb, c * S + s, x);
  }
}

template <typename R, int NX, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_sample_apply(
    int64_t M, int64_t Mp, int chunk, int64_t Cs, int64_t nch, int S, const R* __restrict__ elems,
    const R* __restrict__ eps, const R* __restrict__ xb, R* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nch * S) return;
  const int64_t c = t / S;
  const int s = (int)(t - c * S);
  const StateMap sm{chunk};
  const int64_t i0 = c * Cs;
  const int64_t i1 = min(i0 + Cs, M);
  R x[NX];
  if (c + 1 < nch) ld_aos(xb, (c + 1) * S + s, x);
  else set_zero(x);
  for (int64_t k = i1 - 1; k >= i0; --k) {
    R E[NX][NX], e[NX], xn[NX];
    sample_element<R, NX, SQRT>(elems, Mp, sm(k), eps + (k * S + s) * NX, E, e);
    mv(E, x, xn);
#pragma unroll
    for (int i = 0; i < NX; ++i) x[i] = xn[i] + e[i];
    st_aos(out, k * S + s, x);
  }
}

}  // namespace psm
