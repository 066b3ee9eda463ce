// Sequential filter and smoother (parallel=False): one lane per sequence runs
// the Kalman recursion through its steps, linearizing as it goes.
//
// Without a nominal trajectory the reference linearizes at the running
// estimate -- the transition at the filtered state, the observation at the
// predicted state (parsmooth/sequential/_filtering.py:31-43), and in the
// smoother the transition at the smoothed successor
// (parsmooth/sequential/_smoothing.py:27-33) -- which no parallel scan can
// reproduce, since each linearization point depends on the previous step.
// With a nominal trajectory both read it instead.  A ragged batch of many
// short sequences fills the device one sequence per lane with no scan work.
#pragma once
#include "psm_kernels.hpp"

namespace psm {

// Predicted distribution (m1, L1) of one transition step: the linearization
// point of the observation when there is no nominal trajectory.
template <typename R, int NX, bool SQRT>
PSM_DEV void seq_predict(const R (&F)[NX][NX], const R (&cq)[NX][NX], const R (&bt)[NX],
                         const R (&m)[NX], const R (&L)[NX][NX], R (&m1)[NX], R (&L1)[NX][NX]) {
  mv(F, m, m1);
#pragma unroll
  for (int i = 0; i < NX; ++i) m1[i] += bt[i];
  if constexpr (SQRT) {
    R M[NX][2 * NX];
#pragma unroll
    for (int i = 0; i < NX; ++i)
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        R acc = R(0);
#pragma unroll
        for (int k = 0; k < NX; ++k) acc = fma(F[i][k], L[k][j], acc);
        M[i][j] = acc;
        M[i][NX + j] = cq[i][j];
      }
    tria_inplace<R, NX, 2 * NX>(M);
#pragma unroll
    for (int i = 0; i < NX; ++i)
#pragma unroll
      for (int j = 0; j < NX; ++j) L1[i][j] = M[i][j];
  } else {
    R FP[NX][NX];
#pragma unroll
    for (int i = 0; i < NX; ++i)
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        R acc = R(0);
#pragma unroll
        for (int k = 0; k < NX; ++k) acc = fma(F[i][k], L[k][j], acc);
        FP[i][j] = acc;
      }
#pragma unroll
    for (int i = 0; i < NX; ++i)
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        R acc = cq[i][j];
#pragma unroll
        for (int k = 0; k < NX; ++k) acc = fma(FP[i][k], F[j][k], acc);
        L1[i][j] = acc;
      }
  }
}

// Lane s filters sequence s: states [soff[s], soff[s+1]), state soff[s] the
// prior.  nom_m == nullptr: linearize at the running estimate.  ell (may be
// null) receives the per-sequence log-likelihood, accumulated in FP64.
template <typename R, int NX, int NY, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_seq_filter(
    int nseq, const int64_t* __restrict__ soff, ModelView<R> tr, ModelView<R> ob, SigmaView<R> sp,
    int use_sigma, const R* __restrict__ nom_m, const R* __restrict__ nom_L,
    const R* __restrict__ ys, const R* __restrict__ m0, const R* __restrict__ L0,
    R* __restrict__ out_m, R* __restrict__ out_L, typename AccOf<R>::type* __restrict__ ell,
    Fallback fb) {
  using Acc = typename AccOf<R>::type;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseq) return;
  const int64_t a = soff[s], e = soff[s + 1];
  R m[NX], L[NX][NX];
  ld_aos(m0, s, m);
  ld_aos(L0, s, L);
  st_aos(out_m, a, m);
  st_aos(out_L, a, L);
  Acc acc = Acc(0.0);
  for (int64_t i = a + 1; i < e; ++i) {
    Step<R, NX, NY> st_;
    R rm[NX], rL[NX][NX];
    if (nom_m != nullptr) {
      ld_aos(nom_m, i - 1, rm);
      ld_aos(nom_L, i - 1, rL);
    } else {
      copy(m, rm);
      copy(L, rL);
    }
    tr.row = i - 1;
    linearize_point_w<R, NX, NX, SQRT>(tr, sp, use_sigma, rm, rL, st_.F, st_.cq, st_.bt, fb);
    if (nom_m != nullptr) {
      ld_aos(nom_m, i, rm);
      ld_aos(nom_L, i, rL);
    } else {
      seq_predict<R, NX, SQRT>(st_.F, st_.cq, st_.bt, m, L, rm, rL);
    }
    ob.row = i;
    linearize_point_w<R, NX, NY, SQRT>(ob, sp, use_sigma, rm, rL, st_.H, st_.cr, st_.c, fb);
    R y[NY];
    ld_aos(ys, i - s - 1, y);
    acc += kf_step_fb<R, NX, NY, SQRT>(st_, y, m, L, fb);
    st_aos(out_m, i, m);
    st_aos(out_L, i, L);
  }
  if (ell != nullptr) ell[s] = acc;
}

// Lane s smooths sequence s backwards from its last filtered state (the
// Rauch-Tung-Striebel recursion as a smoother element applied to the
// smoothed successor).  nom_m == nullptr: the transition out of state i is
// linearized at the smoothed state i+1.
template <typename R, int NX, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_seq_smoother(
    int nseq, const int64_t* __restrict__ soff, ModelView<R> tr, SigmaView<R> sp, int use_sigma,
    const R* __restrict__ nom_m, const R* __restrict__ nom_L, const R* __restrict__ fm,
    const R* __restrict__ fL, R* __restrict__ out_m, R* __restrict__ out_L, Fallback fb) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseq) return;
  const int64_t a = soff[s], e = soff[s + 1];
  R m[NX], D[NX][NX];
  ld_aos(fm, e - 1, m);
  ld_aos(fL, e - 1, D);
  st_aos(out_m, e - 1, m);
  st_aos(out_L, e - 1, D);
  for (int64_t i = e - 2; i >= a; --i) {
    R rm[NX], rL[NX][NX];
    if (nom_m != nullptr) {
      ld_aos(nom_m, i, rm);
      ld_aos(nom_L, i, rL);
    } else {
      copy(m, rm);
      copy(D, rL);
    }
    R F[NX][NX], cq[NX][NX], bt[NX];
    tr.row = i;
    linearize_point_w<R, NX, NX, SQRT>(tr, sp, use_sigma, rm, rL, F, cq, bt, fb);
    R xm[NX], xL[NX][NX];
    ld_aos(fm, i, xm);
    ld_aos(fL, i, xL);
    SAgg<R, NX> el;
    smoother_element_fb<R, NX, SQRT>(F, cq, bt, xm, xL, el, fb);
    smoother_apply_fb<R, NX, SQRT>(el, m, D, fb);
    st_aos(out_m, i, m);
    st_aos(out_L, i, D);
  }
}

}  // namespace psm
