// Kernels and host launchers of the parallel square-root smoother, templated
// on the real type, the state/observation dimensions and the form
// (square-root or covariance).
//
// Data layout in HBM
//  * Caller-visible trajectories are array-of-structs, exactly the shapes of the
//    reference: mean (n_states, nx), chol/cov (n_states, nx, nx).
//  * Internal per-state buffers (linearized model, smoother elements) and the
//    per-chunk aggregates are struct-of-arrays: component q of item i at
//    base[q * n_items + i], so a wave touching consecutive items reads
//    consecutive addresses.
//
// One filter pass over a ragged batch (all sequences concatenated, n_states
// states) is
//   1. k_linearize_states   one lane per state: transition and observation
//                           linearized at the nominal state
//   2. k_filter_reduce      one lane per chunk of C states: the chunk aggregate
//                           by absorbing its steps one by one
//   3. k_scan_up/k_scan_down exclusive scan of the chunk aggregates with the
//                           full filtering operator (Kogge-Stone inside a
//                           workgroup, recursive across workgroups)
//   4. k_filter_apply       one lane per chunk: the sequential square-root
//                           Kalman filter from the chunk's incoming filtered
//                           state, writing the filtered states and the
//                           per-step log-likelihood terms
// and the smoother pass mirrors it backwards in time (elements, reverse scan,
// application).  Because the first state of every sequence carries its prior
// (and the last state of every sequence is the smoother's terminal element),
// both scans run over the concatenation without segment bookkeeping.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "psm_core.hpp"
#include "psm_scan_mem.hpp"

namespace psm {

constexpr int kScanWG = 64;     // one wave per scan workgroup
constexpr int kLaneWG = 64;     // workgroup size of the lane-per-item kernels

PSM_DEV int seq_of(const int64_t* __restrict__ soff, int nseq, int64_t i) {
  int lo = 0, hi = nseq - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (soff[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <typename R, int N>
PSM_DEV void ld_aos(const R* __restrict__ p, int64_t i, R (&v)[N]) {
#pragma unroll
  for (int a = 0; a < N; ++a) v[a] = p[i * N + a];
}
template <typename R, int M, int N>
PSM_DEV void ld_aos(const R* __restrict__ p, int64_t i, R (&v)[M][N]) {
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) v[a][b] = p[i * M * N + a * N + b];
}
template <typename R, int N>
PSM_DEV void st_aos(R* __restrict__ p, int64_t i, const R (&v)[N]) {
#pragma unroll
  for (int a = 0; a < N; ++a) p[i * N + a] = v[a];
}
template <typename R, int M, int N>
PSM_DEV void st_aos(R* __restrict__ p, int64_t i, const R (&v)[M][N]) {
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) p[i * M * N + a * N + b] = v[a][b];
}

struct Lin {
  int use_sigma;
};

// ============================================================================
// 1. linearization at every nominal state
// ============================================================================
template <typename R, int NIN, int NOUT, bool SQRT>
PSM_DEV void linearize_point(const ModelView<R>& md, const SigmaView<R>& sp, int use_sigma,
                             const R (&m)[NIN], const R (&Lx)[NIN][NIN], R (&F)[NOUT][NIN],
                             R (&C)[NOUT][NOUT], R (&b)[NOUT]) {
  if (use_sigma) {
    if constexpr (SQRT) {
      linearize_sigma<R, NIN, NOUT, true>(md, sp, m, Lx, F, C, b);
    } else {
      R Lc[NIN][NIN];
      cholesky(Lx, Lc);   // get_mvnsqrt (parsmooth/linearization/_common.py:9-14)
      linearize_sigma<R, NIN, NOUT, false>(md, sp, m, Lc, F, C, b);
    }
  } else {
    linearize_extended<R, NIN, NOUT>(md, m, F, C, b);
  }
}

template <typename R, int NX, int NY, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_linearize_states(
    int64_t M, ModelView<R> tr, ModelView<R> ob, int do_obs, SigmaView<R> sp, int use_sigma,
    const R* __restrict__ nom_m, const R* __restrict__ nom_L, R* __restrict__ lin) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  R m[NX], L[NX][NX];
  ld_aos(nom_m, i, m);
  ld_aos(nom_L, i, L);
  {
    R F[NX][NX], C[NX][NX], b[NX];
    linearize_point<R, NX, NX, SQRT>(tr, sp, use_sigma, m, L, F, C, b);
    int q = 0;
    st(lin, M, i, q, F);
    st(lin, M, i, q, C);
    st(lin, M, i, q, b);
  }
  if (do_obs) {
    R H[NY][NX], C[NY][NY], c[NY];
    linearize_point<R, NX, NY, SQRT>(ob, sp, use_sigma, m, L, H, C, c);
    int q = Step<R, NX, NY>::NTRANS;
    st(lin, M, i, q, H);
    st(lin, M, i, q, C);
    st(lin, M, i, q, c);
  }
}

// ============================================================================
// 2. chunk aggregates of the filter
// ============================================================================
template <typename R, int NX, int NY, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_filter_reduce(
    int64_t M, int chunk, int64_t nchunks, const int64_t* __restrict__ soff, int nseq,
    const R* __restrict__ lin, const R* __restrict__ ys, const R* __restrict__ m0,
    const R* __restrict__ L0, R* __restrict__ agg, Fallback fb) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int64_t i0 = c * chunk;
  const int64_t i1 = min(i0 + chunk, M);
  int s = seq_of(soff, nseq, i0);
  FAgg<R, NX> a;
  set_identity(a);
  for (int64_t i = i0; i < i1; ++i) {
    while (i >= soff[s + 1]) ++s;
    if (i == soff[s]) {
      set_zero(a.A);
      ld_aos(m0, s, a.b);
      ld_aos(L0, s, a.U);
      set_zero(a.eta);
      set_zero(a.Z);
      a.flag = FLAG_HEAD;
    } else {
      Step<R, NX, NY> st_;
      load_trans(lin, M, i - 1, st_);
      load_obs(lin, M, i, st_);
      R y[NY];
      ld_aos(ys, i - s - 1, y);
      absorb_fb<R, NX, NY, SQRT>(a, st_, y, fb);
    }
  }
  store_agg(agg, nchunks, c, a);
}

// ============================================================================
// 3. exclusive scan of aggregates (filter: forward, smoother: reverse)
// ============================================================================
// in/out: n items (SoA, stride n).  Item index of scan position p is
// (reverse ? n - 1 - p : p).  Workgroup w covers positions [w*WG*K, (w+1)*WG*K):
// every lane folds its K items, a Kogge-Stone scan over the lanes runs in a
// double-buffered LDS array, and the lane then writes the exclusive prefix of
// each of its items (within the workgroup) to out_excl; totals[w] receives the
// workgroup's inclusive total.  Aggregates live in LDS / HBM, never whole in
// registers: the operator reads its operands from memory as it needs them.
template <typename Agg, typename R, bool SQRT>
__global__ __launch_bounds__(kScanWG) void k_scan_up(const R* __restrict__ in, int64_t n, int K,
                                                     int reverse, R* __restrict__ out_excl,
                                                     R* __restrict__ totals, int64_t ntot,
                                                     Fallback fb) {
  using V = typename ViewOf<Agg>::type;
  constexpr int NX = Agg_nx<Agg>::value;
  constexpr int NQ = Agg::NQ;
  __shared__ R lds[2 * NQ * kScanWG];
  const int t = threadIdx.x;
  auto slot = [&](int buf, int lane) { return V{lds + buf * NQ * kScanWG, (int64_t)kScanWG, (int64_t)lane}; };
  R* inm = const_cast<R*>(in);
  const int64_t base = ((int64_t)blockIdx.x * kScanWG + t) * K;
  int cur = 0;
  view_identity(slot(0, t));
  for (int j = 0; j < K; ++j) {
    const int64_t p = base + j;
    if (p >= n) break;
    const V e{inm, n, reverse ? n - 1 - p : p};
    combine_mem_fb<R, NX, SQRT>(slot(cur, t), e, slot(1 - cur, t), fb);
    cur = 1 - cur;
  }
  if (cur == 1) view_copy(slot(1, t), slot(0, t));
  cur = 0;
  for (int d = 1; d < kScanWG; d <<= 1) {
    __syncthreads();
    if (t >= d) combine_mem_fb<R, NX, SQRT>(slot(cur, t - d), slot(cur, t), slot(1 - cur, t), fb);
    else view_copy(slot(cur, t), slot(1 - cur, t));
    cur = 1 - cur;
  }
  __syncthreads();
  if (t == kScanWG - 1 && totals != nullptr)
    view_copy(slot(cur, t), V{totals, ntot, (int64_t)blockIdx.x});
  // running exclusive prefix in this lane's slot of the other buffer
  const V run = slot(1 - cur, t);
  if (t == 0) view_identity(run);
  else view_copy(slot(cur, t - 1), run);
  for (int j = 0; j < K; ++j) {
    const int64_t p = base + j;
    if (p >= n) break;
    const int64_t item = reverse ? n - 1 - p : p;
    const V o{out_excl, n, item};
    view_copy(run, o);
    combine_mem_fb<R, NX, SQRT>(o, V{inm, n, item}, run, fb);
  }
}

// dst[item] <- upper[w] (+) out_excl[item] (workgroup w > 0; w == 0 copies):
// the complete exclusive prefixes, written out of place so the operands stay
// intact for the FP64 fallback.
template <typename Agg, typename R, bool SQRT>
__global__ __launch_bounds__(kScanWG) void k_scan_down(const R* __restrict__ out_excl, int64_t n,
                                                       int K, int reverse,
                                                       const R* __restrict__ upper,
                                                       int64_t nupper, R* __restrict__ dst,
                                                       Fallback fb) {
  using V = typename ViewOf<Agg>::type;
  constexpr int NX = Agg_nx<Agg>::value;
  const int t = threadIdx.x;
  const int64_t base = ((int64_t)blockIdx.x * kScanWG + t) * K;
  const V up{const_cast<R*>(upper), nupper, (int64_t)blockIdx.x};
  for (int j = 0; j < K; ++j) {
    const int64_t p = base + j;
    if (p >= n) break;
    const int64_t item = reverse ? n - 1 - p : p;
    const V e{const_cast<R*>(out_excl), n, item};
    const V o{dst, n, item};
    if (blockIdx.x == 0) view_copy(e, o);
    else combine_mem_fb<R, NX, SQRT>(up, e, o, fb);
  }
}

// ============================================================================
// 4. filter application
// ============================================================================
template <typename R, int NX, int NY, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_filter_apply(
    int64_t M, int chunk, int64_t nchunks, const int64_t* __restrict__ soff, int nseq,
    const R* __restrict__ lin, const R* __restrict__ ys, const R* __restrict__ m0,
    const R* __restrict__ L0, const R* __restrict__ prefix, R* __restrict__ out_m,
    R* __restrict__ out_L, double* __restrict__ ell_state, Fallback fb) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int64_t i0 = c * chunk;
  const int64_t i1 = min(i0 + chunk, M);
  int s = seq_of(soff, nseq, i0);
  R m[NX], L[NX][NX];
  {
    // incoming filtered state: (b, U) of the exclusive prefix
    int q = NX * NX;
    ld(prefix, nchunks, c, q, m);
    ld(prefix, nchunks, c, q, L);
  }
  for (int64_t i = i0; i < i1; ++i) {
    while (i >= soff[s + 1]) ++s;
    double e = 0.0;
    if (i == soff[s]) {
      ld_aos(m0, s, m);
      ld_aos(L0, s, L);
    } else {
      Step<R, NX, NY> st_;
      load_trans(lin, M, i - 1, st_);
      load_obs(lin, M, i, st_);
      R y[NY];
      ld_aos(ys, i - s - 1, y);
      e = kf_step_fb<R, NX, NY, SQRT>(st_, y, m, L, fb);
    }
    st_aos(out_m, i, m);
    st_aos(out_L, i, L);
    if (ell_state != nullptr) ell_state[i] = e;
  }
}

// per-sequence sum of the per-state log-likelihood terms (FP64)
template <int BS>
__global__ __launch_bounds__(BS) void k_ell_reduce(const double* __restrict__ ell_state,
                                                    const int64_t* __restrict__ soff,
                                                    double* __restrict__ ell) {
  __shared__ double red[BS];
  const int s = blockIdx.x;
  const int64_t a = soff[s], b = soff[s + 1];
  double acc = 0.0;
  for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) acc += ell_state[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) ell[s] = red[0];
}

// ============================================================================
// smoother passes
// ============================================================================
template <typename R, int NX, int NY, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_smoother_reduce(
    int64_t M, int chunk, int64_t nchunks, const int64_t* __restrict__ soff, int nseq,
    const R* __restrict__ lin, const R* __restrict__ fm, const R* __restrict__ fL,
    R* __restrict__ elems, R* __restrict__ agg, Fallback fb) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int64_t i0 = c * chunk;
  const int64_t i1 = min(i0 + chunk, M);
  int s = seq_of(soff, nseq, i1 - 1);
  SAgg<R, NX> acc;
  set_identity(acc);
  for (int64_t i = i1 - 1; i >= i0; --i) {
    while (i < soff[s]) --s;
    R m[NX], L[NX][NX];
    ld_aos(fm, i, m);
    ld_aos(fL, i, L);
    SAgg<R, NX> e;
    if (i == soff[s + 1] - 1) {
      copy(m, e.g);
      set_zero(e.E);
      copy(L, e.D);
      e.flag = FLAG_HEAD;
    } else {
      R F[NX][NX], cq[NX][NX], bt[NX];
      int q = 0;
      ld(lin, M, i, q, F);
      ld(lin, M, i, q, cq);
      ld(lin, M, i, q, bt);
      smoother_element_fb<R, NX, SQRT>(F, cq, bt, m, L, e, fb);
    }
    store_agg(elems, M, i, e);
    combine_into_fb<R, NX, SQRT>(acc, e, fb);
    copy_agg(e, acc);
  }
  store_agg(agg, nchunks, c, acc);
}

template <typename R, int NX, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_smoother_apply(
    int64_t M, int chunk, int64_t nchunks, const R* __restrict__ elems,
    const R* __restrict__ prefix, R* __restrict__ out_m, R* __restrict__ out_L, Fallback fb) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int64_t i0 = c * chunk;
  const int64_t i1 = min(i0 + chunk, M);
  R m[NX], D[NX][NX];
  {
    int q = 0;
    ld(prefix, nchunks, c, q, m);
    q += NX * NX;
    ld(prefix, nchunks, c, q, D);
  }
  for (int64_t i = i1 - 1; i >= i0; --i) {
    SAgg<R, NX> e;
    load_agg(elems, M, i, e);
    if (e.flag == FLAG_HEAD) {
      copy(e.g, m);
      copy(e.D, D);
    } else {
      smoother_apply_fb<R, NX, SQRT>(e, m, D, fb);
    }
    st_aos(out_m, i, m);
    st_aos(out_L, i, D);
  }
}

// ============================================================================
// operator-level batched kernels (array-of-structs in and out)
// ============================================================================
template <typename R, int NIN, int NOUT, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_linearize_points(
    int64_t n, ModelView<R> md, SigmaView<R> sp, int use_sigma, const R* __restrict__ xm,
    const R* __restrict__ xL, R* __restrict__ F, R* __restrict__ C, R* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R m[NIN], L[NIN][NIN], Fo[NOUT][NIN], Co[NOUT][NOUT], bo[NOUT];
  ld_aos(xm, i, m);
  ld_aos(xL, i, L);
  linearize_point<R, NIN, NOUT, SQRT>(md, sp, use_sigma, m, L, Fo, Co, bo);
  st_aos(F, i, Fo);
  st_aos(C, i, Co);
  st_aos(b, i, bo);
}

template <typename R, int NX, int NY, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_filter_elements(
    int64_t n, const R* __restrict__ F, const R* __restrict__ cq, const R* __restrict__ bt,
    const R* __restrict__ H, const R* __restrict__ cr, const R* __restrict__ c,
    const R* __restrict__ m, const R* __restrict__ L, const R* __restrict__ y, R* __restrict__ A,
    R* __restrict__ b, R* __restrict__ U, R* __restrict__ eta, R* __restrict__ Z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Step<R, NX, NY> s;
  ld_aos(F, i, s.F);
  ld_aos(cq, i, s.cq);
  ld_aos(bt, i, s.bt);
  ld_aos(H, i, s.H);
  ld_aos(cr, i, s.cr);
  ld_aos(c, i, s.c);
  R mm_[NX], LL[NX][NX], yy[NY];
  ld_aos(m, i, mm_);
  ld_aos(L, i, LL);
  ld_aos(y, i, yy);
  FAgg<R, NX> e;
  filter_element<R, NX, NY, SQRT>(s, mm_, LL, yy, e);
  st_aos(A, i, e.A);
  st_aos(b, i, e.b);
  st_aos(U, i, e.U);
  st_aos(eta, i, e.eta);
  st_aos(Z, i, e.Z);
}

template <typename R, int NX>
struct FPtrs {
  const R* p[5];
};
template <typename R, int NX>
struct FOut {
  R* p[5];
};

template <typename R, int NX, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_filter_combine(int64_t n, FPtrs<R, NX> a,
                                                            FPtrs<R, NX> b, FOut<R, NX> o) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  FAgg<R, NX> e1, e2;
  ld_aos(a.p[0], i, e1.A); ld_aos(a.p[1], i, e1.b); ld_aos(a.p[2], i, e1.U);
  ld_aos(a.p[3], i, e1.eta); ld_aos(a.p[4], i, e1.Z);
  ld_aos(b.p[0], i, e2.A); ld_aos(b.p[1], i, e2.b); ld_aos(b.p[2], i, e2.U);
  ld_aos(b.p[3], i, e2.eta); ld_aos(b.p[4], i, e2.Z);
  e1.flag = FLAG_NORMAL;
  e2.flag = FLAG_NORMAL;
  combine_into<R, NX, SQRT>(e1, e2);
  st_aos(o.p[0], i, e2.A); st_aos(o.p[1], i, e2.b); st_aos(o.p[2], i, e2.U);
  st_aos(o.p[3], i, e2.eta); st_aos(o.p[4], i, e2.Z);
}

template <typename R, int NX, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_smoother_elements(
    int64_t n, const R* __restrict__ F, const R* __restrict__ cq, const R* __restrict__ bt,
    const R* __restrict__ m, const R* __restrict__ L, R* __restrict__ g, R* __restrict__ E,
    R* __restrict__ D) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R Fi[NX][NX], cqi[NX][NX], bti[NX], mi[NX], Li[NX][NX];
  ld_aos(F, i, Fi);
  ld_aos(cq, i, cqi);
  ld_aos(bt, i, bti);
  ld_aos(m, i, mi);
  ld_aos(L, i, Li);
  SAgg<R, NX> e;
  smoother_element<R, NX, SQRT>(Fi, cqi, bti, mi, Li, e);
  st_aos(g, i, e.g);
  st_aos(E, i, e.E);
  st_aos(D, i, e.D);
}

template <typename R, int NX>
struct SPtrs {
  const R* p[3];
};
template <typename R, int NX>
struct SOut {
  R* p[3];
};

template <typename R, int NX, bool SQRT>
__global__ __launch_bounds__(kLaneWG) void k_smoother_combine(int64_t n, SPtrs<R, NX> a,
                                                              SPtrs<R, NX> b, SOut<R, NX> o) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SAgg<R, NX> e1, e2;
  ld_aos(a.p[0], i, e1.g); ld_aos(a.p[1], i, e1.E); ld_aos(a.p[2], i, e1.D);
  ld_aos(b.p[0], i, e2.g); ld_aos(b.p[1], i, e2.E); ld_aos(b.p[2], i, e2.D);
  e1.flag = FLAG_NORMAL;
  e2.flag = FLAG_NORMAL;
  combine_into<R, NX, SQRT>(e1, e2);
  st_aos(o.p[0], i, e2.g); st_aos(o.p[1], i, e2.E); st_aos(o.p[2], i, e2.D);
}

}  // namespace psm
