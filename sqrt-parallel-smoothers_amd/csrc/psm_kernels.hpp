// Kernels and host launchers of the parallel square-root smoother, templated
// on the real type, the state/observation dimensions and the form
// (square-root or covariance).
//
// Data layout in HBM
//  * Caller-visible trajectories are array-of-structs, exactly the shapes of the
//    reference: mean (n_states, nx), chol/cov (n_states, nx, nx).
//  * Internal per-state buffers (linearized model, smoother elements) and the
//    scan items are struct-of-arrays, component q of an item at
//    base[q * stride + position], with positions interleaved (StateMap,
//    LevelMap) so that the items a wave works on in one step are consecutive
//    words.
//
// One filter pass over a ragged batch (all sequences concatenated, n_states
// states) is
//   1. k_linearize_states   one lane per state: transition and observation
//                           linearized at the nominal state
//   2. k_filter_reduce      one lane per chunk of C states: the chunk aggregate
//                           by absorbing its steps one by one
//   3. k_scan_up/k_scan_down exclusive scan of the chunk aggregates with the
//                           full filtering operator (Kogge-Stone inside a
//                           workgroup, recursive across workgroups; no
//                           level-0 down-sweep, see 4)
//   4. k_filter_apply       one lane per chunk: joins the chunk's level-0
//                           prefix with its workgroup's prefix, then runs the
//                           sequential square-root Kalman filter from that
//                           incoming filtered state, writing the filtered
//                           states and the per-step log-likelihood terms
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
#include "psm_team.hpp"
#include "psm_quad.hpp"

namespace psm {

constexpr int kScanWG = 64;     // one wave per scan workgroup
// Scan workgroup size for an aggregate type: one wave, or half a wave when three
// LDS buffers of 64 aggregates would not fit the 160 KB of LDS (tangent types).
template <typename Agg, typename R>
constexpr int scan_wg() {
  return (3 * Agg::NQ * 64 * (int)sizeof(R) <= 150 * 1024) ? 64 : 32;
}
constexpr int kLaneWG = 64;     // workgroup size of the lane-per-item kernels

// Positions in the internal struct-of-arrays buffers.  The lane-per-chunk
// kernels run 64 consecutive chunks in one wave and step through their C
// states in lock step, so slot k of chunk c is stored at
//   (c / 64) * 64 C + 64 k + c % 64,
// and one component of one slot is 64 consecutive words for the wave (one
// full-line load instead of a C-strided gather).
struct StateMap {
  int C;
  PSM_DEV int64_t operator()(int64_t i) const {
    const int64_t c = i / C;
    const int64_t k = i - c * C;
    return (c >> 6) * ((int64_t)C << 6) + (k << 6) + (c & 63);
  }
};
inline int64_t state_stride(int64_t M, int C) {
  const int64_t g = (int64_t)C * 64;
  return (M + g - 1) / g * g;
}

// A scan level: n items in scan order, K per lane, WG lanes per workgroup.
// Scan position p of lane l, fold step j (p = blk * WG K + l K + j) is stored at
// blk * WG K + j WG + l, so every fold step of a workgroup touches WG
// consecutive words.  Level 0 holds the chunk aggregates: the smoother scans
// them in reverse (rev = 1), chunk c sitting at scan position n - 1 - c.
struct LevelMap {
  int64_t n;       // items
  int64_t stride;  // padded item count (multiple of WG K): component stride
  int K;
  int WG;
  int rev;
  PSM_DEV int64_t at(int64_t p) const {
    const int64_t blk = (int64_t)K * WG;
    const int64_t b = p / blk;
    const int64_t r = p - b * blk;
    const int64_t l = r / K;
    return b * blk + (r - l * K) * WG + l;
  }
  PSM_DEV int64_t pos(int64_t c) const { return at(rev ? n - 1 - c : c); }
};

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

// Trajectories (a mean and a factor per state) come in two layouts: the API's
// array-of-structs (state i at m + i NX, L + i NX^2; inter == 0), or, inside a
// launch sequence that owns them, the state-interleaved layout of the
// workspace (inter != 0: component q of state i at m[q Mp + pos], pos its
// interleaved position, means first, then a factor's lower triangle or a
// covariance in full (SYM), so that a launch sequence and the host loop over
// the same passes agree bit for bit), which lane-per-chunk kernels read and
// write coalesced.
template <bool SYM, typename R, int NX>
PSM_DEV void traj_load(const R* __restrict__ m_, const R* __restrict__ L_, int inter, int64_t Mp,
                       int64_t i, int64_t pos, R (&m)[NX], R (&L)[NX][NX]) {
  if (inter) {
    int q = 0;
    ld(m_, Mp, pos, q, m);
    if constexpr (SYM) ld(m_, Mp, pos, q, L);
    else ld_tri<false>(m_, Mp, pos, q, L);
  } else {
    ld_aos(m_, i, m);
    ld_aos(L_, i, L);
  }
}
template <bool SYM, typename R, int NX>
PSM_DEV void traj_store(R* __restrict__ m_, R* __restrict__ L_, int inter, int64_t Mp, int64_t i,
                        int64_t pos, const R (&m)[NX], const R (&L)[NX][NX]) {
  if (inter) {
    int q = 0;
    st(m_, Mp, pos, q, m);
    if constexpr (SYM) st(m_, Mp, pos, q, L);
    else st_tri(m_, Mp, pos, q, L);
  } else {
    st_aos(m_, i, m);
    st_aos(L_, i, L);
  }
}


struct Lin {
  int use_sigma;
};

// p[0..n) <- v (flag words and counters; a kernel rather than hipMemsetAsync,
// so that a captured graph replays it like any other launch)
template <typename T>
__global__ void k_fill(T* __restrict__ p, int n, T v) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = v;
}

// ============================================================================
// 1. linearization at every nominal state
// ============================================================================
template <typename R, int NIN, int NOUT, bool SQRT, typename S>
PSM_DEV void linearize_point(const ModelView<R, S>& md, const SigmaView<R>& sp, int use_sigma,
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

// Linearization of the single-precision paths.  In FP32 arithmetic the
// textbook formulas cancel: the sigma-point regression subtracts F P F^T from
// the sigma-point covariance, which is ~0.1 against a process noise Q ~ 3e-9
// on the bearings-only model; the coordinated turn's turn-rate derivatives are
// differences of two numbers ~dt that agree to O((w dt)^2); b = f(m) - F m
// cancels at large |m|.  The FP32 path therefore uses the cancellation-free
// forms of psm_models.hpp (centred residuals, series in w dt, closed-form
// offsets), all in FP32.  Rules with a negative weight and conditional-moments
// models keep the reference's downdate form, evaluated in FP64 from the FP32
// inputs and rounded back.  In square-root form, a regression noise factor
// whose min|diag| / max|diag| falls below the fallback threshold is recomputed
// that way too (and counted), like every other triangular factor.
template <typename R> struct WiderOf { using type = R; };
template <> struct WiderOf<float> { using type = double; };
template <> struct WiderOf<DualT<float>> { using type = DualT<double>; };

template <typename A, int N, typename B>
PSM_DEV void widen(const B (&a)[N], A (&b)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) b[i] = A(a[i]);
}
template <typename A, int M, int N, typename B>
PSM_DEV void widen(const B (&a)[M][N], A (&b)[M][N]) {
#pragma unroll
  for (int i = 0; i < M; ++i) widen(a[i], b[i]);
}

template <typename R, int NIN, int NOUT, bool SQRT>
PSM_DEV void linearize_point_wide(const ModelView<R>& md, const SigmaView<R>& sp, int use_sigma,
                                  const R (&m)[NIN], const R (&Lx)[NIN][NIN], R (&F)[NOUT][NIN],
                                  R (&C)[NOUT][NOUT], R (&b)[NOUT]) {
  using W = typename WiderOf<R>::type;
  const ModelView<W, R> mw{md.kind, md.params, md.q_mean, md.q_chol, md.row, md.pt, md.npts};
  const SigmaView<W> sw{sp.npts, sp.xi, sp.wm, sp.wc};
  W mW[NIN], LW[NIN][NIN], FW[NOUT][NIN], CW[NOUT][NOUT], bW[NOUT];
  widen(m, mW);
  widen(Lx, LW);
  linearize_point<W, NIN, NOUT, SQRT>(mw, sw, use_sigma, mW, LW, FW, CW, bW);
  widen(FW, F);
  widen(CW, C);
  widen(bW, b);
}

template <typename R, int NIN, int NOUT, bool SQRT, typename FB>
PSM_DEV void linearize_point_w(const ModelView<R>& md, const SigmaView<R>& sp, int use_sigma,
                               const R (&m)[NIN], const R (&Lx)[NIN][NIN], R (&F)[NOUT][NIN],
                               R (&C)[NOUT][NOUT], R (&b)[NOUT], const FB& fb,
                               const PointCache<R>& pc = PointCache<R>{nullptr, 0, 0}) {
  if constexpr (kFallback<R, SQRT> && !FB::kRedo) {
    // FP32 square-root pass without FP64 code: a point the centred regression
    // cannot take (conditional moments, a negative weight) or whose factor is
    // ill-conditioned raises the flag, and the guarded second launch
    // linearizes every point again with the FP64 path below
    if (!use_sigma) {
      linearize_extended<R, NIN, NOUT>(md, m, F, C, b);
      return;
    }
    const bool done = md.kind != PSM_MODEL_CONDITIONAL &&
                      linearize_sigma_centred<R, NIN, NOUT, true>(md, sp, m, Lx, F, C, b, pc);
    if (!done || (fb.thr > 0.0f && ill_conditioned(C, fb.thr))) fb.raise();
  } else if constexpr (IsSingle<R>::value) {
    if (!use_sigma) {
      linearize_extended<R, NIN, NOUT>(md, m, F, C, b);
      return;
    }
    bool done = false;
    if (md.kind != PSM_MODEL_CONDITIONAL) {
      if constexpr (SQRT) {
        done = linearize_sigma_centred<R, NIN, NOUT, true>(md, sp, m, Lx, F, C, b, pc);
        if constexpr (kFallback<R, SQRT>) {
          if (done && fb.thr > 0.0f && ill_conditioned(C, fb.thr)) {
            count_fallback(fb);
            done = false;
          }
        }
      } else {
        R Lc[NIN][NIN];
        cholesky(Lx, Lc);   // get_mvnsqrt (parsmooth/linearization/_common.py:9-14)
        done = linearize_sigma_centred<R, NIN, NOUT, false>(md, sp, m, Lc, F, C, b, pc);
      }
    }
    if (!done) linearize_point_wide<R, NIN, NOUT, SQRT>(md, sp, use_sigma, m, Lx, F, C, b);
  } else {
    linearize_point<R, NIN, NOUT, SQRT>(md, sp, use_sigma, m, Lx, F, C, b);
  }
}

// Thread t fills position t of the state-interleaved buffer (coalesced stores);
// its state is the inverse of StateMap.
template <typename R, int NX, int NY, bool SQRT, typename FB = Fallback>
__global__ __launch_bounds__(kLaneWG) void k_linearize_states(
    int64_t M, int64_t Mp, int C, ModelView<R> tr, ModelView<R> ob, int do_obs, SigmaView<R> sp,
    int use_sigma, const R* __restrict__ nom_m, const R* __restrict__ nom_L, int nom_inter,
    R* __restrict__ lin, FB fb) {
  if (fb.skip()) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mp) return;
  const int64_t g = (int64_t)C << 6;
  const int64_t blk = t / g;
  const int64_t r = t - blk * g;
  const int64_t i = ((blk << 6) + (r & 63)) * C + (r >> 6);
  if (i >= M) return;
  tr.row = i;
  ob.row = i;
  R m[NX], L[NX][NX];
  traj_load<!SQRT, R, NX>(nom_m, nom_L, nom_inter, Mp, i, t, m, L);
  // FP32: the regression's increments wait in LDS for its residual pass
  PointCache<R> pc{nullptr, 0, 0};
  if constexpr (std::is_same<R, float>::value) {
    __shared__ float fcache[(2 * NX + 1) * NX * kLaneWG];
    pc = PointCache<R>{fcache + threadIdx.x, kLaneWG, 2 * NX + 1};
  }
  {
    R F[NX][NX], Cq[NX][NX], b[NX];
    linearize_point_w<R, NX, NX, SQRT>(tr, sp, use_sigma, m, L, F, Cq, b, fb, pc);
    int q = 0;
    st(lin, Mp, t, q, F);
    if constexpr (SQRT) lower_factor(Cq);
    st_tri(lin, Mp, t, q, Cq);
    st(lin, Mp, t, q, b);
  }
  if (do_obs) {
    R H[NY][NX], Cr[NY][NY], c[NY];
    linearize_point_w<R, NX, NY, SQRT>(ob, sp, use_sigma, m, L, H, Cr, c, fb);
    int q = Step<R, NX, NY>::NTRANS;
    st(lin, Mp, t, q, H);
    if constexpr (SQRT) lower_factor(Cr);
    st_tri(lin, Mp, t, q, Cr);
    st(lin, Mp, t, q, c);
  }
}

// ============================================================================
// 2. chunk aggregates of the filter
// ============================================================================
template <typename R, int NX, int NY, bool SQRT, typename FB = Fallback>
__global__ __launch_bounds__(kLaneWG) void k_filter_reduce(
    int64_t M, int64_t Mp, int chunk, int64_t nchunks, const int64_t* __restrict__ soff, int nseq,
    const R* __restrict__ lin, const R* __restrict__ ys, const R* __restrict__ m0,
    const R* __restrict__ L0, R* __restrict__ agg, LevelMap lm, FB fb) {
  if (fb.skip()) return;
  const StateMap sm{chunk};
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
      load_trans<!SQRT>(lin, Mp, sm(i - 1), st_);
      load_obs<!SQRT>(lin, Mp, sm(i), st_);
      R y[NY];
      ld_aos(ys, i - s - 1, y);
      absorb_fb<R, NX, NY, SQRT>(a, st_, y, fb);
    }
  }
  store_agg(agg, lm.stride, lm.pos(c), a);
}

// ============================================================================
// 3. exclusive scan of aggregates (filter: forward, smoother: reverse)
// ============================================================================
// in/out: the items of one level (LevelMap: n items in scan order, stored
// interleaved).  Workgroup w covers scan positions [w*WG*K, (w+1)*WG*K): every
// lane folds its K items, a Kogge-Stone scan over the lanes runs in a
// double-buffered LDS array, and the lane then writes the exclusive prefix of
// each of its items (within the workgroup) to out_excl; the workgroup's
// inclusive total becomes item w of the next level (tm).  Aggregates live in
// LDS / HBM, never whole in registers: the operator reads its operands from
// memory as it needs them.
template <typename Agg, typename R, bool SQRT, int WG = scan_wg<Agg, R>(), typename FB = Fallback>
__global__ __launch_bounds__(WG) void k_scan_up(const R* __restrict__ in, LevelMap lm,
                                                R* __restrict__ out_excl, R* __restrict__ totals,
                                                LevelMap tm, FB fb) {
  if (fb.skip()) return;
  using V = typename ViewOf<Agg>::type;
  constexpr int NX = Agg_nx<Agg>::value;
  constexpr int NQ = Agg::NQ;
  // buffers 0 and 1: running / scanned aggregates; buffer 2: the lane's next
  // input item, copied from HBM in one burst so that the combination reads
  // only LDS
  __shared__ R lds[3 * NQ * WG];
  const int t = threadIdx.x;
  const int K = lm.K;
  const int64_t n = lm.n;
  auto slot = [&](int buf, int lane) { return V{lds + buf * NQ * WG, (int64_t)WG, (int64_t)lane}; };
  R* inm = const_cast<R*>(in);
  const V stage = slot(2, t);
  const int64_t base = ((int64_t)blockIdx.x * WG + t) * K;
  int cur = 0;
  if (base < n) view_copy(V{inm, lm.stride, lm.at(base)}, slot(0, t));
  else view_identity(slot(0, t));
  for (int j = 1; j < K; ++j) {
    const int64_t p = base + j;
    if (p >= n) break;
    view_copy(V{inm, lm.stride, lm.at(p)}, stage);
    combine_mem_fb<R, NX, SQRT>(slot(cur, t), stage, slot(1 - cur, t), fb);
    cur = 1 - cur;
  }
  if (cur == 1) view_copy(slot(1, t), slot(0, t));
  cur = 0;
  for (int d = 1; d < WG; d <<= 1) {
    __syncthreads();
    if (t >= d) combine_mem_fb<R, NX, SQRT>(slot(cur, t - d), slot(cur, t), slot(1 - cur, t), fb);
    else view_copy(slot(cur, t), slot(1 - cur, t));
    cur = 1 - cur;
  }
  __syncthreads();
  if (t == WG - 1 && totals != nullptr)
    view_copy(slot(cur, t), V{totals, tm.stride, tm.at((int64_t)blockIdx.x)});
  // running exclusive prefix in this lane's slot of the other buffer; once
  // every lane has read its predecessor's inclusive prefix, slot(cur, t) is
  // free for the next running prefix
  V run = slot(1 - cur, t);
  if (t == 0) view_identity(run);
  else view_copy(slot(cur, t - 1), run);
  __syncthreads();
  V tmp = slot(cur, t);
  for (int j = 0; j < K; ++j) {
    const int64_t p = base + j;
    if (p >= n) break;
    const int64_t item = lm.at(p);
    view_copy(run, V{out_excl, lm.stride, item});
    // the running prefix after the lane's last item is its inclusive total,
    // which the scan over the lanes already produced
    if (j + 1 == K || p + 1 == n) break;
    view_copy(V{inm, lm.stride, item}, stage);
    combine_mem_fb<R, NX, SQRT>(run, stage, tmp, fb);
    const V x = run;
    run = tmp;
    tmp = x;
  }
}

// dst[p] <- upper[w] (+) out_excl[p] (workgroup w > 0; w == 0 copies): the
// complete exclusive prefixes, written out of place so the operands stay
// intact for the FP64 fallback.  upper holds the complete exclusive prefixes of
// the next level (um).  The combinations are independent, so every thread
// takes one storage position q of the level (consecutive threads, consecutive
// positions) and recovers its scan position p and workgroup w from q.
template <typename Agg, typename R, bool SQRT, int WG = scan_wg<Agg, R>(), typename FB = Fallback>
__global__ __launch_bounds__(WG) void k_scan_down(const R* __restrict__ out_excl, LevelMap lm,
                                                  const R* __restrict__ upper, LevelMap um,
                                                  R* __restrict__ dst, FB fb) {
  if (fb.skip()) return;
  using V = typename ViewOf<Agg>::type;
  constexpr int NX = Agg_nx<Agg>::value;
  const int64_t q = (int64_t)blockIdx.x * WG + threadIdx.x;
  const int64_t span = (int64_t)lm.K * lm.WG;
  const int64_t w = q / span;
  const int64_t r = q - w * span;
  const int64_t p = w * span + (r % lm.WG) * lm.K + r / lm.WG;
  if (p >= lm.n) return;
  const V e{const_cast<R*>(out_excl), lm.stride, q};
  const V o{dst, lm.stride, q};
  if (w == 0) view_copy(e, o);
  else combine_mem_fb<R, NX, SQRT>(V{const_cast<R*>(upper), um.stride, um.at(w)}, e, o, fb);
}

// Team versions of the two scan kernels for the square-root form: kTeamVL
// teams of TL lanes per workgroup; each team is one virtual lane of the scan
// above (folds K items, joins the Kogge-Stone scan over the teams, rewrites
// its items' exclusive prefixes) and runs every combination cooperatively.
template <typename Agg, typename R>
struct TeamCfg {
  using V = typename ViewOf<Agg>::type;
  static constexpr int TL = kTeamLanes;
  static constexpr int VL = kTeamVL;
  static constexpr int WGT = VL * TL;
  static constexpr int SCR = team_scratch_bytes<V>();
};

template <typename Agg, typename R, typename FB = Fallback>
__global__ __launch_bounds__((TeamCfg<Agg, R>::WGT)) void k_scan_up_team(
    const R* __restrict__ in, LevelMap lm, R* __restrict__ out_excl, R* __restrict__ totals,
    LevelMap tm, FB fb) {
  if (fb.skip()) return;
  using Cfg = TeamCfg<Agg, R>;
  using V = typename Cfg::V;
  constexpr int NX = Agg_nx<Agg>::value;
  constexpr int NQ = Agg::NQ;
  constexpr int TL = Cfg::TL;
  constexpr int VL = Cfg::VL;
  __shared__ R lds[2 * NQ * VL];
  __shared__ __align__(16) char scr[VL * Cfg::SCR];
  const int t = (int)threadIdx.x / TL;
  const int r = (int)threadIdx.x % TL;
  void* my = scr + t * Cfg::SCR;
  auto slot = [&](int buf, int lane) { return V{lds + buf * NQ * VL, (int64_t)VL, (int64_t)lane}; };
  const int K = lm.K;
  const int64_t n = lm.n;
  R* inm = const_cast<R*>(in);
  const int64_t base = ((int64_t)blockIdx.x * VL + t) * K;
  int cur = 0;
  team_identity(slot(0, t), r);
  for (int j = 0; j < K; ++j) {
    const int64_t p = base + j;
    if (p >= n) break;
    team_combine<R, NX>(slot(cur, t), V{inm, lm.stride, lm.at(p)}, slot(1 - cur, t), r, my, fb);
    cur = 1 - cur;
  }
  if (cur == 1) {
    team_sync();
    team_copy(slot(1, t), slot(0, t), r);
  }
  cur = 0;
  for (int d = 1; d < VL; d <<= 1) {
    __syncthreads();
    if (t >= d) team_combine<R, NX>(slot(cur, t - d), slot(cur, t), slot(1 - cur, t), r, my, fb);
    else team_copy(slot(cur, t), slot(1 - cur, t), r);
    cur = 1 - cur;
  }
  __syncthreads();
  if (t == VL - 1 && totals != nullptr)
    team_copy(slot(cur, t), V{totals, tm.stride, tm.at((int64_t)blockIdx.x)}, r);
  // running exclusive prefix of this team: run = slot(1 - cur, t); once every
  // team has read its predecessor's inclusive prefix, slot(cur, t) is free
  V run = slot(1 - cur, t);
  if (t == 0) team_identity(run, r);
  else team_copy(slot(cur, t - 1), run, r);
  __syncthreads();
  V tmp = slot(cur, t);
  for (int j = 0; j < K; ++j) {
    const int64_t p = base + j;
    if (p >= n) break;
    const int64_t item = lm.at(p);
    team_sync();
    team_copy(run, V{out_excl, lm.stride, item}, r);
    if (j + 1 == K || p + 1 == n) break;  // the team's inclusive total is not needed
    team_combine<R, NX>(run, V{inm, lm.stride, item}, tmp, r, my, fb);
    const V x = run;
    run = tmp;
    tmp = x;
  }
}

// One item per team: block b covers workgroup b / K of the level and item
// b % K of each of its teams.
template <typename Agg, typename R, typename FB = Fallback>
__global__ __launch_bounds__((TeamCfg<Agg, R>::WGT)) void k_scan_down_team(
    const R* __restrict__ out_excl, LevelMap lm, const R* __restrict__ upper, LevelMap um,
    R* __restrict__ dst, FB fb) {
  if (fb.skip()) return;
  using Cfg = TeamCfg<Agg, R>;
  using V = typename Cfg::V;
  constexpr int NX = Agg_nx<Agg>::value;
  constexpr int TL = Cfg::TL;
  constexpr int VL = Cfg::VL;
  __shared__ __align__(16) char scr[VL * Cfg::SCR];
  const int t = (int)threadIdx.x / TL;
  const int r = (int)threadIdx.x % TL;
  void* my = scr + t * Cfg::SCR;
  const int K = lm.K;
  const int64_t w = (int64_t)blockIdx.x / K;
  const int j = (int)((int64_t)blockIdx.x - w * K);
  const int64_t p = (w * VL + t) * K + j;
  if (p >= lm.n) return;  // whole teams leave together; team_combine syncs only within the team
  const int64_t item = lm.at(p);
  const V e{const_cast<R*>(out_excl), lm.stride, item};
  const V o{dst, lm.stride, item};
  if (w == 0) team_copy(e, o, r);
  else team_combine<R, NX>(V{const_cast<R*>(upper), um.stride, um.at(w)}, e, o, r, my, fb);
}

// Quad version of k_scan_up for level 0 of the square-root scans: VL quads
// per workgroup; each quad is one virtual lane of the scan above (folds K
// items, joins the Kogge-Stone scan over the quads, rewrites its items'
// exclusive prefixes) and runs every combination cooperatively.  The level
// layout is that of k_scan_up with VL lanes, so the two are interchangeable.
template <typename Agg, typename R, int VL, typename FB = Fallback>
__global__ __launch_bounds__(VL * kQuadLanes) void k_scan_up_quad(
    const R* __restrict__ in, LevelMap lm, R* __restrict__ out_excl, R* __restrict__ totals,
    LevelMap tm, FB fb) {
  if (fb.skip()) return;
  using V = typename ViewOf<Agg>::type;
  constexpr int NX = Agg_nx<Agg>::value;
  constexpr int NQ = Agg::NQ;
  // buffers 0 and 1: running / scanned aggregates; buffer 2: the quad's next
  // input item, copied from HBM before it is combined
  __shared__ R lds[3 * NQ * VL];
  const int t = (int)threadIdx.x / kQuadLanes;
  const int r = (int)threadIdx.x % kQuadLanes;
  auto slot = [&](int buf, int lane) { return V{lds + buf * NQ * VL, (int64_t)VL, (int64_t)lane}; };
  R* inm = const_cast<R*>(in);
  const V stage = slot(2, t);
  const int K = lm.K;
  const int64_t n = lm.n;
  const int64_t base = ((int64_t)blockIdx.x * VL + t) * K;
  int cur = 0;
  if (base < n) quad_copy(V{inm, lm.stride, lm.at(base)}, slot(0, t), r);
  else quad_identity(slot(0, t), r);
  for (int j = 1; j < K; ++j) {
    const int64_t p = base + j;
    if (p >= n) break;
    team_sync();
    quad_copy(V{inm, lm.stride, lm.at(p)}, stage, r);
    quad_combine<R, NX>(slot(cur, t), stage, slot(1 - cur, t), r, fb);
    cur = 1 - cur;
  }
  if (cur == 1) {
    team_sync();
    quad_copy(slot(1, t), slot(0, t), r);
  }
  cur = 0;
  for (int d = 1; d < VL; d <<= 1) {
    __syncthreads();
    if (t >= d) quad_combine<R, NX>(slot(cur, t - d), slot(cur, t), slot(1 - cur, t), r, fb);
    else quad_copy(slot(cur, t), slot(1 - cur, t), r);
    cur = 1 - cur;
  }
  __syncthreads();
  if (t == VL - 1 && totals != nullptr)
    quad_copy(slot(cur, t), V{totals, tm.stride, tm.at((int64_t)blockIdx.x)}, r);
  // running exclusive prefix of this quad: run = slot(1 - cur, t); once every
  // quad has read its predecessor's inclusive prefix, slot(cur, t) is free
  V run = slot(1 - cur, t);
  if (t == 0) quad_identity(run, r);
  else quad_copy(slot(cur, t - 1), run, r);
  __syncthreads();
  V tmp = slot(cur, t);
  for (int j = 0; j < K; ++j) {
    const int64_t p = base + j;
    if (p >= n) break;
    const int64_t item = lm.at(p);
    team_sync();
    quad_copy(run, V{out_excl, lm.stride, item}, r);
    if (j + 1 == K || p + 1 == n) break;  // the quad's inclusive total is not needed
    quad_copy(V{inm, lm.stride, item}, stage, r);
    quad_combine<R, NX>(run, stage, tmp, r, fb);
    const V x = run;
    run = tmp;
    tmp = x;
  }
}

// ============================================================================
// 4. filter application
// ============================================================================
// Complete exclusive prefix of level-0 item c (chunk c, in scan order when
// lm.rev).  The level-0 up-sweep leaves each item's prefix within its
// workgroup; items of workgroup w > 0 still need the prefix of the workgroups
// before them (level 1: upper[w]).  That last combination -- what a level-0
// down-sweep would do for every item -- is done here by the lane that uses it,
// into joined (the free level-0 input buffer).  upper == nullptr: a one-level
// scan, whose prefixes are already complete.  Returns the buffer holding the
// prefix and its position *pc.
template <typename Agg, typename R, bool SQRT, typename FB>
PSM_DEV const R* level0_prefix(int64_t c, const R* __restrict__ prefix, const LevelMap& lm,
                               const R* __restrict__ upper, const LevelMap& um,
                               R* __restrict__ joined, const FB& fb, int64_t* pc) {
  using V = typename ViewOf<Agg>::type;
  constexpr int NX = Agg_nx<Agg>::value;
  const int64_t p = lm.rev ? lm.n - 1 - c : c;
  *pc = lm.at(p);
  const int64_t w = p / ((int64_t)lm.K * lm.WG);
  if (upper == nullptr || w == 0) return prefix;
  combine_mem_fb<R, NX, SQRT>(V{const_cast<R*>(upper), um.stride, um.at(w)},
                              V{const_cast<R*>(prefix), lm.stride, *pc}, V{joined, lm.stride, *pc}, fb);
  return joined;
}

template <typename R, int NX, int NY, bool SQRT, typename FB = Fallback>
__global__ __launch_bounds__(kLaneWG) void k_filter_apply(
    int64_t M, int64_t Mp, int chunk, int64_t nchunks, const int64_t* __restrict__ soff, int nseq,
    const R* __restrict__ lin, const R* __restrict__ ys, const R* __restrict__ m0,
    const R* __restrict__ L0, const R* __restrict__ prefix, LevelMap lm,
    const R* __restrict__ upper, LevelMap um, R* __restrict__ joined, R* __restrict__ out_m,
    R* __restrict__ out_L, int out_inter, typename AccOf<R>::type* __restrict__ ell_state, FB fb) {
  if (fb.skip()) return;
  using Acc = typename AccOf<R>::type;
  const StateMap sm{chunk};
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int64_t i0 = c * chunk;
  const int64_t i1 = min(i0 + chunk, M);
  int s = seq_of(soff, nseq, i0);
  R m[NX], L[NX][NX];
  {
    // incoming filtered state: (b, U) of the complete exclusive prefix
    int64_t pc;
    const R* src = level0_prefix<FAgg<R, NX>, R, SQRT>(c, prefix, lm, upper, um, joined, fb, &pc);
    int q = NX * NX;
    ld(src, lm.stride, pc, q, m);
    ld(src, lm.stride, pc, q, L);
  }
  for (int64_t i = i0; i < i1; ++i) {
    while (i >= soff[s + 1]) ++s;
    Acc e = Acc(0.0);
    if (i == soff[s]) {
      ld_aos(m0, s, m);
      ld_aos(L0, s, L);
    } else {
      Step<R, NX, NY> st_;
      load_trans<!SQRT>(lin, Mp, sm(i - 1), st_);
      load_obs<!SQRT>(lin, Mp, sm(i), st_);
      R y[NY];
      ld_aos(ys, i - s - 1, y);
      e = kf_step_fb<R, NX, NY, SQRT>(st_, y, m, L, fb);
    }
    traj_store<!SQRT, R, NX>(out_m, out_L, out_inter, Mp, i, sm(i), m, L);
    if (ell_state != nullptr) ell_state[i] = e;
  }
}

// per-sequence sum of the per-state log-likelihood terms (FP64)
template <int BS, typename Acc>
__global__ __launch_bounds__(BS) void k_ell_reduce(const Acc* __restrict__ ell_state,
                                                    const int64_t* __restrict__ soff,
                                                    Acc* __restrict__ ell) {
  __shared__ Acc red[BS];
  const int s = blockIdx.x;
  const int64_t a = soff[s], b = soff[s + 1];
  Acc acc = Acc(0.0);
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
template <typename R, int NX, int NY, bool SQRT, typename FB = Fallback>
__global__ __launch_bounds__(kLaneWG) void k_smoother_reduce(
    int64_t M, int64_t Mp, int chunk, int64_t nchunks, const int64_t* __restrict__ soff, int nseq,
    const R* __restrict__ lin, const R* __restrict__ fm, const R* __restrict__ fL, int f_inter,
    R* __restrict__ elems, R* __restrict__ agg, LevelMap lm, FB fb) {
  if (fb.skip()) return;
  const StateMap sm{chunk};
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
    traj_load<!SQRT, R, NX>(fm, fL, f_inter, Mp, i, sm(i), m, L);
    SAgg<R, NX> e;
    if (i == soff[s + 1] - 1) {
      copy(m, e.g);
      set_zero(e.E);
      copy(L, e.D);
      e.flag = FLAG_HEAD;
    } else {
      R F[NX][NX], cq[NX][NX], bt[NX];
      int q = 0;
      const int64_t pi = sm(i);
      ld(lin, Mp, pi, q, F);
      ld_tri<!SQRT>(lin, Mp, pi, q, cq);
      ld(lin, Mp, pi, q, bt);
      smoother_element_fb<R, NX, SQRT>(F, cq, bt, m, L, e, fb);
    }
    store_agg(elems, Mp, sm(i), e);
    combine_into_fb<R, NX, SQRT>(acc, e, fb);
    copy_agg(e, acc);
  }
  store_agg(agg, lm.stride, lm.pos(c), acc);
}

template <typename R, int NX, bool SQRT, typename FB = Fallback>
__global__ __launch_bounds__(kLaneWG) void k_smoother_apply(
    int64_t M, int64_t Mp, int chunk, int64_t nchunks, const R* __restrict__ elems,
    const R* __restrict__ prefix, LevelMap lm, const R* __restrict__ upper, LevelMap um,
    R* __restrict__ joined, R* __restrict__ out_m, R* __restrict__ out_L, int out_inter, FB fb) {
  if (fb.skip()) return;
  const StateMap sm{chunk};
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int64_t i0 = c * chunk;
  const int64_t i1 = min(i0 + chunk, M);
  R m[NX], D[NX][NX];
  {
    int64_t pc;
    const R* src = level0_prefix<SAgg<R, NX>, R, SQRT>(c, prefix, lm, upper, um, joined, fb, &pc);
    int q = 0;
    ld(src, lm.stride, pc, q, m);
    q += NX * NX;
    ld(src, lm.stride, pc, q, D);
  }
  for (int64_t i = i1 - 1; i >= i0; --i) {
    SAgg<R, NX> e;
    load_agg(elems, Mp, sm(i), e);
    if (e.flag == FLAG_HEAD) {
      copy(e.g, m);
      copy(e.D, D);
    } else {
      smoother_apply_fb<R, NX, SQRT>(e, m, D, fb);
    }
    traj_store<!SQRT, R, NX>(out_m, out_L, out_inter, Mp, i, sm(i), m, D);
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
  md.row = i;
  R m[NIN], L[NIN][NIN], Fo[NOUT][NIN], Co[NOUT][NOUT], bo[NOUT];
  ld_aos(xm, i, m);
  ld_aos(xL, i, L);
  linearize_point_w<R, NIN, NOUT, SQRT>(md, sp, use_sigma, m, L, Fo, Co, bo, Fallback{0.0f, nullptr});
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
