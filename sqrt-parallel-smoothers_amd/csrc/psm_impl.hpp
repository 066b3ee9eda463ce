// Host-side launchers for one (real type, nx, ny, form) instantiation, and the
// function table the C ABI dispatches through.
#pragma once
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "psm_adjoint.hpp"
#include "psm_kernels.hpp"
#include "psm_sample.hpp"
#include "psm_seq.hpp"

namespace psm {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);   // defined in psm_api.hip

inline int hip_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return PSM_E_HIP;
  }
  return PSM_OK;
}

// ---------------------------------------------------------------- kernel timing
// When enabled (psm_timing_enable), every launch of the whole-pass entry points
// is bracketed by two hipEvents recorded on the launching stream; the report
// aggregates elapsed times per kernel.
int timing_begin(const char* name, hipStream_t st);   // -1 when disabled
void timing_end(int slot, hipStream_t st);

struct TimedScope {
  int slot;
  hipStream_t st;
  TimedScope(const char* name, hipStream_t s) : slot(timing_begin(name, s)), st(s) {}
  ~TimedScope() { timing_end(slot, st); }
};

// ---------------------------------------------------------------- workspace plan
struct ScanLevel {
  int64_t n;      // items at this level
  int K;          // items per lane
  int64_t nwg;    // workgroups
  int64_t stride; // padded item count (component stride of the interleaved layout)
  int wg;         // lanes (or teams) per scan workgroup at this level
  size_t off_in;  // byte offsets in the workspace
  size_t off_out;
};

struct Plan {
  int64_t M = 0;
  int64_t Mp = 0;  // padded state count (component stride of the per-state buffers)
  int chunk = 0;
  int wg = 0;      // scan workgroup size
  bool team = false;  // team-cooperative kernels above level 0
  bool quad = false;  // quad-cooperative level-0 up-sweep
  int64_t nchunks = 0;
  int nlevels = 0;
  ScanLevel lv[16];
  size_t off_lin = 0, off_ell = 0, off_selem = 0, off_guard = 0, off_traj = 0, total = 0;
};

// Per-workgroup flag words of the guarded second launches of the FP32
// square-root kernels (FallbackFlag / Fallback::guard): one per workgroup of
// the largest grid of a pass (the lane-per-state linearization).
inline int64_t guard_words(int64_t Mp) { return (Mp + kLaneWG - 1) / kLaneWG + 1; }

// Integer environment setting within [lo, hi]; unset, empty, malformed or
// out-of-range values give dflt.
inline int env_int(const char* name, int dflt, int lo = INT_MIN, int hi = INT_MAX) {
  const char* v = std::getenv(name);
  if (v == nullptr || *v == '\0') return dflt;
  char* end = nullptr;
  errno = 0;
  const long x = std::strtol(v, &end, 10);
  if (errno != 0 || end == v || *end != '\0' || x < lo || x > hi) return dflt;
  return (int)x;
}

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Default states per chunk: 4, or larger so that the chunk aggregates fit one
// round of 64-lane, fan-in-2 level-0 workgroups on 256 CUs (32768 items; one
// aggregate more would run a nearly empty workgroup in a second round), at
// most 16 (the throughput regime's chunk).
inline int default_chunk(int64_t M) {
  const int64_t c = (M + 32767) / 32768;
  return c < 4 ? 4 : (c > 16 ? 16 : (int)c);
}

// wg: lanes per workgroup of level 0 (one lane per combination); wgu: lanes or
// teams per workgroup of the levels above.
inline Plan make_plan(const psm_problem_t* p, size_t rsize, int nlin, int nqf, int nqs,
                      size_t accsize = sizeof(double), int wg = kScanWG, int wgu = 0,
                      int kdef = 4) {
  if (wgu <= 0) wgu = wg;
  Plan pl;
  pl.M = p->n_states;
  pl.chunk = p->chunk > 0 ? p->chunk : env_int("PSM_CHUNK", default_chunk(pl.M), 1, 4096);
  pl.nchunks = (pl.M + pl.chunk - 1) / pl.chunk;
  pl.Mp = state_stride(pl.M, pl.chunk);
  pl.wg = wg;
  const int nq = nqf > nqs ? nqf : nqs;
  size_t off = 0;
  pl.off_lin = off;
  off = align_up(off + (size_t)nlin * pl.Mp * rsize);
  pl.off_ell = off;
  off = align_up(off + (size_t)pl.M * accsize);
  pl.off_selem = off;
  off = align_up(off + (size_t)nqs * pl.Mp * rsize);
  const int kreq = p->scan_k & PSM_SCAN_K_MASK;
  const int kmax = kreq > 0 ? kreq : env_int("PSM_SCAN_K", kdef, 1, PSM_SCAN_K_MASK);
  const int ktop_req = (p->scan_k & PSM_SCAN_TOP_MASK) >> PSM_SCAN_TOP_SHIFT;
  const int ktop = ktop_req > kmax ? ktop_req : kmax;
  int64_t n = pl.nchunks;
  int l = 0;
  while (true) {
    ScanLevel& s = pl.lv[l];
    s.n = n;
    s.wg = l == 0 ? wg : wgu;
    const int wg_l = s.wg;
    // a level that fits one workgroup with at most ktop items per lane is the
    // top; the others fold kmax items per lane
    if (n <= (int64_t)wg_l * ktop) {
      s.K = (int)((n + wg_l - 1) / wg_l);
      if (s.K < 1) s.K = 1;
      s.nwg = 1;
    } else {
      s.K = kmax;
      s.nwg = (n + (int64_t)wg_l * kmax - 1) / ((int64_t)wg_l * kmax);
    }
    const int64_t blk = (int64_t)wg_l * s.K;
    s.stride = (n + blk - 1) / blk * blk;
    s.off_in = off;
    off = align_up(off + (size_t)nq * s.stride * rsize);
    s.off_out = off;
    off = align_up(off + (size_t)nq * s.stride * rsize);
    ++l;
    if (s.nwg == 1 || l == 16) break;
    n = s.nwg;
  }
  pl.nlevels = l;
  pl.off_guard = off;
  off = align_up(off + (size_t)guard_words(pl.Mp) * sizeof(unsigned));
  pl.total = off;
  return pl;
}

template <typename R>
inline ModelView<R> model_view(const psm_model_t* m) {
  ModelView<R> v;
  v.kind = m->kind;
  v.params = static_cast<const R*>(m->params);
  v.q_mean = static_cast<const R*>(m->q_mean);
  v.q_chol = static_cast<const R*>(m->q_chol);
  return v;
}

template <typename R>
inline SigmaView<R> sigma_view(const psm_sigma_t* s, int* use) {
  SigmaView<R> v{0, nullptr, nullptr, nullptr};
  *use = 0;
  if (s != nullptr && s->npts > 0) {
    v.npts = s->npts;
    v.xi = static_cast<const double*>(s->xi);
    v.wm = static_cast<const double*>(s->wm);
    v.wc = static_cast<const double*>(s->wc);
    *use = 1;
  }
  return v;
}

inline unsigned grid_for(int64_t n, int wg) { return (unsigned)((n + wg - 1) / wg); }

// ---------------------------------------------------------------- the function table
struct OpsTable {
  int (*linearize_trans)(const psm_model_t*, const psm_sigma_t*, int64_t, const void*, const void*,
                         void*, void*, void*, hipStream_t);
  int (*linearize_obs)(const psm_model_t*, const psm_sigma_t*, int64_t, const void*, const void*,
                       void*, void*, void*, hipStream_t);
  int (*filter_elements)(int64_t, const void*, const void*, const void*, const void*, const void*,
                         const void*, const void*, const void*, const void*, void*, void*, void*,
                         void*, void*, hipStream_t);
  int (*filter_combine)(int64_t, const void* const*, const void* const*, void* const*, hipStream_t);
  int (*smoother_elements)(int64_t, const void*, const void*, const void*, const void*, const void*,
                           void*, void*, void*, hipStream_t);
  int (*smoother_combine)(int64_t, const void* const*, const void* const*, void* const*,
                          hipStream_t);
  size_t (*workspace_bytes)(const psm_problem_t*);
  int (*filter)(const psm_problem_t*, const psm_model_t*, const psm_model_t*, const psm_sigma_t*,
                const void*, const void*, const void*, const void*, const void*, void*, void*,
                double*, void*, size_t, hipStream_t, int lin_done);
  int (*smoother)(const psm_problem_t*, const psm_model_t*, const psm_sigma_t*, const void*,
                  const void*, const void*, const void*, void*, void*, void*, size_t, hipStream_t,
                  int lin_done);
  size_t (*sample_workspace_bytes)(const psm_problem_t*, int);
  int (*sample)(const psm_problem_t*, const psm_model_t*, const psm_sigma_t*, const void*,
                const void*, const void*, const void*, const void*, int, void*, void*, size_t,
                hipStream_t);
  int (*seq_filter)(const psm_problem_t*, const psm_model_t*, const psm_model_t*,
                    const psm_sigma_t*, const void*, const void*, const void*, const void*,
                    const void*, void*, void*, double*, hipStream_t);
  int (*seq_smoother)(const psm_problem_t*, const psm_model_t*, const psm_sigma_t*, const void*,
                      const void*, const void*, const void*, void*, void*, hipStream_t);
  // the launch schedule a problem resolves to (environment defaults applied),
  // packed: chunk | level-0 fan-in << 16 | levels << 24 | team << 32 |
  // level-0 workgroup size << 33 | quad level 0 << 49; a captured launch
  // sequence is valid only for the same signature
  uint64_t (*plan_signature)(const psm_problem_t*);
  // reverse-mode adjoint (square-root form, float / double primal)
  size_t (*adjoint_workspace_bytes)(const psm_problem_t*);
  int (*adjoint_prepare)(const psm_problem_t*, const psm_model_t*, const psm_model_t*, const psm_sigma_t*,
                         const void*, const void*, const void*, const void*, const void*, const void*,
                         void*, size_t, hipStream_t);
  int (*adjoint_vjp)(const psm_problem_t*, int, const psm_model_t*, const psm_model_t*, const psm_sigma_t*,
                     const void*, const void*, const void*, const void*, const void*, const void*, const void*,
                     const double*, const double*, const double*, const double*, double*, double*, double*,
                     void*, size_t, hipStream_t);
};

template <typename R, int NX, int NY, bool SQRT>
struct Impl {
  using FA = FAgg<R, NX>;
  using SA = SAgg<R, NX>;
  static constexpr int NLIN = Step<R, NX, NY>::NQ;

  using Acc = typename AccOf<R>::type;
  // Level 0 of a scan (the chunk aggregates: many items) runs one lane per
  // combination with WG lanes per workgroup.  The square-root form in float /
  // double runs the levels above (few items, a latency chain) with the
  // team-cooperative kernels, kTeamVL teams per workgroup, when scan_k has the
  // PSM_SCAN_TEAM bit, or by default (scan_k without a fan-in; the
  // PSM_SCAN_TEAM environment variable set to 0 turns the default off).
  static constexpr bool TEAM = SQRT && (std::is_same<R, float>::value ||
                                        std::is_same<R, double>::value);
  static constexpr int WG = scan_wg<FA, R>() < scan_wg<SA, R>() ? scan_wg<FA, R>() : scan_wg<SA, R>();

  static bool team_requested(const psm_problem_t* p) {
    if (p->scan_k & PSM_SCAN_TEAM) return true;
    if (p->scan_k & PSM_SCAN_K_MASK) return false;  // explicit fan-in without the team bit
    return env_int("PSM_SCAN_TEAM", 1) != 0;
  }

  static Plan plan(const psm_problem_t* p) {
    const bool team = TEAM && team_requested(p);
    const bool half = TEAM && (p->scan_k & PSM_SCAN_HALF) != 0;
    Plan pl = make_plan(p, sizeof(R), NLIN, FA::NQ, SA::NQ, sizeof(Acc), half ? WG / 2 : WG,
                        team ? kTeamVL : WG, team ? 2 : 4);
    pl.team = team;
    pl.quad = TEAM && (p->scan_k & PSM_SCAN_QUAD) != 0;
    // one state-interleaved trajectory (mean and factor per state) for launch
    // sequences that pass trajectories between their own passes
    pl.off_traj = pl.total;
    pl.total = align_up(pl.total + (size_t)(NX + NX * NX) * pl.Mp * sizeof(R));
    return pl;
  }

  static size_t workspace_bytes(const psm_problem_t* p) { return plan(p).total; }

  static uint64_t plan_signature(const psm_problem_t* p) {
    const Plan pl = plan(p);
    uint64_t sig = (uint64_t)(uint32_t)pl.chunk & 0xffffu;
    sig |= (uint64_t)(pl.lv[0].K & 0xff) << 16;
    sig |= (uint64_t)(pl.nlevels & 0xff) << 24;
    sig |= (uint64_t)(pl.team ? 1 : 0) << 32;
    sig |= (uint64_t)(pl.lv[0].wg & 0xffff) << 33;
    sig |= (uint64_t)(pl.quad ? 1 : 0) << 49;
    for (int l = 1; l < pl.nlevels; ++l) sig = sig * 1099511628211ull + (uint64_t)pl.lv[l].K;
    return sig;
  }

  template <int NIN, int NOUT>
  static int linearize_points(const psm_model_t* md, const psm_sigma_t* sg, int64_t n,
                              const void* xm, const void* xL, void* F, void* C, void* b,
                              hipStream_t st) {
    if (n <= 0) return PSM_OK;
    int use = 0;
    SigmaView<R> sv = sigma_view<R>(sg, &use);
    ModelView<R> mv = model_view<R>(md);
    mv.npts = use ? sv.npts : 0;
    hipLaunchKernelGGL((k_linearize_points<R, NIN, NOUT, SQRT>), dim3(grid_for(n, kLaneWG)),
                       dim3(kLaneWG), 0, st, n, mv, sv, use,
                       static_cast<const R*>(xm), static_cast<const R*>(xL), static_cast<R*>(F),
                       static_cast<R*>(C), static_cast<R*>(b));
    return hip_check("k_linearize_points");
  }
  static int linearize_trans(const psm_model_t* md, const psm_sigma_t* sg, int64_t n,
                             const void* xm, const void* xL, void* F, void* C, void* b,
                             hipStream_t st) {
    return linearize_points<NX, NX>(md, sg, n, xm, xL, F, C, b, st);
  }
  static int linearize_obs(const psm_model_t* md, const psm_sigma_t* sg, int64_t n, const void* xm,
                           const void* xL, void* F, void* C, void* b, hipStream_t st) {
    return linearize_points<NX, NY>(md, sg, n, xm, xL, F, C, b, st);
  }

  static int filter_elements(int64_t n, const void* F, const void* cq, const void* bt,
                             const void* H, const void* cr, const void* c, const void* m,
                             const void* L, const void* y, void* A, void* b, void* U, void* eta,
                             void* Z, hipStream_t st) {
    if (n <= 0) return PSM_OK;
    hipLaunchKernelGGL((k_filter_elements<R, NX, NY, SQRT>), dim3(grid_for(n, kLaneWG)),
                       dim3(kLaneWG), 0, st, n, (const R*)F, (const R*)cq, (const R*)bt,
                       (const R*)H, (const R*)cr, (const R*)c, (const R*)m, (const R*)L,
                       (const R*)y, (R*)A, (R*)b, (R*)U, (R*)eta, (R*)Z);
    return hip_check("k_filter_elements");
  }

  static int filter_combine(int64_t n, const void* const* e1, const void* const* e2,
                            void* const* out, hipStream_t st) {
    if (n <= 0) return PSM_OK;
    FPtrs<R, NX> a, b;
    FOut<R, NX> o;
    for (int k = 0; k < 5; ++k) {
      a.p[k] = (const R*)e1[k];
      b.p[k] = (const R*)e2[k];
      o.p[k] = (R*)out[k];
    }
    hipLaunchKernelGGL((k_filter_combine<R, NX, SQRT>), dim3(grid_for(n, kLaneWG)), dim3(kLaneWG),
                       0, st, n, a, b, o);
    return hip_check("k_filter_combine");
  }

  static int smoother_elements(int64_t n, const void* F, const void* cq, const void* bt,
                               const void* m, const void* L, void* g, void* E, void* D,
                               hipStream_t st) {
    if (n <= 0) return PSM_OK;
    hipLaunchKernelGGL((k_smoother_elements<R, NX, SQRT>), dim3(grid_for(n, kLaneWG)),
                       dim3(kLaneWG), 0, st, n, (const R*)F, (const R*)cq, (const R*)bt,
                       (const R*)m, (const R*)L, (R*)g, (R*)E, (R*)D);
    return hip_check("k_smoother_elements");
  }

  static int smoother_combine(int64_t n, const void* const* e1, const void* const* e2,
                              void* const* out, hipStream_t st) {
    if (n <= 0) return PSM_OK;
    SPtrs<R, NX> a, b;
    SOut<R, NX> o;
    for (int k = 0; k < 3; ++k) {
      a.p[k] = (const R*)e1[k];
      b.p[k] = (const R*)e2[k];
      o.p[k] = (R*)out[k];
    }
    hipLaunchKernelGGL((k_smoother_combine<R, NX, SQRT>), dim3(grid_for(n, kLaneWG)),
                       dim3(kLaneWG), 0, st, n, a, b, o);
    return hip_check("k_smoother_combine");
  }

  // Exclusive scan of the level-0 aggregates; *final_off receives the workspace
  // offset of the complete exclusive prefixes.
  static LevelMap level_map(const Plan& pl, int l, int rev) {
    const ScanLevel& s = pl.lv[l];
    return LevelMap{s.n, s.stride, s.K, s.wg, rev};
  }

  // FP32 square-root passes launch each kernel twice: without FP64 code and
  // with a FallbackFlag, then with a Fallback guarded by that flag, which
  // redoes the kernel's work only if the first launch met an ill-conditioned
  // result (see Fallback).  Other forms launch once with the Fallback as given.
  // Every guarded pair leaves the flags it used down again (the second launch
  // clears the flags it finds up), so one array serves all pairs of a pass;
  // it is cleared once at the start of the pass.
  struct Guards {
    unsigned* base;
    int64_t words;
  };
  static Guards guards(const Plan& pl, char* ws, hipStream_t st) {
    Guards g{(unsigned*)(ws + pl.off_guard), guard_words(pl.Mp)};
    if constexpr (kFallback<R, SQRT>)
      hipLaunchKernelGGL((k_fill<unsigned>), dim3(grid_for(g.words, 256)), dim3(256), 0, st, g.base,
                         (int)g.words, 0u);
    return g;
  }
  // waves: wavefronts of the launch.  A launch of at most PSM_FB_SINGLE_WAVES
  // wavefronts (default 1024, one per SIMD of an MI355X) runs once with the
  // FP64 code in place: its registers do not limit its occupancy, and the
  // guarded second launch would only add a kernel node to a latency-bound
  // pass.
  static int64_t single_launch_waves() {
    static const int64_t w = env_int("PSM_FB_SINGLE_WAVES", 1024, 0, 1 << 30);
    return w;
  }
  template <typename L>
  static void launch_fb(const Fallback& fb, Guards& g, unsigned grid, int64_t waves, L&& launch) {
    if constexpr (kFallback<R, SQRT>) {
      // no flag word per workgroup, or a small launch with the check on: one
      // launch with the FP64 code
      if ((int64_t)grid > g.words || (fb.thr > 0.f && waves <= single_launch_waves())) {
        launch(fb);
        return;
      }
      launch(FallbackFlag{fb.thr, g.base});
      if (fb.thr > 0.f) {
        Fallback redo = fb;
        redo.guard = g.base;
        launch(redo);
      }
    } else {
      launch(fb);
    }
  }

  // Exclusive scan of the level-0 items (chunk aggregates, already in scan
  // order in the interleaved layout).  The level-0 down-sweep is left to the
  // apply kernel (level0_prefix): *final_off receives the level-0 prefixes
  // within each workgroup and *upper_off the complete level-1 prefixes
  // (SIZE_MAX with a single level, whose prefixes are already complete).
  template <typename Agg>
  static int run_scan(const Plan& pl, char* ws, const Fallback& fb, Guards& g, hipStream_t st,
                      size_t* final_off, size_t* upper_off) {
    constexpr bool filt = std::is_same<Agg, FA>::value;
    for (int l = 0; l < pl.nlevels; ++l) {
      const ScanLevel& s = pl.lv[l];
      R* in = (R*)(ws + s.off_in);
      R* out = (R*)(ws + s.off_out);
      const bool top = l + 1 == pl.nlevels;
      R* tot = top ? nullptr : (R*)(ws + pl.lv[l + 1].off_in);
      const LevelMap tm = top ? LevelMap{0, 0, 1, WG, 0} : level_map(pl, l + 1, 0);
      TimedScope ts_(l == 0 ? (filt ? "k_fscan_up_l0" : "k_sscan_up_l0")
                            : (filt ? "k_fscan_up" : "k_sscan_up"), st);
      // wavefronts per workgroup of the kernel this level runs
      const int64_t wpw = (TEAM && l > 0 && pl.team) ? TeamCfg<Agg, R>::WGT / 64
                          : (TEAM && l == 0 && pl.quad) ? (s.wg * kQuadLanes + 63) / 64
                                                        : (s.wg + 63) / 64;
      launch_fb(fb, g, (unsigned)s.nwg, s.nwg * wpw, [&](auto f) {
        using F = decltype(f);
        bool launched = false;
        if constexpr (TEAM) {
          if (l > 0 && pl.team) {
            hipLaunchKernelGGL((k_scan_up_team<Agg, R, F>), dim3((unsigned)s.nwg),
                               dim3(TeamCfg<Agg, R>::WGT), 0, st, (const R*)in, level_map(pl, l, 0),
                               out, tot, tm, f);
            launched = true;
          }
        }
        if constexpr (TEAM) {
          if (l == 0 && pl.quad) {
            if (s.wg == WG / 2)
              hipLaunchKernelGGL((k_scan_up_quad<Agg, R, WG / 2, F>), dim3((unsigned)s.nwg),
                                 dim3(WG / 2 * kQuadLanes), 0, st, (const R*)in,
                                 level_map(pl, l, 0), out, tot, tm, f);
            else
              hipLaunchKernelGGL((k_scan_up_quad<Agg, R, WG, F>), dim3((unsigned)s.nwg),
                                 dim3(WG * kQuadLanes), 0, st, (const R*)in, level_map(pl, l, 0),
                                 out, tot, tm, f);
            launched = true;
          }
        }
        if constexpr (TEAM) {
          if (!launched && s.wg == WG / 2) {  // half-wave level 0
            hipLaunchKernelGGL((k_scan_up<Agg, R, SQRT, WG / 2, F>), dim3((unsigned)s.nwg), dim3(WG / 2),
                               0, st, (const R*)in, level_map(pl, l, 0), out, tot, tm, f);
            launched = true;
          }
        }
        if (!launched)
          hipLaunchKernelGGL((k_scan_up<Agg, R, SQRT, WG, F>), dim3((unsigned)s.nwg), dim3(WG), 0,
                             st, (const R*)in, level_map(pl, l, 0), out, tot, tm, f);
      });
      int rc = hip_check("k_scan_up");
      if (rc) return rc;
    }
    // complete prefixes of level l >= 1 land in lv[l].in (free once the up-sweep is done)
    for (int l = pl.nlevels - 2; l >= 1; --l) {
      const ScanLevel& s = pl.lv[l];
      const size_t upper_off =
          (l + 1 == pl.nlevels - 1) ? pl.lv[l + 1].off_out : pl.lv[l + 1].off_in;
      TimedScope ts_(filt ? "k_fscan_down" : "k_sscan_down", st);
      const int64_t dwaves = (TEAM && l > 0 && pl.team)
                                 ? s.nwg * s.K * (TeamCfg<Agg, R>::WGT / 64)
                                 : (int64_t)grid_for(s.nwg * s.K * s.wg, WG) * ((WG + 63) / 64);
      launch_fb(fb, g, std::max((unsigned)(s.nwg * s.K), grid_for(s.nwg * s.K * s.wg, WG)), dwaves,
                [&](auto f) {
        using F = decltype(f);
        bool launched = false;
        if constexpr (TEAM) {
          if (l > 0 && pl.team) {
            hipLaunchKernelGGL((k_scan_down_team<Agg, R, F>), dim3((unsigned)(s.nwg * s.K)),
                               dim3(TeamCfg<Agg, R>::WGT), 0, st, (const R*)(ws + s.off_out),
                               level_map(pl, l, 0), (const R*)(ws + upper_off),
                               level_map(pl, l + 1, 0), (R*)(ws + s.off_in), f);
            launched = true;
          }
        }
        if (!launched)
          hipLaunchKernelGGL((k_scan_down<Agg, R, SQRT, WG, F>),
                             dim3((unsigned)grid_for(s.nwg * s.K * s.wg, WG)), dim3(WG), 0,
                             st, (const R*)(ws + s.off_out), level_map(pl, l, 0),
                             (const R*)(ws + upper_off), level_map(pl, l + 1, 0),
                             (R*)(ws + s.off_in), f);
      });
      int rc = hip_check("k_scan_down");
      if (rc) return rc;
    }
    *final_off = pl.lv[0].off_out;
    *upper_off = pl.nlevels == 1 ? SIZE_MAX
                                 : (pl.nlevels == 2 ? pl.lv[1].off_out : pl.lv[1].off_in);
    return PSM_OK;
  }

  static int linearize_states(const Plan& pl, const psm_model_t* tr, const psm_model_t* ob,
                              const psm_sigma_t* sg, const void* nm, const void* nL, char* ws,
                              hipStream_t st, const Fallback& fb, Guards& g) {
    int use = 0;
    SigmaView<R> sv = sigma_view<R>(sg, &use);
    ModelView<R> trv = model_view<R>(tr);
    ModelView<R> obv = ob ? model_view<R>(ob) : model_view<R>(tr);
    trv.npts = use ? sv.npts : 0;
    obv.npts = use ? sv.npts : 0;
    // nm == nullptr: the nominal is the workspace trajectory (state-interleaved)
    const int inter = nm == nullptr ? 1 : 0;
    if (inter) nm = ws + pl.off_traj;
    TimedScope ts_("k_linearize_states", st);
    launch_fb(fb, g, grid_for(pl.Mp, kLaneWG), grid_for(pl.Mp, kLaneWG), [&](auto f) {
      hipLaunchKernelGGL((k_linearize_states<R, NX, NY, SQRT, decltype(f)>), dim3(grid_for(pl.Mp, kLaneWG)),
                         dim3(kLaneWG), 0, st, pl.M, pl.Mp, pl.chunk, trv, obv, ob ? 1 : 0, sv, use,
                         (const R*)nm, (const R*)nL, inter, (R*)(ws + pl.off_lin), f);
    });
    return hip_check("k_linearize_states");
  }

  static int filter(const psm_problem_t* p, const psm_model_t* tr, const psm_model_t* ob,
                    const psm_sigma_t* sg, const void* ys, const void* m0, const void* L0,
                    const void* nm, const void* nL, void* om, void* oL, double* ell, void* wsv,
                    size_t wsb, hipStream_t st, int lin_done) {
    const Plan pl = plan(p);
    if (wsb < pl.total) {
      set_error("workspace too small");
      return PSM_E_WORKSPACE;
    }
    if (pl.M <= 0) return PSM_OK;
    char* ws = (char*)wsv;
    int rc;
    const Fallback fb{p->fallback_ratio, p->fallback_count};
    Guards g = guards(pl, ws, st);
    if (!lin_done) {
      rc = linearize_states(pl, tr, ob, sg, nm, nL, ws, st, fb, g);
      if (rc) return rc;
    }
    const R* lin = (const R*)(ws + pl.off_lin);
    // om == nullptr: the filtered trajectory goes to the workspace trajectory
    // (state-interleaved), where a following smoother pass of the same launch
    // sequence reads it
    const int om_inter = om == nullptr ? 1 : 0;
    if (om_inter) om = ws + pl.off_traj;
    {
    TimedScope ts_("k_filter_reduce", st);
    launch_fb(fb, g, grid_for(pl.nchunks, kLaneWG), grid_for(pl.nchunks, kLaneWG), [&](auto f) {
      hipLaunchKernelGGL((k_filter_reduce<R, NX, NY, SQRT, decltype(f)>), dim3(grid_for(pl.nchunks, kLaneWG)),
                         dim3(kLaneWG), 0, st, pl.M, pl.Mp, pl.chunk, pl.nchunks, p->state_offsets,
                         p->n_seq, lin, (const R*)ys, (const R*)m0, (const R*)L0,
                         (R*)(ws + pl.lv[0].off_in), level_map(pl, 0, 0), f);
    });
    }
    if ((rc = hip_check("k_filter_reduce"))) return rc;
    size_t pre_off = 0, up_off = 0;
    if ((rc = run_scan<FA>(pl, ws, fb, g, st, &pre_off, &up_off))) return rc;
    Acc* ell_state = ell ? (Acc*)(ws + pl.off_ell) : nullptr;
    {
    TimedScope ts_("k_filter_apply", st);
    launch_fb(fb, g, grid_for(pl.nchunks, kLaneWG), grid_for(pl.nchunks, kLaneWG), [&](auto f) {
      hipLaunchKernelGGL((k_filter_apply<R, NX, NY, SQRT, decltype(f)>), dim3(grid_for(pl.nchunks, kLaneWG)),
                         dim3(kLaneWG), 0, st, pl.M, pl.Mp, pl.chunk, pl.nchunks, p->state_offsets,
                         p->n_seq, lin, (const R*)ys, (const R*)m0, (const R*)L0,
                         (const R*)(ws + pre_off), level_map(pl, 0, 0),
                         up_off == SIZE_MAX ? nullptr : (const R*)(ws + up_off),
                         pl.nlevels > 1 ? level_map(pl, 1, 0) : level_map(pl, 0, 0),
                         (R*)(ws + pl.lv[0].off_in), (R*)om, (R*)oL, om_inter, ell_state, f);
    });
    }
    if ((rc = hip_check("k_filter_apply"))) return rc;
    if (ell) {
      hipLaunchKernelGGL((k_ell_reduce<256, Acc>), dim3((unsigned)p->n_seq), dim3(256), 0, st,
                         (const Acc*)ell_state, p->state_offsets, (Acc*)ell);
      if ((rc = hip_check("k_ell_reduce"))) return rc;
    }
    return PSM_OK;
  }

  static int smoother(const psm_problem_t* p, const psm_model_t* tr, const psm_sigma_t* sg,
                      const void* fm, const void* fL, const void* nm, const void* nL, void* om,
                      void* oL, void* wsv, size_t wsb, hipStream_t st, int lin_done) {
    const Plan pl = plan(p);
    if (wsb < pl.total) {
      set_error("workspace too small");
      return PSM_E_WORKSPACE;
    }
    if (pl.M <= 0) return PSM_OK;
    char* ws = (char*)wsv;
    int rc;
    const Fallback fb{p->fallback_ratio, p->fallback_count};
    Guards g = guards(pl, ws, st);
    if (!lin_done) {
      rc = linearize_states(pl, tr, nullptr, sg, nm, nL, ws, st, fb, g);
      if (rc) return rc;
    }
    // fm == nullptr: the filtered trajectory is the workspace trajectory; om ==
    // nullptr: the smoothed one goes there (both state-interleaved)
    const int fm_inter = fm == nullptr ? 1 : 0, om_inter = om == nullptr ? 1 : 0;
    if (fm_inter) fm = ws + pl.off_traj;
    if (om_inter) om = ws + pl.off_traj;
    {
    TimedScope ts_("k_smoother_reduce", st);
    launch_fb(fb, g, grid_for(pl.nchunks, kLaneWG), grid_for(pl.nchunks, kLaneWG), [&](auto f) {
      hipLaunchKernelGGL((k_smoother_reduce<R, NX, NY, SQRT, decltype(f)>),
                         dim3(grid_for(pl.nchunks, kLaneWG)), dim3(kLaneWG), 0, st, pl.M, pl.Mp, pl.chunk,
                         pl.nchunks, p->state_offsets, p->n_seq, (const R*)(ws + pl.off_lin), (const R*)fm,
                         (const R*)fL, fm_inter, (R*)(ws + pl.off_selem), (R*)(ws + pl.lv[0].off_in),
                         level_map(pl, 0, 1), f);
    });
    }
    if ((rc = hip_check("k_smoother_reduce"))) return rc;
    size_t pre_off = 0, up_off = 0;
    if ((rc = run_scan<SA>(pl, ws, fb, g, st, &pre_off, &up_off))) return rc;
    {
      TimedScope ts_("k_smoother_apply", st);
      launch_fb(fb, g, grid_for(pl.nchunks, kLaneWG), grid_for(pl.nchunks, kLaneWG), [&](auto f) {
        hipLaunchKernelGGL((k_smoother_apply<R, NX, SQRT, decltype(f)>), dim3(grid_for(pl.nchunks, kLaneWG)),
                           dim3(kLaneWG), 0, st, pl.M, pl.Mp, pl.chunk, pl.nchunks,
                           (const R*)(ws + pl.off_selem), (const R*)(ws + pre_off),
                           level_map(pl, 0, 1), up_off == SIZE_MAX ? nullptr : (const R*)(ws + up_off),
                           pl.nlevels > 1 ? level_map(pl, 1, 0) : level_map(pl, 0, 0),
                           (R*)(ws + pl.lv[0].off_in), (R*)om, (R*)oL, om_inter, f);
      });
    }
    return hip_check("k_smoother_apply");
  }

  // ---------------------------------------------------------------- sampling
  struct SamplePlan {
    int64_t Cs, nch;
    size_t off_P, off_o, off_xb, total;
  };
  static SamplePlan sample_plan(const Plan& pl, int S) {
    SamplePlan sp;
    sp.Cs = sample_chunk(pl.M);
    sp.nch = (pl.M + sp.Cs - 1) / sp.Cs;
    size_t off = pl.total;
    sp.off_P = off;
    off = align_up(off + (size_t)sp.nch * NX * NX * sizeof(R));
    sp.off_o = off;
    off = align_up(off + (size_t)sp.nch * S * NX * sizeof(R));
    sp.off_xb = off;
    off = align_up(off + (size_t)sp.nch * S * NX * sizeof(R));
    sp.total = off;
    return sp;
  }
  static size_t sample_workspace_bytes(const psm_problem_t* p, int S) {
    return sample_plan(plan(p), S).total;
  }
  static int sample(const psm_problem_t* p, const psm_model_t* tr, const psm_sigma_t* sg,
                    const void* fm, const void* fL, const void* nm, const void* nL, const void* eps,
                    int S, void* out, void* wsv, size_t wsb, hipStream_t st) {
    const Plan pl = plan(p);
    const SamplePlan sp = sample_plan(pl, S);
    if (wsb < sp.total) {
      set_error("workspace too small");
      return PSM_E_WORKSPACE;
    }
    if (pl.M <= 0 || S <= 0) return PSM_OK;
    char* ws = (char*)wsv;
    const Fallback fb{p->fallback_ratio, p->fallback_count};
    Guards g = guards(pl, ws, st);
    int rc = linearize_states(pl, tr, nullptr, sg, nm, nL, ws, st, fb, g);
    if (rc) return rc;
    // smoother elements at every state (the chunk aggregates it also forms are not used)
    launch_fb(fb, g, grid_for(pl.nchunks, kLaneWG), grid_for(pl.nchunks, kLaneWG), [&](auto f) {
      hipLaunchKernelGGL((k_smoother_reduce<R, NX, NY, SQRT, decltype(f)>), dim3(grid_for(pl.nchunks, kLaneWG)),
                         dim3(kLaneWG), 0, st, pl.M, pl.Mp, pl.chunk, pl.nchunks, p->state_offsets,
                         p->n_seq, (const R*)(ws + pl.off_lin), (const R*)fm, (const R*)fL, 0,
                         (R*)(ws + pl.off_selem), (R*)(ws + pl.lv[0].off_in), level_map(pl, 0, 1), f);
    });
    if ((rc = hip_check("k_smoother_reduce"))) return rc;
    const int64_t lanes = sp.nch * S;
    hipLaunchKernelGGL((k_sample_reduce<R, NX, SQRT>), dim3(grid_for(lanes, kLaneWG)), dim3(kLaneWG),
                       0, st, pl.M, pl.Mp, pl.chunk, sp.Cs, sp.nch, S, (const R*)(ws + pl.off_selem),
                       (const R*)eps, (R*)(ws + sp.off_P), (R*)(ws + sp.off_o));
    if ((rc = hip_check("k_sample_reduce"))) return rc;
    hipLaunchKernelGGL((k_sample_bounds<R, NX>), dim3(grid_for(S, kLaneWG)), dim3(kLaneWG), 0, st,
                       sp.nch, S, (const R*)(ws + sp.off_P), (const R*)(ws + sp.off_o),
                       (R*)(ws + sp.off_xb));
    if ((rc = hip_check("k_sample_bounds"))) return rc;
    hipLaunchKernelGGL((k_sample_apply<R, NX, SQRT>), dim3(grid_for(lanes, kLaneWG)), dim3(kLaneWG),
                       0, st, pl.M, pl.Mp, pl.chunk, sp.Cs, sp.nch, S, (const R*)(ws + pl.off_selem),
                       (const R*)eps, (const R*)(ws + sp.off_xb), (R*)out);
    return hip_check("k_sample_apply");
  }

  // Sequential passes (one lane per sequence); real dtypes only.
  static constexpr bool kReal = std::is_same<R, float>::value || std::is_same<R, double>::value;
  static int seq_filter(const psm_problem_t* p, const psm_model_t* tr, const psm_model_t* ob,
                        const psm_sigma_t* sg, const void* ys, const void* m0, const void* L0,
                        const void* nm, const void* nL, void* om, void* oL, double* ell,
                        hipStream_t st) {
    if constexpr (!kReal) {
      set_error("sequential passes support PSM_F32 and PSM_F64 only");
      return PSM_E_UNSUPPORTED;
    } else {
      if (p->n_seq <= 0) return PSM_OK;
      int use = 0;
      SigmaView<R> sv = sigma_view<R>(sg, &use);
      ModelView<R> trv = model_view<R>(tr), obv = model_view<R>(ob);
      trv.npts = use ? sv.npts : 0;
      obv.npts = use ? sv.npts : 0;
      const Fallback fb{p->fallback_ratio, p->fallback_count};
      TimedScope ts_("k_seq_filter", st);
      hipLaunchKernelGGL((k_seq_filter<R, NX, NY, SQRT>), dim3(grid_for(p->n_seq, kLaneWG)),
                         dim3(kLaneWG), 0, st, p->n_seq, p->state_offsets, trv, obv, sv, use,
                         (const R*)nm, (const R*)nL, (const R*)ys, (const R*)m0, (const R*)L0,
                         (R*)om, (R*)oL, ell, fb);
      return hip_check("k_seq_filter");
    }
  }
  static int seq_smoother(const psm_problem_t* p, const psm_model_t* tr, const psm_sigma_t* sg,
                          const void* fm, const void* fL, const void* nm, const void* nL, void* om,
                          void* oL, hipStream_t st) {
    if constexpr (!kReal) {
      set_error("sequential passes support PSM_F32 and PSM_F64 only");
      return PSM_E_UNSUPPORTED;
    } else {
      if (p->n_seq <= 0) return PSM_OK;
      int use = 0;
      SigmaView<R> sv = sigma_view<R>(sg, &use);
      ModelView<R> trv = model_view<R>(tr);
      trv.npts = use ? sv.npts : 0;
      const Fallback fb{p->fallback_ratio, p->fallback_count};
      TimedScope ts_("k_seq_smoother", st);
      hipLaunchKernelGGL((k_seq_smoother<R, NX, SQRT>), dim3(grid_for(p->n_seq, kLaneWG)),
                         dim3(kLaneWG), 0, st, p->n_seq, p->state_offsets, trv, sv, use,
                         (const R*)nm, (const R*)nL, (const R*)fm, (const R*)fL, (R*)om, (R*)oL,
                         fb);
      return hip_check("k_seq_smoother");
    }
  }

  // ---------------------------------------------------------------- adjoint
  // FP64 primal only: an FP32 objective's adjoint runs on an FP64 problem that
  // holds its FP32-rounded arrays (an FP32 checkpoint would carry FP32 rounding
  // into every step's residual), so no FP32 slice carries adjoint kernels.
  static constexpr bool kAdj = SQRT && std::is_same<R, double>::value;
  struct AdjPlan {
    int64_t M = 0, Mp = 0;
    size_t off_ck = 0, off_pa = 0, off_lc = 0, off_ln = 0, off_th = 0;
    int nlv = 0;               // aggregate levels above the states
    int64_t lvn[16] = {}, lvst[16] = {};
    size_t off_lv[16] = {};
    size_t total = 0;
  };
  static AdjPlan adj_plan(const psm_problem_t* p) {
    AdjPlan a;
    constexpr int C = adj::kChunk;
    a.M = p->n_states;
    a.Mp = state_stride(a.M, C);
    size_t off = 0;
    const size_t d = sizeof(double);
    a.off_ck = off;
    off = align_up(off + (size_t)adj::Ckp<NX>::N * a.Mp * d);
    a.off_pa = off;
    off = align_up(off + (size_t)adj::Pas<NX>::N * a.Mp * d);
    a.off_lc = off;
    off = align_up(off + (size_t)adj::Lco<NX, NY>::N * a.Mp * d);
    a.off_ln = off;   // the linearized model at the fixed point, same layout
    off = align_up(off + (size_t)adj::Lco<NX, NY>::N * a.Mp * d);
    a.off_th = off;
    off = align_up(off + (size_t)adj::Th<NX, NY>::N * a.Mp * d);
    // aggregate levels: level 1 holds ceil(M / C) items, ... up to one chunk
    int64_t n = a.M;
    while (n > C && a.nlv < 15) {
      n = (n + C - 1) / C;
      ++a.nlv;
      a.lvn[a.nlv] = n;
      a.lvst[a.nlv] = state_stride(n, C);
      a.off_lv[a.nlv] = off;
      off = align_up(off + (size_t)adj::Sc<NX, true>::NQ * a.lvst[a.nlv] * d);
    }
    a.lvn[0] = a.M;
    a.lvst[0] = a.Mp;
    a.total = off;
    return a;
  }
  static size_t adjoint_workspace_bytes(const psm_problem_t* p) {
    if constexpr (!kAdj) return 0;
    else return adj_plan(p).total;
  }
  static ModelView<double, R> model_view_d(const psm_model_t* m) {
    return ModelView<double, R>{m->kind, static_cast<const R*>(m->params), static_cast<const R*>(m->q_mean),
                                static_cast<const R*>(m->q_chol)};
  }
  static int adj_check(const psm_problem_t* p, const psm_model_t* tr, const psm_model_t* ob, size_t awsb) {
    if constexpr (!kAdj) {
      set_error("the adjoint supports the square-root form in PSM_F64 only (run an FP32 objective's "
                "adjoint on an FP64 problem holding its FP32-rounded arrays)");
      return PSM_E_UNSUPPORTED;
    } else {
      for (const psm_model_t* m : {tr, ob}) {
        if (m->kind == PSM_MODEL_TABULATED || m->kind == PSM_MODEL_CONDITIONAL) {
          set_error("the adjoint needs models the kernels evaluate (not tabulated or conditional-moments models)");
          return PSM_E_UNSUPPORTED;
        }
      }
      if (awsb < adj_plan(p).total) {
        set_error("adjoint workspace too small");
        return PSM_E_WORKSPACE;
      }
      return PSM_OK;
    }
  }

  // Checkpoint of the pass at nominal (nm, nL) whose filtered (fm, fL) and
  // smoothed (sm, sL) trajectories the caller computed.
  static int adjoint_prepare(const psm_problem_t* p, const psm_model_t* tr, const psm_model_t* ob,
                             const psm_sigma_t* sg, const void* nm, const void* nL, const void* fm,
                             const void* fL, const void* sm, const void* sL, void* awsv, size_t awsb,
                             hipStream_t st) {
    int rc = adj_check(p, tr, ob, awsb);
    if (rc) return rc;
    if constexpr (kAdj) {
      (void)sm;
      const AdjPlan a = adj_plan(p);
      if (a.M <= 0) return PSM_OK;
      int use = 0;
      SigmaView<double> sv = sigma_view<double>(sg, &use);
      ModelView<double, R> trv = model_view_d(tr), obv = model_view_d(ob);
      char* aws = (char*)awsv;
      TimedScope ts_("k_adj_prep", st);
      hipLaunchKernelGGL((adj::k_adj_prep<R, NX, NY>), dim3(grid_for(a.Mp, 64)), dim3(64), 0, st, a.M, a.Mp,
                         p->state_offsets, p->n_seq, trv, obv, sv, use, (const R*)nm, (const R*)nL,
                         (const R*)fm, (const R*)fL, (const R*)sL, (double*)(aws + a.off_ck),
                         (double*)(aws + a.off_ln));
      return hip_check("k_adj_prep");
    }
    return PSM_OK;
  }

  // chunked scan of one recursion (levels above the states as planned)
  template <bool CONG, int SRC>
  static int adj_scan(const AdjPlan& a, char* aws, hipStream_t st) {
    using namespace adj;
    const double* ck = (const double*)(aws + a.off_ck);
    double* pa = (double*)(aws + a.off_pa);
    auto lvbuf = [&](int k) { return (double*)(aws + a.off_lv[k]); };
    const ScanIO<NX, CONG, SRC> io0{ck, pa, nullptr, a.Mp, 0, a.M};
    constexpr int C = kChunk;
    // up-sweep: aggregates of level k into level k + 1
    for (int k = 0; k < a.nlv; ++k) {
      const unsigned g = grid_for((a.lvn[k] + C - 1) / C, 64);
      if (k == 0) {
        hipLaunchKernelGGL((k_ascan_reduce<NX, CONG, SRC>), dim3(g), dim3(64), 0, st, a.lvn[0], io0,
                           lvbuf(1), a.lvst[1]);
      } else {
        const ScanIO<NX, CONG, SRC_LEVEL> io{ck, pa, lvbuf(k), a.Mp, a.lvst[k], a.M};
        hipLaunchKernelGGL((k_ascan_reduce<NX, CONG, SRC_LEVEL>), dim3(g), dim3(64), 0, st, a.lvn[k], io,
                           lvbuf(k + 1), a.lvst[k + 1]);
      }
      int rc = hip_check("k_ascan_reduce");
      if (rc) return rc;
    }
    // down-sweep: the top level from zero, then every level from the one above
    for (int k = a.nlv; k >= 0; --k) {
      const unsigned g = grid_for((a.lvn[k] + C - 1) / C, 64);
      const double* up = k == a.nlv ? nullptr : lvbuf(k + 1);
      const int64_t ust = k == a.nlv ? 0 : a.lvst[k + 1];
      if (k == 0) {
        hipLaunchKernelGGL((k_ascan_apply<NX, CONG, SRC>), dim3(g), dim3(64), 0, st, a.lvn[0], io0, up, ust);
      } else {
        const ScanIO<NX, CONG, SRC_LEVEL> io{ck, pa, lvbuf(k), a.Mp, a.lvst[k], a.M};
        hipLaunchKernelGGL((k_ascan_apply<NX, CONG, SRC_LEVEL>), dim3(g), dim3(64), 0, st, a.lvn[k], io, up,
                           ust);
      }
      int rc = hip_check("k_ascan_apply");
      if (rc) return rc;
    }
    return PSM_OK;
  }

  // mode 0: cotangents (cot_m, cot_L) of the smoothed trajectory -> nominal
  // cotangents (out_m, out_L) [+ (add_m, add_L)] and parameter cotangents theta
  // (n_seq x NTH); mode 1: the same for the pass's log-likelihood (cot_* unused).
  static int adjoint_vjp(const psm_problem_t* p, int mode, const psm_model_t* tr, const psm_model_t* ob,
                         const psm_sigma_t* sg, const void* ys, const void* m0, const void* L0, const void* nm,
                         const void* nL, const void* sm, const void* sL, const double* cot_m,
                         const double* cot_L, const double* add_m, const double* add_L, double* out_m,
                         double* out_L, double* theta, void* awsv, size_t awsb, hipStream_t st) {
    int rc = adj_check(p, tr, ob, awsb);
    if (rc) return rc;
    if constexpr (kAdj) {
      using namespace adj;
      const AdjPlan a = adj_plan(p);
      if (a.M <= 0) return PSM_OK;
      int use = 0;
      SigmaView<double> sv = sigma_view<double>(sg, &use);
      ModelView<double, R> trv = model_view_d(tr), obv = model_view_d(ob);
      char* aws = (char*)awsv;
      const double* ck = (const double*)(aws + a.off_ck);
      double* pa = (double*)(aws + a.off_pa);
      double* lc = (double*)(aws + a.off_lc);
      const double* ln = (const double*)(aws + a.off_ln);
      double* th = (double*)(aws + a.off_th);
      const unsigned g = grid_for(a.Mp, 64);
      if (mode == MODE_OUTPUT) {
        TimedScope ts_("k_adj_scans", st);
        hipLaunchKernelGGL((k_adj_items<R, NX>), dim3(g), dim3(64), 0, st, a.M, a.Mp, (const R*)sL, cot_m,
                           cot_L, ck, pa);
        if ((rc = hip_check("k_adj_items"))) return rc;
        if ((rc = adj_scan<false, SRC_FWD_VEC>(a, aws, st))) return rc;
        hipLaunchKernelGGL((k_adj_items2<NX>), dim3(g), dim3(64), 0, st, a.M, a.Mp, ck, pa);
        if ((rc = hip_check("k_adj_items2"))) return rc;
        if ((rc = adj_scan<false, SRC_BWD_VEC>(a, aws, st))) return rc;
        if ((rc = adj_scan<true, SRC_BWD_CONG>(a, aws, st))) return rc;
        if ((rc = adj_scan<true, SRC_FWD_CONG>(a, aws, st))) return rc;
      }
      {
        TimedScope ts_("k_adj_chain", st);
        if (mode == MODE_OUTPUT)
          hipLaunchKernelGGL((k_adj_chain<R, NX, NY, MODE_OUTPUT>), dim3(g), dim3(64), 0, st, a.M, a.Mp,
                             p->state_offsets, p->n_seq, trv, obv, sv, use, (const R*)ys, (const R*)m0,
                             (const R*)L0, (const R*)nm, (const R*)nL, (const R*)sm, ck, ln, (const double*)pa,
                             lc, th, a.Mp);
        else
          hipLaunchKernelGGL((k_adj_chain<R, NX, NY, MODE_FISHER>), dim3(g), dim3(64), 0, st, a.M, a.Mp,
                             p->state_offsets, p->n_seq, trv, obv, sv, use, (const R*)ys, (const R*)m0,
                             (const R*)L0, (const R*)nm, (const R*)nL, (const R*)sm, ck, ln, (const double*)pa,
                             lc, th, a.Mp);
        if ((rc = hip_check("k_adj_chain"))) return rc;
      }
      {
        TimedScope ts_("k_adj_lin", st);
        auto lin = [&](auto us, auto part) {
          hipLaunchKernelGGL((k_adj_lin<R, NX, NY, decltype(us)::value, decltype(part)::value>), dim3(g), dim3(64),
                             0, st, a.M, a.Mp, p->state_offsets, p->n_seq, trv, obv, sv, (const R*)nm,
                             (const R*)nL, ln, (const double*)lc, add_m, add_L, out_m, out_L, th, a.Mp);
        };
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        if (use) {
          lin(I1{}, I0{});
          lin(I1{}, I1{});
        } else {
          lin(I0{}, I0{});
          lin(I0{}, I1{});
        }
        if ((rc = hip_check("k_adj_lin"))) return rc;
      }
      if (theta != nullptr) {
        hipLaunchKernelGGL((k_adj_theta<NX, NY>), dim3((unsigned)p->n_seq, (unsigned)Th<NX, NY>::N), dim3(256),
                           0, st, a.Mp, p->state_offsets, (const double*)th, a.Mp, theta);
        if ((rc = hip_check("k_adj_theta"))) return rc;
      }
    }
    return PSM_OK;
  }

  static const OpsTable* table() {
    static const OpsTable t = {&linearize_trans, &linearize_obs, &filter_elements,
                               &filter_combine,  &smoother_elements, &smoother_combine,
                               &workspace_bytes, &filter,          &smoother,
                               &sample_workspace_bytes, &sample, &seq_filter, &seq_smoother,
                               &plan_signature, &adjoint_workspace_bytes, &adjoint_prepare, &adjoint_vjp};
    return &t;
  }
};

}  // namespace psm

// One translation unit instantiates every ny for a given (dtype, form, nx).
#define PSM_INSTANTIATE(TAG, R, NX, NY, SQRT)                        \
  const psm::OpsTable* TAG() { return psm::Impl<R, NX, NY, SQRT>::table(); }
