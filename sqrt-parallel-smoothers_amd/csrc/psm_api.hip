// extern "C" entry points of libparsmooth_hip.so (declared in include/psm.h):
// argument checking, dispatch to the compiled (dtype, form, nx, ny) slice,
// the per-thread error string, the stopping-criterion reduction and the
// hipGraph cache of psm_filter_smoother / psm_iterated_smoother_graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "psm.h"
#include "psm_impl.hpp"
#include "psm_registry.hpp"
#include "psm_util.hpp"

#define PSM_DECL(T, NX, NY) const psm::OpsTable* psm_ops_##T##_##NX##_##NY();
PSM_FOR_ALL(PSM_DECL)
PSM_FOR_DUAL(PSM_DECL)
#undef PSM_DECL

namespace psm {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

static const OpsTable* lookup(int dtype, int form, int nx, int ny) {
  if (nx < 1 || nx > PSM_MAX_NX || ny < 1 || ny > PSM_MAX_NY) return nullptr;
  if (dtype < PSM_F32 || dtype > PSM_F64_TANGENT) return nullptr;
  if (form != PSM_FORM_SQRT && form != PSM_FORM_STD) return nullptr;
  static const OpsTable* table[4][2][PSM_MAX_NX + 1][PSM_MAX_NY + 1] = {};
  static std::once_flag once;
  std::call_once(once, [] {
#define PSM_FILL(T, NX, NY) fill_##T(NX, NY, psm_ops_##T##_##NX##_##NY());
    auto fill_f32s = [](int nx, int ny, const OpsTable* t) { table[PSM_F32][0][nx][ny] = t; };
    auto fill_f32c = [](int nx, int ny, const OpsTable* t) { table[PSM_F32][1][nx][ny] = t; };
    auto fill_f64s = [](int nx, int ny, const OpsTable* t) { table[PSM_F64][0][nx][ny] = t; };
    auto fill_f64c = [](int nx, int ny, const OpsTable* t) { table[PSM_F64][1][nx][ny] = t; };
    auto fill_d32s = [](int nx, int ny, const OpsTable* t) { table[PSM_F32_TANGENT][0][nx][ny] = t; };
    auto fill_d64s = [](int nx, int ny, const OpsTable* t) { table[PSM_F64_TANGENT][0][nx][ny] = t; };
    PSM_FOR_ALL(PSM_FILL)
    PSM_FOR_DUAL(PSM_FILL)
#undef PSM_FILL
  });
  return table[dtype][form][nx][ny];
}

static const OpsTable* need(int dtype, int form, int nx, int ny) {
  const OpsTable* t = lookup(dtype, form, nx, ny);
  if (!t)
    set_error("unsupported combination dtype=" + std::to_string(dtype) + " form=" +
              std::to_string(form) + " nx=" + std::to_string(nx) + " ny=" + std::to_string(ny));
  return t;
}

static int check_model(const psm_model_t* m, int nin, int nout) {
  if (!m) { set_error("null model"); return PSM_E_ARG; }
  if (m->nin != nin || m->nout != nout) {
    set_error("model dims (" + std::to_string(m->nin) + "->" + std::to_string(m->nout) +
              ") do not match (" + std::to_string(nin) + "->" + std::to_string(nout) + ")");
    return PSM_E_ARG;
  }
  if (!m->q_mean || !m->q_chol) { set_error("model noise pointers are null"); return PSM_E_ARG; }
  switch (m->kind) {
    case PSM_MODEL_LINEAR:
      if (!m->params) { set_error("linear model needs its matrix"); return PSM_E_ARG; }
      return PSM_OK;
    case PSM_MODEL_COORDINATED_TURN:
      if (nin != 5 || nout != 5) { set_error("coordinated turn needs nx = 5"); return PSM_E_ARG; }
      if (!m->params) { set_error("coordinated turn needs dt"); return PSM_E_ARG; }
      return PSM_OK;
    case PSM_MODEL_BEARINGS:
      if (nout != 2 || nin < 2) { set_error("bearings needs ny = 2 and nx >= 2"); return PSM_E_ARG; }
      if (!m->params) { set_error("bearings needs sensor positions"); return PSM_E_ARG; }
      return PSM_OK;
    case PSM_MODEL_TABULATED:
    case PSM_MODEL_CONDITIONAL:
      if (!m->params) { set_error("tabulated model needs its value table"); return PSM_E_ARG; }
      return PSM_OK;
    case PSM_MODEL_SIN: case PSM_MODEL_COS: case PSM_MODEL_EXP: case PSM_MODEL_ARCTAN:
      if (nin != nout) { set_error("elementwise model needs nin == nout"); return PSM_E_ARG; }
      return PSM_OK;
    default:
      set_error("unknown model kind " + std::to_string(m->kind));
      return PSM_E_ARG;
  }
}

static int check_problem(const psm_problem_t* p) {
  if (!p) { set_error("null problem"); return PSM_E_ARG; }
  if (p->n_seq < 1 || p->n_states < p->n_seq || !p->state_offsets) {
    set_error("bad sequence layout");
    return PSM_E_ARG;
  }
  if (p->chunk < 0) { set_error("chunk must be >= 0"); return PSM_E_ARG; }
  return PSM_OK;
}

// ---------------------------------------------------------------- criterion
template <typename R>
__global__ __launch_bounds__(1024) void k_mean_sq_diff(int64_t n, const R* __restrict__ a,
                                                       const R* __restrict__ b,
                                                       double* __restrict__ out) {
  __shared__ double red[1024];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double d = (double)a[i] - (double)b[i];
    acc += d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = n > 0 ? red[0] / (double)n : 0.0;
}

// ---------------------------------------------------------------- FP64 division / square root
// The reciprocal and square root every FP64 factorisation uses (psm_linalg.hpp),
// one value per lane, for the tests that compare them with correctly rounded results.
__global__ __launch_bounds__(256) void k_math_check(int64_t n, const double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[i] = rcp_(x[i]);
  q[i] = sqrt_(x[i]);
}

// ---------------------------------------------------------------- graph cache
// Captured launch sequences, keyed by everything the launches depend on (dims,
// the resolved schedule, models, every buffer address, the caller's stream).
// A caller on a stream of its own gets its graph captured on and replayed on
// that stream; callers on the legacy / per-thread default stream (which cannot
// be captured) share one non-blocking side stream, ordered after and before
// the caller's stream by two events created once.  At most kMaxGraphs graphs
// are kept; the least recently used one is dropped (after its last launch has
// finished, which the entry's own event tells) to make room.
struct GraphEntry {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;   // recorded after every launch
  uint64_t used = 0;           // graph-clock value of the last launch
};
constexpr size_t kMaxGraphs = 64;
static std::mutex g_graph_mu;
static std::map<std::vector<uint64_t>, GraphEntry> g_graphs;
static uint64_t g_graph_clock = 0;
static hipStream_t g_side_stream = nullptr;
static hipEvent_t g_side_in = nullptr, g_side_out = nullptr;

static void push(std::vector<uint64_t>& k, const void* p) { k.push_back((uint64_t)(uintptr_t)p); }
static void push_model(std::vector<uint64_t>& k, const psm_model_t* m) {
  if (!m) { k.push_back(~0ull); return; }
  k.push_back((uint64_t)m->kind);
  k.push_back(((uint64_t)m->nin << 32) | (uint32_t)m->nout);
  push(k, m->params); push(k, m->q_mean); push(k, m->q_chol);
}
static void push_problem(std::vector<uint64_t>& k, const psm_problem_t* p, const OpsTable* t) {
  k.push_back(((uint64_t)p->dtype << 48) | ((uint64_t)p->form << 40) | ((uint64_t)p->nx << 32) |
              ((uint64_t)p->ny << 16));
  k.push_back(t->plan_signature(p));
  k.push_back((uint64_t)p->n_seq);
  k.push_back((uint64_t)p->n_states);
  push(k, p->state_offsets);
  uint32_t fr;
  memcpy(&fr, &p->fallback_ratio, sizeof(fr));
  k.push_back((uint64_t)fr);
  push(k, p->fallback_count);
}
static void push_sigma(std::vector<uint64_t>& k, const psm_sigma_t* sigma) {
  if (sigma && sigma->npts > 0) {
    k.push_back((uint64_t)sigma->npts);
    push(k, sigma->xi); push(k, sigma->wm); push(k, sigma->wc);
  } else {
    k.push_back(0);
  }
}

static bool default_stream(hipStream_t st) {
  return st == nullptr || st == hipStreamLegacy || st == hipStreamPerThread;
}

static void drop_graph(GraphEntry& ge) {
  if (ge.done) {
    (void)hipEventSynchronize(ge.done);
    (void)hipEventDestroy(ge.done);
  }
  if (ge.exec) (void)hipGraphExecDestroy(ge.exec);
  if (ge.graph) (void)hipGraphDestroy(ge.graph);
}

// Run `run` (the launch sequence, on the stream it is given) as a graph: captured
// on first use of `key`, replayed afterwards.  A caller whose stream is being
// captured already gets the launches recorded into its own graph instead.
template <typename Fn>
static int run_as_graph(std::vector<uint64_t>& key, hipStream_t st, Fn&& run) {
  hipStreamCaptureStatus cap_st = hipStreamCaptureStatusNone;
  if (!default_stream(st) && hipStreamIsCapturing(st, &cap_st) == hipSuccess &&
      cap_st != hipStreamCaptureStatusNone)
    return run(st);
  push(key, st);
  std::lock_guard<std::mutex> lock(g_graph_mu);
  const bool side = default_stream(st);
  if (side && g_side_stream == nullptr) {
    if (hipStreamCreateWithFlags(&g_side_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&g_side_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g_side_out, hipEventDisableTiming) != hipSuccess) {
      set_error("psm graph: side stream / event creation failed");
      return PSM_E_HIP;
    }
  }
  hipStream_t cs = side ? g_side_stream : st;
  auto it = g_graphs.find(key);
  if (it == g_graphs.end()) {
    if (g_graphs.size() >= kMaxGraphs) {
      auto lru = g_graphs.begin();
      for (auto j = g_graphs.begin(); j != g_graphs.end(); ++j)
        if (j->second.used < lru->second.used) lru = j;
      drop_graph(lru->second);
      g_graphs.erase(lru);
    }
    GraphEntry ge;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      set_error("hipStreamBeginCapture failed");
      return PSM_E_HIP;
    }
    int r = run(cs);
    hipError_t e = hipStreamEndCapture(cs, &ge.graph);
    if (r) { if (ge.graph) (void)hipGraphDestroy(ge.graph); return r; }
    if (e != hipSuccess) { set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e)); return PSM_E_HIP; }
    e = hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(ge.graph);
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      return PSM_E_HIP;
    }
    e = hipEventCreateWithFlags(&ge.done, hipEventDisableTiming);
    if (e != hipSuccess) {
      ge.done = nullptr;
      drop_graph(ge);
      set_error(std::string("hipEventCreate: ") + hipGetErrorString(e));
      return PSM_E_HIP;
    }
    it = g_graphs.emplace(key, ge).first;
  }
  hipError_t e;
  if (side) {
    (void)hipEventRecord(g_side_in, st);
    (void)hipStreamWaitEvent(cs, g_side_in, 0);
    e = hipGraphLaunch(it->second.exec, cs);
    (void)hipEventRecord(g_side_out, cs);
    (void)hipStreamWaitEvent(st, g_side_out, 0);
  } else {
    e = hipGraphLaunch(it->second.exec, cs);
  }
  (void)hipEventRecord(it->second.done, cs);
  it->second.used = ++g_graph_clock;
  if (e != hipSuccess) {
    set_error(std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    return PSM_E_HIP;
  }
  return PSM_OK;
}

// ---------------------------------------------------------------- timing
struct TimingRec {
  const char* name;
  hipEvent_t a, b;
};
static std::mutex g_tmu;
static bool g_timing = false;
static std::vector<TimingRec> g_trecs;

// set while psm_filter_smoother_noalloc runs on this thread: no timing events
// are created then (creating them allocates)
static thread_local bool g_no_alloc = false;

int timing_begin(const char* name, hipStream_t st) {
  if (!g_timing || g_no_alloc) return -1;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone) return -1;
  TimingRec r{name, nullptr, nullptr};
  if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
  (void)hipEventRecord(r.a, st);
  std::lock_guard<std::mutex> lock(g_tmu);
  g_trecs.push_back(r);
  return (int)g_trecs.size() - 1;
}

void timing_end(int slot, hipStream_t st) {
  if (slot < 0) return;
  hipEvent_t b;
  {
    std::lock_guard<std::mutex> lock(g_tmu);
    if (slot >= (int)g_trecs.size()) return;
    b = g_trecs[slot].b;
  }
  (void)hipEventRecord(b, st);
}

}  // namespace psm

using namespace psm;

extern "C" {

int psm_version(void) { return 1; }

const char* psm_last_error(void) { return g_err.c_str(); }

int psm_supported(int32_t dtype, int32_t form, int32_t nx, int32_t ny) {
  return lookup(dtype, form, nx, ny) != nullptr ? 1 : 0;
}

int psm_linearize(int32_t dtype, int32_t form, const psm_model_t* model, const psm_sigma_t* sigma,
                  int64_t n, const void* x_mean, const void* x_chol, void* F, void* C, void* b,
                  void* stream) {
  if (!model) { set_error("null model"); return PSM_E_ARG; }
  const int nx = model->nin, nout = model->nout;
  int rc = check_model(model, nx, nout);
  if (rc) return rc;
  if (nout == nx) {
    const OpsTable* t = need(dtype, form, nx, 1);
    if (!t) return PSM_E_UNSUPPORTED;
    return t->linearize_trans(model, sigma, n, x_mean, x_chol, F, C, b, (hipStream_t)stream);
  }
  const OpsTable* t = need(dtype, form, nx, nout);
  if (!t) return PSM_E_UNSUPPORTED;
  return t->linearize_obs(model, sigma, n, x_mean, x_chol, F, C, b, (hipStream_t)stream);
}

int psm_filter_elements(int32_t dtype, int32_t form, int32_t nx, int32_t ny, int64_t n,
                        const void* F, const void* cq, const void* bt, const void* H,
                        const void* cr, const void* c, const void* m, const void* L,
                        const void* y, void* A, void* b, void* U, void* eta, void* Z,
                        void* stream) {
  const OpsTable* t = need(dtype, form, nx, ny);
  if (!t) return PSM_E_UNSUPPORTED;
  return t->filter_elements(n, F, cq, bt, H, cr, c, m, L, y, A, b, U, eta, Z,
                            (hipStream_t)stream);
}

int psm_filter_combine(int32_t dtype, int32_t form, int32_t nx, int64_t n, const void* const* e1,
                       const void* const* e2, void* const* out, void* stream) {
  const OpsTable* t = need(dtype, form, nx, 1);
  if (!t) return PSM_E_UNSUPPORTED;
  if (!e1 || !e2 || !out) { set_error("null element pointer array"); return PSM_E_ARG; }
  return t->filter_combine(n, e1, e2, out, (hipStream_t)stream);
}

int psm_smoother_elements(int32_t dtype, int32_t form, int32_t nx, int64_t n, const void* F,
                          const void* cq, const void* bt, const void* m, const void* L, void* g,
                          void* E, void* D, void* stream) {
  const OpsTable* t = need(dtype, form, nx, 1);
  if (!t) return PSM_E_UNSUPPORTED;
  return t->smoother_elements(n, F, cq, bt, m, L, g, E, D, (hipStream_t)stream);
}

int psm_smoother_combine(int32_t dtype, int32_t form, int32_t nx, int64_t n, const void* const* e1,
                         const void* const* e2, void* const* out, void* stream) {
  const OpsTable* t = need(dtype, form, nx, 1);
  if (!t) return PSM_E_UNSUPPORTED;
  if (!e1 || !e2 || !out) { set_error("null element pointer array"); return PSM_E_ARG; }
  return t->smoother_combine(n, e1, e2, out, (hipStream_t)stream);
}

size_t psm_workspace_bytes(const psm_problem_t* p) {
  if (check_problem(p)) return 0;
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return 0;
  return t->workspace_bytes(p);
}

int psm_chunk_size(const psm_problem_t* p) {
  if (check_problem(p)) return 0;
  return make_plan(p, 8, 1, 1, 1).chunk;
}

int psm_filter(const psm_problem_t* p, const psm_model_t* trans, const psm_model_t* obs,
               const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
               const void* nom_m, const void* nom_L, void* out_m, void* out_L, double* ell,
               void* ws, size_t wsb, void* stream) {
  int rc = check_problem(p);
  if (rc) return rc;
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  if ((rc = check_model(obs, p->nx, p->ny))) return rc;
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  if (!ys || !m0 || !L0 || !nom_m || !nom_L || !out_m || !out_L) {
    set_error("psm_filter: null array");
    return PSM_E_ARG;
  }
  return t->filter(p, trans, obs, sigma, ys, m0, L0, nom_m, nom_L, out_m, out_L, ell, ws, wsb,
                   (hipStream_t)stream, 0);
}

int psm_smoother(const psm_problem_t* p, const psm_model_t* trans, const psm_sigma_t* sigma,
                 const void* fm, const void* fL, const void* nom_m, const void* nom_L, void* out_m,
                 void* out_L, void* ws, size_t wsb, void* stream) {
  int rc = check_problem(p);
  if (rc) return rc;
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  if (!fm || !fL || !nom_m || !nom_L || !out_m || !out_L) {
    set_error("psm_smoother: null array");
    return PSM_E_ARG;
  }
  return t->smoother(p, trans, sigma, fm, fL, nom_m, nom_L, out_m, out_L, ws, wsb,
                     (hipStream_t)stream, 0);
}

int psm_filter_smoother(const psm_problem_t* p, const psm_model_t* trans, const psm_model_t* obs,
                        const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                        const void* nom_m, const void* nom_L, void* filt_m, void* filt_L,
                        void* out_m, void* out_L, void* ws, size_t wsb, int32_t use_graph,
                        void* stream) {
  int rc = check_problem(p);
  if (rc) return rc;
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  if ((rc = check_model(obs, p->nx, p->ny))) return rc;
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  if (!ys || !m0 || !L0 || !nom_m || !nom_L || !filt_m || !filt_L || !out_m || !out_L) {
    set_error("psm_filter_smoother: null array");
    return PSM_E_ARG;
  }
  auto run = [&](hipStream_t s) -> int {
    int r = t->filter(p, trans, obs, sigma, ys, m0, L0, nom_m, nom_L, filt_m, filt_L, nullptr, ws,
                      wsb, s, 0);
    if (r) return r;
    return t->smoother(p, trans, sigma, filt_m, filt_L, nom_m, nom_L, out_m, out_L, ws, wsb, s, 1);
  };
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return run(st);
  std::vector<uint64_t> key = {0x46534d4f4f5448ull};  // one filter-smoother pass
  push_problem(key, p, t);
  push_model(key, trans);
  push_model(key, obs);
  push_sigma(key, sigma);
  for (const void* q : {ys, m0, L0, nom_m, nom_L, (const void*)filt_m, (const void*)filt_L,
                        (const void*)out_m, (const void*)out_L, (const void*)ws})
    push(key, q);
  key.push_back((uint64_t)wsb);
  return run_as_graph(key, st, run);
}

int psm_iterated_smoother_graph(const psm_problem_t* p, const psm_model_t* trans,
                                const psm_model_t* obs, const psm_sigma_t* sigma, const void* ys,
                                const void* m0, const void* L0, const void* init_m,
                                const void* init_L, int32_t n_iter, void* filt_m, void* filt_L,
                                void* out0_m, void* out0_L, void* out1_m, void* out1_L,
                                double* ell, void* ws, size_t wsb, int32_t use_graph,
                                void* stream) {
  int rc = check_problem(p);
  if (rc) return rc;
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  if ((rc = check_model(obs, p->nx, p->ny))) return rc;
  if (n_iter < 1) { set_error("psm_iterated_smoother_graph: n_iter must be >= 1"); return PSM_E_ARG; }
  if (!ys || !m0 || !L0 || !init_m || !init_L || !filt_m || !filt_L || !out0_m || !out0_L ||
      (n_iter > 1 && (!out1_m || !out1_L))) {
    set_error("psm_iterated_smoother_graph: null array");
    return PSM_E_ARG;
  }
  if (out0_m == init_m || out0_L == init_L) {
    set_error("psm_iterated_smoother_graph: out0 must not alias the initial nominal");
    return PSM_E_ARG;
  }
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  void* bm[2] = {out0_m, out1_m};
  void* bL[2] = {out0_L, out1_L};
  // Between passes the filtered and smoothed trajectories stay in the
  // workspace in the state-interleaved layout (null pointers select it), which
  // the lane-per-chunk kernels read and write coalesced; the last pass writes
  // its smoothed trajectory to out0 / out1.
  auto run = [&](hipStream_t s) -> int {
    const void* nm = init_m;
    const void* nL = init_L;
    for (int k = 0; k < n_iter; ++k) {
      const bool last = k + 1 == n_iter;
      int r = t->filter(p, trans, obs, sigma, ys, m0, L0, nm, nL, nullptr, nullptr, nullptr, ws, wsb, s, 0);
      if (r) return r;
      r = t->smoother(p, trans, sigma, nullptr, nullptr, nm, nL, last ? bm[k & 1] : nullptr,
                      last ? bL[k & 1] : nullptr, ws, wsb, s, 1);
      if (r) return r;
      nm = last ? bm[k & 1] : nullptr;
      nL = last ? bL[k & 1] : nullptr;
    }
    if (ell != nullptr)
      return t->filter(p, trans, obs, sigma, ys, m0, L0, nm, nL, filt_m, filt_L, ell, ws, wsb, s, 0);
    return PSM_OK;
  };
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return run(st);
  std::vector<uint64_t> key = {0x49544552415445ull, (uint64_t)n_iter};  // the whole iteration
  push_problem(key, p, t);
  push_model(key, trans);
  push_model(key, obs);
  push_sigma(key, sigma);
  for (const void* q : {ys, m0, L0, init_m, init_L, (const void*)filt_m, (const void*)filt_L,
                        (const void*)out0_m, (const void*)out0_L, (const void*)out1_m,
                        (const void*)out1_L, (const void*)ell, (const void*)ws})
    push(key, q);
  key.push_back((uint64_t)wsb);
  return run_as_graph(key, st, run);
}

// ---------------------------------------------------------------- library-owned workspace
// The workspace of psm_filter_smoother_noalloc: a static device array that each
// GPU holds from the library's load until the process ends.  Host code never
// uses its name as an address: the device address of each GPU's copy is taken
// once per device with hipGetSymbolAddress and kept in g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[PSM_STATIC_WORKSPACE_BYTES];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static int scratch_of_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error(std::string("psm_filter_smoother_noalloc: no usable current device (") +
              hipGetErrorString(e) + ")");
    return PSM_E_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      set_error(std::string("psm_filter_smoother_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return PSM_E_HIP;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return PSM_OK;
}

int psm_filter_smoother_noalloc(const psm_problem_t* p, const psm_model_t* trans,
                                const psm_model_t* obs, const psm_sigma_t* sigma, const void* ys,
                                const void* m0, const void* L0, const void* nom_m,
                                const void* nom_L, void* filt_m, void* filt_L, void* out_m,
                                void* out_L, void* stream) {
  int rc = check_problem(p);
  if (rc) return rc;
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  if ((rc = check_model(obs, p->nx, p->ny))) return rc;
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  if (!ys || !m0 || !L0 || !nom_m || !nom_L || !filt_m || !filt_L || !out_m || !out_L) {
    set_error("psm_filter_smoother_noalloc: null array");
    return PSM_E_ARG;
  }
  const size_t need_bytes = t->workspace_bytes(p);
  if (need_bytes > (size_t)PSM_STATIC_WORKSPACE_BYTES) {
    set_error("psm_filter_smoother_noalloc: this problem needs " + std::to_string(need_bytes) +
              " bytes of workspace; the library's static workspace holds " +
              std::to_string((size_t)PSM_STATIC_WORKSPACE_BYTES) +
              " (use psm_filter_smoother with a caller-supplied workspace)");
    return PSM_E_WORKSPACE;
  }
  void* ws = nullptr;
  if ((rc = scratch_of_current_device(&ws))) return rc;
  // every kernel of the pass writes each workspace word before reading it, so
  // nothing there needs clearing between calls
  g_no_alloc = true;
  hipStream_t st = (hipStream_t)stream;
  rc = t->filter(p, trans, obs, sigma, ys, m0, L0, nom_m, nom_L, filt_m, filt_L, nullptr, ws,
                 PSM_STATIC_WORKSPACE_BYTES, st, 0);
  if (rc == PSM_OK)
    rc = t->smoother(p, trans, sigma, filt_m, filt_L, nom_m, nom_L, out_m, out_L, ws,
                     PSM_STATIC_WORKSPACE_BYTES, st, 1);
  g_no_alloc = false;
  return rc;
}

// ---------------------------------------------------------------- reverse-mode adjoint
size_t psm_adjoint_workspace_bytes(const psm_problem_t* p) {
  if (check_problem(p)) return 0;
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return 0;
  return t->adjoint_workspace_bytes(p);
}

int psm_adjoint_prepare(const psm_problem_t* p, const psm_model_t* trans, const psm_model_t* obs,
                        const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                        const void* nom_m, const void* nom_L, void* filt_m, void* filt_L, void* out_m,
                        void* out_L, void* ws, size_t wsb, void* aws, size_t awsb, void* stream) {
  int rc = check_problem(p);
  if (rc) return rc;
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  if ((rc = check_model(obs, p->nx, p->ny))) return rc;
  if (!ys || !m0 || !L0 || !nom_m || !nom_L || !filt_m || !filt_L || !out_m || !out_L || !aws) {
    set_error("psm_adjoint_prepare: null array");
    return PSM_E_ARG;
  }
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  rc = t->filter(p, trans, obs, sigma, ys, m0, L0, nom_m, nom_L, filt_m, filt_L, nullptr, ws, wsb, st, 0);
  if (rc) return rc;
  rc = t->smoother(p, trans, sigma, filt_m, filt_L, nom_m, nom_L, out_m, out_L, ws, wsb, st, 1);
  if (rc) return rc;
  return t->adjoint_prepare(p, trans, obs, sigma, nom_m, nom_L, filt_m, filt_L, out_m, out_L, aws, awsb, st);
}

static int adjoint_common(const psm_problem_t* p, const psm_model_t* trans, const psm_model_t* obs,
                          const OpsTable** t) {
  int rc = check_problem(p);
  if (rc) return rc;
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  if ((rc = check_model(obs, p->nx, p->ny))) return rc;
  *t = need(p->dtype, p->form, p->nx, p->ny);
  return *t ? PSM_OK : PSM_E_UNSUPPORTED;
}

int psm_filter_smoother_vjp(const psm_problem_t* p, const psm_model_t* trans, const psm_model_t* obs,
                            const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                            const void* nom_m, const void* nom_L, const void* sm_m, const void* sm_L,
                            const double* cot_m, const double* cot_L, const double* add_m,
                            const double* add_L, double* out_m, double* out_L, double* theta_bar,
                            void* aws, size_t awsb, void* stream) {
  const OpsTable* t = nullptr;
  int rc = adjoint_common(p, trans, obs, &t);
  if (rc) return rc;
  if (!cot_m || !cot_L || !out_m || !out_L || !sm_m || !sm_L || !aws || (!add_m) != (!add_L)) {
    set_error("psm_filter_smoother_vjp: null array");
    return PSM_E_ARG;
  }
  return t->adjoint_vjp(p, 0, trans, obs, sigma, ys, m0, L0, nom_m, nom_L, sm_m, sm_L, cot_m, cot_L, add_m,
                        add_L, out_m, out_L, theta_bar, aws, awsb, (hipStream_t)stream);
}

int psm_filter_vjp(const psm_problem_t* p, const psm_model_t* trans, const psm_model_t* obs,
                   const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                   const void* nom_m, const void* nom_L, const void* sm_m, const void* sm_L, double* out_m,
                   double* out_L, double* theta_bar, void* aws, size_t awsb, void* stream) {
  const OpsTable* t = nullptr;
  int rc = adjoint_common(p, trans, obs, &t);
  if (rc) return rc;
  if (!out_m || !out_L || !sm_m || !sm_L || !aws) {
    set_error("psm_filter_vjp: null array");
    return PSM_E_ARG;
  }
  return t->adjoint_vjp(p, 1, trans, obs, sigma, ys, m0, L0, nom_m, nom_L, sm_m, sm_L, nullptr, nullptr,
                        nullptr, nullptr, out_m, out_L, theta_bar, aws, awsb, (hipStream_t)stream);
}

size_t psm_sample_workspace_bytes(const psm_problem_t* p, int32_t n_samples) {
  if (check_problem(p)) return 0;
  if (n_samples < 1) { set_error("n_samples must be >= 1"); return 0; }
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return 0;
  return t->sample_workspace_bytes(p, n_samples);
}

int psm_sample(const psm_problem_t* p, const psm_model_t* trans, const psm_sigma_t* sigma,
               const void* filt_m, const void* filt_L, const void* nom_m, const void* nom_L,
               const void* eps, int32_t n_samples, void* out, void* ws, size_t wsb,
               void* stream) {
  int rc = check_problem(p);
  if (rc) return rc;
  if (p->dtype != PSM_F32 && p->dtype != PSM_F64) {
    set_error("psm_sample: dtype must be PSM_F32 or PSM_F64");
    return PSM_E_ARG;
  }
  if (n_samples < 1) { set_error("n_samples must be >= 1"); return PSM_E_ARG; }
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  if (!filt_m || !filt_L || !nom_m || !nom_L || !eps || !out) {
    set_error("psm_sample: null array");
    return PSM_E_ARG;
  }
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  return t->sample(p, trans, sigma, filt_m, filt_L, nom_m, nom_L, eps, n_samples, out, ws, wsb,
                   (hipStream_t)stream);
}

// ---------------------------------------------------------------- sequential passes
static bool caller_evaluated(const psm_model_t* m) {
  return m->kind == PSM_MODEL_TABULATED || m->kind == PSM_MODEL_CONDITIONAL;
}

static int check_seq(const psm_problem_t* p, const void* nom_m, const void* nom_L,
                     const psm_model_t* a, const psm_model_t* b, const char* who) {
  if (p->dtype != PSM_F32 && p->dtype != PSM_F64) {
    set_error(std::string(who) + ": dtype must be PSM_F32 or PSM_F64");
    return PSM_E_ARG;
  }
  if ((nom_m == nullptr) != (nom_L == nullptr)) {
    set_error(std::string(who) + ": nom_m and nom_L must both be given or both be NULL");
    return PSM_E_ARG;
  }
  if (nom_m == nullptr && (caller_evaluated(a) || (b && caller_evaluated(b)))) {
    set_error(std::string(who) +
              ": models evaluated by the caller need a nominal trajectory (the running "
              "estimate is only known inside the recursion)");
    return PSM_E_ARG;
  }
  return PSM_OK;
}

int psm_seq_filter(const psm_problem_t* p, const psm_model_t* trans, const psm_model_t* obs,
                   const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                   const void* nom_m, const void* nom_L, void* out_m, void* out_L, double* ell,
                   void* stream) {
  int rc = check_problem(p);
  if (rc) return rc;
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  if ((rc = check_model(obs, p->nx, p->ny))) return rc;
  if ((rc = check_seq(p, nom_m, nom_L, trans, obs, "psm_seq_filter"))) return rc;
  if (!ys || !m0 || !L0 || !out_m || !out_L) {
    set_error("psm_seq_filter: null array");
    return PSM_E_ARG;
  }
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  return t->seq_filter(p, trans, obs, sigma, ys, m0, L0, nom_m, nom_L, out_m, out_L, ell,
                       (hipStream_t)stream);
}

int psm_seq_smoother(const psm_problem_t* p, const psm_model_t* trans, const psm_sigma_t* sigma,
                     const void* fm, const void* fL, const void* nom_m, const void* nom_L,
                     void* out_m, void* out_L, void* stream) {
  int rc = check_problem(p);
  if (rc) return rc;
  if ((rc = check_model(trans, p->nx, p->nx))) return rc;
  if ((rc = check_seq(p, nom_m, nom_L, trans, nullptr, "psm_seq_smoother"))) return rc;
  if (!fm || !fL || !out_m || !out_L) {
    set_error("psm_seq_smoother: null array");
    return PSM_E_ARG;
  }
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  return t->seq_smoother(p, trans, sigma, fm, fL, nom_m, nom_L, out_m, out_L, (hipStream_t)stream);
}

int psm_mean_sq_diff(int32_t dtype, int64_t n, const void* a, const void* b, double* out,
                     void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSM_F64)
    hipLaunchKernelGGL(k_mean_sq_diff<double>, dim3(1), dim3(1024), 0, st, n, (const double*)a,
                       (const double*)b, out);
  else if (dtype == PSM_F32)
    hipLaunchKernelGGL(k_mean_sq_diff<float>, dim3(1), dim3(1024), 0, st, n, (const float*)a,
                       (const float*)b, out);
  else { set_error("bad dtype"); return PSM_E_ARG; }
  return hip_check("k_mean_sq_diff");
}

int psm_math_check(int64_t n, const double* x, double* out_rcp, double* out_sqrt, void* stream) {
  if (n < 0 || (n > 0 && (!x || !out_rcp || !out_sqrt))) {
    set_error("psm_math_check: bad arguments");
    return PSM_E_ARG;
  }
  if (n == 0) return PSM_OK;
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) {
    set_error("psm_math_check: n too large");
    return PSM_E_ARG;
  }
  hipLaunchKernelGGL(k_math_check, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, n, x,
                     out_rcp, out_sqrt);
  return hip_check("k_math_check");
}

// Grid of one lane per batch item for the runtime-size helpers; 0 when nb is
// out of range (the caller reports it).
static unsigned util_grid(int64_t nb) {
  const int64_t g = (nb + kUtilWG - 1) / kUtilWG;
  return (g < 1 || g > 0x7fffffff) ? 0u : (unsigned)g;
}

int psm_tria(int32_t dtype, int64_t nb, int32_t rows, int32_t cols, void* A, void* stream) {
  if (nb < 0 || rows < 1 || cols < 1 || (nb > 0 && !A) || (nb > 0 && util_grid(nb) == 0)) {
    set_error("psm_tria: bad arguments");
    return PSM_E_ARG;
  }
  if (nb == 0) return PSM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSM_F64)
    hipLaunchKernelGGL(k_util_tria<double>, dim3(util_grid(nb)), dim3(kUtilWG), 0, st, nb, rows, cols,
                       (double*)A);
  else if (dtype == PSM_F32)
    hipLaunchKernelGGL(k_util_tria<float>, dim3(util_grid(nb)), dim3(kUtilWG), 0, st, nb, rows, cols,
                       (float*)A);
  else { set_error("psm_tria: bad dtype"); return PSM_E_ARG; }
  return hip_check("k_util_tria");
}

int psm_cholesky_update(int32_t dtype, int64_t nb, int32_t n, int32_t nv, void* L, const void* V,
                        double multiplier, void* work, void* stream) {
  if (nb < 0 || n < 1 || nv < 0 || (nb > 0 && (!L || !work || (nv > 0 && !V))) ||
      (nb > 0 && util_grid(nb) == 0)) {
    set_error("psm_cholesky_update: bad arguments");
    return PSM_E_ARG;
  }
  if (nb == 0 || nv == 0) return PSM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSM_F64)
    hipLaunchKernelGGL(k_util_chol_update<double>, dim3(util_grid(nb)), dim3(kUtilWG), 0, st, nb, n, nv,
                       (double*)L, (const double*)V, (double*)work, multiplier);
  else if (dtype == PSM_F32)
    hipLaunchKernelGGL(k_util_chol_update<float>, dim3(util_grid(nb)), dim3(kUtilWG), 0, st, nb, n, nv,
                       (float*)L, (const float*)V, (float*)work, (float)multiplier);
  else { set_error("psm_cholesky_update: bad dtype"); return PSM_E_ARG; }
  return hip_check("k_util_chol_update");
}

int psm_mvn_loglikelihood(int32_t dtype, int64_t nb, int32_t n, const void* x, const void* L,
                          double* out, double* work, void* stream) {
  if (nb < 0 || n < 1 || (nb > 0 && (!x || !L || !out || !work)) || (nb > 0 && util_grid(nb) == 0)) {
    set_error("psm_mvn_loglikelihood: bad arguments");
    return PSM_E_ARG;
  }
  if (nb == 0) return PSM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSM_F64)
    hipLaunchKernelGGL(k_util_mvn_ll<double>, dim3(util_grid(nb)), dim3(kUtilWG), 0, st, nb, n,
                       (const double*)x, (const double*)L, out, work);
  else if (dtype == PSM_F32)
    hipLaunchKernelGGL(k_util_mvn_ll<float>, dim3(util_grid(nb)), dim3(kUtilWG), 0, st, nb, n,
                       (const float*)x, (const float*)L, out, work);
  else { set_error("psm_mvn_loglikelihood: bad dtype"); return PSM_E_ARG; }
  return hip_check("k_util_mvn_ll");
}

void psm_timing_enable(int32_t enable) {
  std::lock_guard<std::mutex> lock(g_tmu);
  g_timing = enable != 0;
}

int psm_timing_report(char* buf, size_t len) {
  std::vector<TimingRec> recs;
  {
    std::lock_guard<std::mutex> lock(g_tmu);
    recs.swap(g_trecs);
  }
  std::map<std::string, std::pair<double, long>> agg;
  for (auto& r : recs) {
    float ms = 0.f;
    (void)hipEventSynchronize(r.b);
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto& e = agg[r.name];
      e.first += ms;
      e.second += 1;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  std::string out = "{";
  bool first = true;
  for (auto& kv : agg) {
    char tmp[256];
    snprintf(tmp, sizeof(tmp), "%s\"%s\": [%.6f, %ld]", first ? "" : ", ", kv.first.c_str(),
             kv.second.first, kv.second.second);
    out += tmp;
    first = false;
  }
  out += "}";
  if (buf && len > 0) {
    size_t n = out.size() < len - 1 ? out.size() : len - 1;
    memcpy(buf, out.data(), n);
    buf[n] = 0;
  }
  return (int)out.size();
}

// ---------------------------------------------------------------- autotuner
struct TuneResult {
  int32_t chunk, scan_k;
  float ms;
};
static std::mutex g_tune_mu;
static std::map<std::vector<uint64_t>, TuneResult> g_tuned;

int psm_autotune(const psm_problem_t* p, const psm_model_t* trans, const psm_model_t* obs,
                 const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                 const void* nom_m, const void* nom_L, void* filt_m, void* filt_L, void* out_m,
                 void* out_L, void* ws, size_t wsb, void* stream, int32_t* best_chunk,
                 int32_t* best_scan_k, float* best_ms) {
  int rc = check_problem(p);
  if (rc) return rc;
  if (best_chunk == nullptr || best_scan_k == nullptr) {
    set_error("psm_autotune: null output pointer");
    return PSM_E_ARG;
  }
  const OpsTable* t = need(p->dtype, p->form, p->nx, p->ny);
  if (!t) return PSM_E_UNSUPPORTED;
  std::vector<uint64_t> key = {(uint64_t)p->dtype, (uint64_t)p->form, (uint64_t)p->nx,
                               (uint64_t)p->ny, (uint64_t)p->n_seq, (uint64_t)p->n_states,
                               (uint64_t)(trans ? trans->kind : -1), (uint64_t)(obs ? obs->kind : -1),
                               (uint64_t)(sigma ? sigma->npts : 0)};
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto hit = g_tuned.find(key);
  if (hit != g_tuned.end()) {
    *best_chunk = hit->second.chunk;
    *best_scan_k = hit->second.scan_k;
    if (best_ms) *best_ms = hit->second.ms;
    return PSM_OK;
  }
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    set_error("psm_autotune: hipEventCreate failed");
    return PSM_E_HIP;
  }
  // time one candidate: a warm-up pass, then kReps passes between events
  constexpr int kReps = 2;
  auto time_one = [&](int32_t c, int32_t k, float* ms) -> int {
    psm_problem_t q = *p;
    q.chunk = c;
    q.scan_k = k;
    *ms = -1.0f;
    if (t->workspace_bytes(&q) > wsb) return PSM_OK;  // skipped: does not fit
    int r = psm_filter_smoother(&q, trans, obs, sigma, ys, m0, L0, nom_m, nom_L, filt_m, filt_L,
                                out_m, out_L, ws, wsb, 0, stream);
    if (r) return r;
    (void)hipEventRecord(e0, st);
    for (int i = 0; i < kReps && r == 0; ++i)
      r = psm_filter_smoother(&q, trans, obs, sigma, ys, m0, L0, nom_m, nom_L, filt_m, filt_L,
                              out_m, out_L, ws, wsb, 0, stream);
    (void)hipEventRecord(e1, st);
    if (r) return r;
    if (hipEventSynchronize(e1) != hipSuccess) {
      set_error("psm_autotune: hipEventSynchronize failed");
      return PSM_E_HIP;
    }
    float v = 0.0f;
    (void)hipEventElapsedTime(&v, e0, e1);
    *ms = v / kReps;
    return PSM_OK;
  };
  // stage 1: power-of-two chunks x fan-in {2, 4, 8} x {per-lane, team upper
  // levels, half-wave level 0, both, team + quad level 0, team + quad +
  // half-wave}; stage 2: chunks between the stage-1
  // sizes (5/8 .. 7/4 of the best one) with the three best scan modes.  Chunk
  // sizes that are not powers of two avoid a nearly empty last workgroup when
  // the state count is just above a power of two (T + 1 states).
  static const int32_t chunks[] = {2, 4, 8, 16, 32};
  static const int32_t ks[] = {2, 4, 8};
  const bool team_ok = p->form == PSM_FORM_SQRT && (p->dtype == PSM_F32 || p->dtype == PSM_F64);
  std::vector<TuneResult> res;
  for (int32_t c : chunks) {
    static const int32_t kModes[] = {0, PSM_SCAN_TEAM, PSM_SCAN_HALF, PSM_SCAN_TEAM | PSM_SCAN_HALF,
                                     PSM_SCAN_TEAM | PSM_SCAN_QUAD,
                                     PSM_SCAN_TEAM | PSM_SCAN_QUAD | PSM_SCAN_HALF};
    for (int32_t kk = 0; kk < 18 && rc == 0; ++kk) {
      const int32_t mode = kk / 3;
      const int32_t k = ks[kk % 3] | kModes[mode];
      if (mode > 0 && !team_ok) continue;
      float ms;
      rc = time_one(c, k, &ms);
      if (rc == 0 && ms >= 0.0f) res.push_back(TuneResult{c, k, ms});
    }
    if (rc) break;
  }
  if (rc == 0 && !res.empty()) {
    std::sort(res.begin(), res.end(),
              [](const TuneResult& a, const TuneResult& b) { return a.ms < b.ms; });
    std::vector<int32_t> modes;
    for (const TuneResult& r : res) {
      if (std::find(modes.begin(), modes.end(), r.scan_k) == modes.end()) modes.push_back(r.scan_k);
      if (modes.size() == 3) break;
    }
    const int32_t c0 = res.front().chunk;
    std::vector<int32_t> fine;
    for (int32_t num : {5, 6, 7, 10, 12, 14}) {  // eighths of c0
      const int32_t c = (c0 * num + 4) / 8;
      if (c < 2 || c > 64 || (c & (c - 1)) == 0) continue;
      if (std::find(fine.begin(), fine.end(), c) == fine.end()) fine.push_back(c);
    }
    for (int32_t c : fine) {
      for (int32_t k : modes) {
        float ms;
        rc = time_one(c, k, &ms);
        if (rc) break;
        if (ms >= 0.0f) res.push_back(TuneResult{c, k, ms});
      }
      if (rc) break;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (res.empty()) {
    set_error("psm_autotune: no candidate schedule fits the workspace");
    return PSM_E_WORKSPACE;
  }
  TuneResult best = res.front();
  for (const TuneResult& r : res)
    if (r.ms < best.ms) best = r;
  g_tuned[key] = best;
  *best_chunk = best.chunk;
  *best_scan_k = best.scan_k;
  if (best_ms) *best_ms = best.ms;
  return PSM_OK;
}

void psm_clear_graphs(void) {
  std::lock_guard<std::mutex> lock(g_graph_mu);
  for (auto& kv : g_graphs) drop_graph(kv.second);
  g_graphs.clear();
}

}  // extern "C"
