/*
 * psm.h -- C ABI of libparsmooth_hip.so, the MI355X (gfx950) implementation of
 * parsmooth's parallel square-root filtering / smoothing path.
 *
 * Every entry point is stream-ordered on the caller's HIP stream, takes plain
 * device pointers and sizes, never allocates or frees caller memory, and
 * returns 0 on success or a negative PSM_E* status (psm_last_error() gives the
 * message of the last failure on the calling thread).  Scratch memory for the
 * scans is a caller-supplied workspace sized by psm_workspace_bytes().
 *
 * Arrays are C-contiguous and use the reference's shapes: trajectories are
 * (n_states, nx) means and (n_states, nx, nx) chol factors (square-root form)
 * or covariances (covariance form); observations are (n_obs, ny).
 *
 * A ragged batch of sequences is laid out back to back: sequence s owns
 * states [state_offsets[s], state_offsets[s+1]) (T_s + 1 states: the prior
 * state 0 then one state per observation) and observations
 * [state_offsets[s] - s, state_offsets[s+1] - s - 1).
 */
#ifndef PSM_H_
#define PSM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dtypes */
#define PSM_F32 0
#define PSM_F64 1
/* Tangent (forward-mode) dtypes, square-root form only: every array element is
 * a (value, tangent) pair of float / double (so a (n, nx) trajectory is stored
 * as (n, nx, 2)); the log-likelihood output is a (value, tangent) pair of
 * doubles per sequence.  Passing tangents of the inputs (model parameters and
 * noise, prior, observations, nominal trajectory) yields the directional
 * derivative of every output along them. */
#define PSM_F32_TANGENT 2
#define PSM_F64_TANGENT 3

/* forms: MVNSqrt (chol factors) or MVNStandard (covariances) */
#define PSM_FORM_SQRT 0
#define PSM_FORM_STD 1

/* built-in model functions (FunctionalModel.function) */
#define PSM_MODEL_LINEAR 0            /* x -> A x, params = A (nout x nin) */
#define PSM_MODEL_COORDINATED_TURN 1  /* params = [dt], nin = nout = 5 */
#define PSM_MODEL_BEARINGS 2          /* params = [s1x, s1y, s2x, s2y], nout = 2 */
#define PSM_MODEL_SIN 3               /* elementwise, nin = nout */
#define PSM_MODEL_COS 4
#define PSM_MODEL_EXP 5
#define PSM_MODEL_ARCTAN 6
/* Values evaluated by the caller (any function the caller can evaluate,
 * e.g. a torch callable), one row per linearization point -- per state for
 * psm_filter / psm_smoother / psm_filter_smoother (the transition of step
 * k -> k+1 and the observation of state k both use row k), per point for
 * psm_linearize.  Row layout: with a sigma table, f at the points
 * m + L xi_p (npts x nout, p in table order, L the chol of the nominal);
 * without one, f(m) followed by its Jacobian (nout + nout x nin). */
#define PSM_MODEL_TABULATED 7
/* Conditional-moments model (E[y|x], chol or covariance of y|x) evaluated by
 * the caller, as PSM_MODEL_TABULATED with each value block followed by the
 * conditional factor (nout x nout): with a sigma table npts x (nout + nout^2),
 * without one nout + nout x nin + nout^2 (mean, Jacobian, factor at the
 * mean).  q_mean / q_chol must point to zeros (no additive noise). */
#define PSM_MODEL_CONDITIONAL 8

/* status codes */
#define PSM_OK 0
#define PSM_E_ARG (-1)
#define PSM_E_UNSUPPORTED (-2)
#define PSM_E_HIP (-3)
#define PSM_E_WORKSPACE (-4)

/* FunctionalModel(function, MVN(q_mean, q_chol_or_cov)); all pointers on device,
 * in the problem's dtype. */
typedef struct psm_model {
  int32_t kind;
  int32_t nin;
  int32_t nout;
  int32_t pad_;
  const void* params;
  const void* q_mean;   /* (nout,) */
  const void* q_chol;   /* (nout, nout): chol in sqrt form, covariance in covariance form */
} psm_model_t;

/* Unit sigma-point table: points are m + L xi_p.  A NULL table (or npts == 0)
 * selects the extended (Taylor) linearization.  The table is FP64 (double)
 * for every dtype: the regression subtracts F P F^T from the sigma-point
 * covariance, and FP32-rounded weights (which no longer sum to one) leave a
 * residual larger than a small process noise. */
typedef struct psm_sigma {
  int32_t npts;
  int32_t pad_;
  const double* xi;  /* (npts, n) */
  const double* wm;  /* (npts,) */
  const double* wc;  /* (npts,) */
} psm_sigma_t;

#define PSM_SCAN_K_MASK 0xff
#define PSM_SCAN_TEAM 0x100
#define PSM_SCAN_HALF 0x200
#define PSM_SCAN_QUAD 0x400
#define PSM_SCAN_TOP_SHIFT 16
#define PSM_SCAN_TOP_MASK 0xff0000

/* A ragged batch of sequences sharing dims, models and linearization. */
typedef struct psm_problem {
  int32_t dtype;
  int32_t form;
  int32_t nx;
  int32_t ny;
  int32_t n_seq;
  int32_t chunk;         /* states per lane in the chunk passes (0: default) */
  int64_t n_states;      /* total states = sum(T_s + 1) */
  const int64_t* state_offsets;  /* device, n_seq + 1 */
  /* FP32 square-root form only: when a triangular factor produced by a Kalman
   * step, an absorbed step, a smoother element / application or a scan
   * combination has min|diag| / max|diag| below fallback_ratio, that operation
   * is recomputed in FP64 from the same FP32 inputs and rounded back.  0
   * disables the check.  fallback_count (device, may be NULL) is incremented
   * once per recomputation; the caller zeroes it. */
  float fallback_ratio;
  /* bits 0-7: scan items folded per lane (0: the PSM_SCAN_K environment
   * variable, else 2 with team-cooperative upper levels and 4 without);
   * PSM_SCAN_TEAM: levels above the first run team-cooperative kernels
   * (square-root form, float / double); with bits 0-7 clear they do so by
   * default, unless the PSM_SCAN_TEAM environment variable is 0;
   * PSM_SCAN_HALF: level 0 of the scan runs 32-lane workgroups (square-root
   * form, float / double): twice the workgroups, half the LDS each, one
   * scan-over-lanes step less;
   * PSM_SCAN_QUAD: level 0 of the scan runs quad-cooperative kernels (four
   * lanes per combination, 64 or, with PSM_SCAN_HALF, 32 quads per
   * workgroup; square-root form, float / double): the level layout is the
   * same, each combination's chain of dependent operations about a third;
   * bits 16-23 (PSM_SCAN_TOP_MASK): a level whose items fit one workgroup
   * with at most this many items per lane is the top level (values below the
   * fan-in, 0 included, mean the fan-in) */
  int32_t scan_k;
  unsigned long long* fallback_count;
} psm_problem_t;

int psm_version(void);
const char* psm_last_error(void);
/* 1 if (dtype, form, nx, ny) has compiled kernels: nx 1..8 and ny 1..4 for
 * PSM_F32 / PSM_F64 in both forms; the tangent dtypes only for the square-root
 * slices listed in the tangent registry (nx <= 5).  The reference takes any
 * size from the array shapes (parsmooth/parallel/_operators.py:43-77). */
int psm_supported(int32_t dtype, int32_t form, int32_t nx, int32_t ny);

/* ---------------------------------------------------------------------------
 * Linearization of a FunctionalModel at n points.
 * Replaces parsmooth/linearization/_extended.py:9-40 (sigma == NULL) and
 * _cubature.py:10-34 / _gh.py:13-41 / _unscented.py:10-39 through
 * _sigma_points.py:63-100 (sigma != NULL, table built by the caller).
 * x_chol is the chol (sqrt form) or covariance (covariance form) of x.
 * Outputs F (n, nout, nin), C (n, nout, nout), b (n, nout).
 * ------------------------------------------------------------------------- */
int psm_linearize(int32_t dtype, int32_t form, const psm_model_t* model, const psm_sigma_t* sigma,
                  int64_t n, const void* x_mean, const void* x_chol, void* F, void* C, void* b,
                  void* stream);

/* ---------------------------------------------------------------------------
 * Filter elements for n steps from the linearized model and a prior (m, L)
 * per step.  Replaces parsmooth/parallel/_filtering.py:115-146
 * (_sqrt_associative_params_one) and :78-97 (_standard_associative_params_one).
 * Inputs: F (n,nx,nx) cq (n,nx,nx) bt (n,nx) H (n,ny,nx) cr (n,ny,ny) c (n,ny)
 * m (n,nx) L (n,nx,nx) y (n,ny).  Outputs: A b U eta Z (U/Z or C/J in
 * covariance form).
 * ------------------------------------------------------------------------- */
int psm_filter_elements(int32_t dtype, int32_t form, int32_t nx, int32_t ny, int64_t n,
                        const void* F, const void* cq, const void* bt, const void* H,
                        const void* cr, const void* c, const void* m, const void* L,
                        const void* y, void* A, void* b, void* U, void* eta, void* Z,
                        void* stream);

/* Batched filtering operator out = e1 (+) e2 (e1 earlier).  Replaces
 * parsmooth/parallel/_operators.py:43-77 (sqrt) and :7-40 (covariance).
 * e1, e2, out: arrays of 5 device pointers (A, b, U|C, eta, Z|J). */
int psm_filter_combine(int32_t dtype, int32_t form, int32_t nx, int64_t n, const void* const* e1,
                       const void* const* e2, void* const* out, void* stream);

/* Smoother elements for n states from the linearized transition and the
 * filtered (m, L).  Replaces parsmooth/parallel/_smoothing.py:72-85 and :60-69.
 * Outputs g (n,nx), E (n,nx,nx), D (n,nx,nx) (D: chol or covariance). */
int psm_smoother_elements(int32_t dtype, int32_t form, int32_t nx, int64_t n, const void* F,
                          const void* cq, const void* bt, const void* m, const void* L, void* g,
                          void* E, void* D, void* stream);

/* Batched smoothing operator (e1 = later aggregate, e2 = earlier).  Replaces
 * parsmooth/parallel/_operators.py:104-125 and :80-101.  3 pointers each. */
int psm_smoother_combine(int32_t dtype, int32_t form, int32_t nx, int64_t n, const void* const* e1,
                         const void* const* e2, void* const* out, void* stream);

/* ---------------------------------------------------------------------------
 * Whole passes over a ragged batch.
 * ------------------------------------------------------------------------- */
size_t psm_workspace_bytes(const psm_problem_t* prob);

/* States per lane of the chunk passes this problem resolves to (prob->chunk,
 * else the PSM_CHUNK environment variable, else a size-dependent default). */
int psm_chunk_size(const psm_problem_t* prob);

/* Parallel filter at a nominal trajectory.  Replaces
 * parsmooth/parallel/_filtering.py:13-61 (linearization at every step, element
 * construction, associative scan, x0 prepended, log-likelihood).
 * ys (n_obs,ny); m0 (n_seq,nx); L0 (n_seq,nx,nx); nominal (n_states,...);
 * out (n_states,...).  ell (n_seq doubles, device) may be NULL. */
int psm_filter(const psm_problem_t* prob, const psm_model_t* trans, const psm_model_t* obs,
               const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
               const void* nom_m, const void* nom_L, void* out_m, void* out_L, double* ell,
               void* workspace, size_t workspace_bytes, void* stream);

/* Parallel smoother at a nominal trajectory.  Replaces
 * parsmooth/parallel/_smoothing.py:14-57 (elements, last element (m_T, 0, L_T),
 * reverse associative scan). */
int psm_smoother(const psm_problem_t* prob, const psm_model_t* trans, const psm_sigma_t* sigma,
                 const void* filt_m, const void* filt_L, const void* nom_m, const void* nom_L,
                 void* out_m, void* out_L, void* workspace, size_t workspace_bytes, void* stream);

/* Sequential filter (parallel=False): one lane per sequence runs predict /
 * update through its steps.  Replaces parsmooth/sequential/_filtering.py:11-53
 * (:56-108 for the two forms' predict and update; log-likelihood summed per
 * sequence in FP64 into ell, which may be NULL).  nom_m == NULL: the
 * transition is linearized at the filtered state and the observation at the
 * predicted state (:36-41); else both at the nominal trajectory.  Models of
 * kind PSM_MODEL_TABULATED / PSM_MODEL_CONDITIONAL need a nominal.  PSM_F32
 * and PSM_F64 only; no workspace. */
int psm_seq_filter(const psm_problem_t* prob, const psm_model_t* trans, const psm_model_t* obs,
                   const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                   const void* nom_m, const void* nom_L, void* out_m, void* out_L, double* ell,
                   void* stream);

/* Sequential (Rauch-Tung-Striebel) smoother, one lane per sequence.  Replaces
 * parsmooth/sequential/_smoothing.py:12-77.  nom_m == NULL: the transition out
 * of state k is linearized at the smoothed state k+1 (:27-33). */
int psm_seq_smoother(const psm_problem_t* prob, const psm_model_t* trans, const psm_sigma_t* sigma,
                     const void* filt_m, const void* filt_L, const void* nom_m, const void* nom_L,
                     void* out_m, void* out_L, void* stream);

/* One iteration of the fixed-point map: filter then smoother at the same
 * nominal (parsmooth/methods.py:38-47), sharing the transition linearization.
 * filt_m/filt_L receive the filtered trajectory.  use_graph != 0 captures the
 * launch sequence of this argument set in a hipGraph on first use and replays
 * it afterwards. */
int psm_filter_smoother(const psm_problem_t* prob, const psm_model_t* trans,
                        const psm_model_t* obs, const psm_sigma_t* sigma, const void* ys,
                        const void* m0, const void* L0, const void* nom_m, const void* nom_L,
                        void* filt_m, void* filt_L, void* out_m, void* out_L, void* workspace,
                        size_t workspace_bytes, int32_t use_graph, void* stream);

/* Size of the library-owned workspace of psm_filter_smoother_noalloc (per GPU). */
#define PSM_STATIC_WORKSPACE_BYTES (64u << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): psm_filter_smoother without a caller-supplied workspace
 * and without any allocation (no hipMalloc / hipFree, no event or graph
 * creation), eager on the caller's stream.  The scan workspace is a static
 * device buffer of PSM_STATIC_WORKSPACE_BYTES that each GPU holds until the
 * process ends; problems whose psm_workspace_bytes exceeds it are rejected
 * with PSM_E_WORKSPACE (64 MiB holds a single FP64 nx=5 sequence of about
 * 41,000 steps).  Every call on one GPU shares that buffer: calls that
 * overlap on one GPU (concurrent streams or threads) race and must be
 * serialised by the caller.  Nothing in it needs to start at zero. */
int psm_filter_smoother_noalloc(const psm_problem_t* prob, const psm_model_t* trans,
                                const psm_model_t* obs, const psm_sigma_t* sigma, const void* ys,
                                const void* m0, const void* L0, const void* nom_m,
                                const void* nom_L, void* filt_m, void* filt_L, void* out_m,
                                void* out_L, void* stream);

/* The whole fixed-count iteration of iterated smoothing as one launch
 * sequence: n_iter filter + smoother passes, pass k at the previous pass's
 * smoothed trajectory (pass 0 at init), then, when ell != NULL, the
 * log-likelihood filter pass at the final nominal (ell: n_seq doubles; the
 * filtered trajectory of that pass is left in filt_m / filt_L, which nothing
 * else writes).  Between passes the filtered and smoothed trajectories stay in
 * the workspace, state-interleaved.  Replaces
 * parsmooth/methods.py:54-76 with criterion i < n_iter (the fixed_point
 * while_loop of parsmooth/_utils.py:136-146 run on the device).  The passes
 * write out0 and out1 alternately: the result is in out0 when n_iter is odd,
 * in out1 when it is even (out1 may be NULL for n_iter == 1); out0 must not
 * alias init.  use_graph != 0 captures the whole sequence in one hipGraph on
 * first use of this argument set and replays it afterwards (no host work
 * between passes). */
int psm_iterated_smoother_graph(const psm_problem_t* prob, const psm_model_t* trans,
                                const psm_model_t* obs, const psm_sigma_t* sigma, const void* ys,
                                const void* m0, const void* L0, const void* init_m,
                                const void* init_L, int32_t n_iter, void* filt_m, void* filt_L,
                                void* out0_m, void* out0_L, void* out1_m, void* out1_L,
                                double* ell, void* workspace, size_t workspace_bytes,
                                int32_t use_graph, void* stream);

/* ---------------------------------------------------------------------------
 * Reverse-mode adjoint (vector-Jacobian products) of one filter + smoother
 * pass at a nominal trajectory and of its log-likelihood, for the implicit
 * gradient of iterated smoothing.  Replaces the custom VJP of the fixed point,
 * parsmooth/_utils.py:118-133 (_fixed_point_rev / _rev_iter: vjp of
 * filter_smoother with respect to the nominal and the model parameters) and
 * jax.grad of the log-likelihood of parsmooth/parallel/_filtering.py:53-60.
 * Square-root form, PSM_F64 problems (all adjoint arithmetic and every
 * cotangent array is FP64; other dtypes return PSM_E_UNSUPPORTED and a
 * workspace size of 0), models the kernels evaluate.  An FP32 objective runs
 * these on an FP64 problem holding its FP32-rounded observations and model
 * arrays, as parsmooth.gradient does: the checkpoint of an FP32 primal would
 * carry FP32 rounding of the smoothed means into every step's residual.  Cotangents of
 * chol factors are lower triangular (the factors' lower triangles are the
 * variables).
 *
 * Parameter cotangents theta_bar: (n_seq, PSM_ADJ_NTH(nx, ny)) doubles per
 * call, per sequence, laid out as [transition params (nx*nx slot; the model
 * kind's parameter count is a prefix: linear nx*nx, coordinated turn 1) |
 * transition noise mean (nx) | transition noise chol (nx*nx) | observation
 * params (ny*nx slot: linear ny*nx, bearings 4) | observation noise mean (ny)
 * | observation noise chol (ny*ny) | prior mean m0 (nx) | prior chol L0
 * (nx*nx)].  theta_bar may be NULL.
 *
 * Workspace: caller-owned, psm_adjoint_workspace_bytes; psm_adjoint_prepare
 * writes the checkpoint there (the primal pass's filtered and smoothed
 * covariances, gains, filter-element matrices and linearized models at the
 * nominal, FP64), which the VJP calls read, so it must stay intact until the
 * last VJP of that nominal.
 * ------------------------------------------------------------------------- */
#define PSM_ADJ_NTH(nx, ny) (3 * (nx) * (nx) + 2 * (nx) + (ny) * (nx) + (ny) + (ny) * (ny))

size_t psm_adjoint_workspace_bytes(const psm_problem_t* prob);

/* The primal pass at (nom_m, nom_L) -- filter then smoother, as
 * psm_filter_smoother, writing filt_* and out_* -- followed by the adjoint
 * checkpoint of that pass in aws.  ws / wsb: the primal's workspace. */
int psm_adjoint_prepare(const psm_problem_t* prob, const psm_model_t* trans, const psm_model_t* obs,
                        const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                        const void* nom_m, const void* nom_L, void* filt_m, void* filt_L, void* out_m,
                        void* out_L, void* workspace, size_t workspace_bytes, void* aws, size_t aws_bytes,
                        void* stream);

/* J^T u of the pass prepared at (nom_m, nom_L): for cotangents (cot_m, cot_L)
 * of its smoothed trajectory (sm_m, sm_L, as psm_adjoint_prepare returned
 * it), the cotangents of the nominal trajectory (out_m (n_states, nx), out_L
 * (n_states, nx, nx)), plus (add_m, add_L) when those are given (both or
 * neither), and of the model parameters (theta_bar).  One call evaluates both
 * products whatever the number of parameters: the reverse sum
 * u <- x_bar + J_x^T u of the fixed point is add = x_bar. */
int psm_filter_smoother_vjp(const psm_problem_t* prob, const psm_model_t* trans, const psm_model_t* obs,
                            const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                            const void* nom_m, const void* nom_L, const void* sm_m, const void* sm_L,
                            const double* cot_m, const double* cot_L, const double* add_m,
                            const double* add_L, double* out_m, double* out_L, double* theta_bar,
                            void* aws, size_t aws_bytes, void* stream);

/* Gradient of the per-sequence log-likelihood of the filter at (nom_m,
 * nom_L) (the ell of psm_filter) with respect to the nominal trajectory
 * (out_m, out_L) and the model parameters (theta_bar), by Fisher's identity
 * over the smoothed distribution of the linearized model (sm_m, sm_L from
 * psm_adjoint_prepare). */
int psm_filter_vjp(const psm_problem_t* prob, const psm_model_t* trans, const psm_model_t* obs,
                   const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                   const void* nom_m, const void* nom_L, const void* sm_m, const void* sm_L, double* out_m,
                   double* out_L, double* theta_bar, void* aws, size_t aws_bytes, void* stream);

/* Pathwise samples of the smoothing distribution at a nominal trajectory.
 * Replaces parsmooth/_pathwise_sampler.py:13-38 (parallel sampling, reached
 * from parsmooth/methods.py:79-91) with its elements :64-127.  From the
 * filtered trajectory (filt_m, filt_L) and the transition linearized at the
 * nominal, every sample path follows x_k = E_k x_{k+1} + g_k + D_k eps_k
 * backwards from x_T = m_T + L_T eps_T (smoother elements; D_k's Cholesky
 * factor in covariance form).  eps: standard normals drawn by the caller,
 * (n_states, n_samples, nx); out: samples, same shape.  PSM_F32 / PSM_F64. */
size_t psm_sample_workspace_bytes(const psm_problem_t* prob, int32_t n_samples);
int psm_sample(const psm_problem_t* prob, const psm_model_t* trans, const psm_sigma_t* sigma,
               const void* filt_m, const void* filt_L, const void* nom_m, const void* nom_L,
               const void* eps, int32_t n_samples, void* out, void* workspace,
               size_t workspace_bytes, void* stream);

/* mean((a - b)^2) over n values, accumulated in FP64 into *out (device).
 * The default stopping criterion of parsmooth/methods.py:50-51. */
int psm_mean_sq_diff(int32_t dtype, int64_t n, const void* a, const void* b, double* out,
                     void* stream);

/* Runtime-size helpers of parsmooth._utils (parsmooth/_utils.py), one lane per
 * batch item, each item row-major and contiguous in device memory; PSM_F32 /
 * PSM_F64.  They serve the Python utilities; the passes keep their blocks in
 * registers at the compiled sizes.
 * psm_tria replaces tria (_utils.py:22-24): A (nb x rows x cols) is
 *   triangularised in place; its first min(rows, cols) columns become L (lower,
 *   non-negative diagonal; the reference keeps QR's signs) with L L^T = A A^T,
 *   the other columns zero.
 * psm_cholesky_update replaces _cholesky_update and cholesky_update_many
 *   (_utils.py:13-20, 39-81): L (nb x n x n, lower, in place) becomes the factor
 *   of L L^T + multiplier v v^T for each of the nv vectors V (nb x nv x n) in
 *   turn; after each vector the entries above the diagonal are dropped and
 *   non-finite entries set to 0, as in the reference.  work: nb x n values of
 *   the dtype.
 * psm_mvn_loglikelihood replaces mvn_loglikelihood (_utils.py:152-160):
 *   out[b] = log N(x_b; 0, L_b L_b^T) for x (nb x n) and lower L (nb x n x n),
 *   in FP64.  work: nb x n doubles. */
int psm_tria(int32_t dtype, int64_t nb, int32_t rows, int32_t cols, void* A, void* stream);
int psm_cholesky_update(int32_t dtype, int64_t nb, int32_t n, int32_t nv, void* L, const void* V,
                        double multiplier, void* work, void* stream);
int psm_mvn_loglikelihood(int32_t dtype, int64_t nb, int32_t n, const void* x, const void* L,
                          double* out, double* work, void* stream);

/* Diagnostic, no reference counterpart (the reference divides and takes
 * square roots through XLA): out_rcp[i] = 1 / x[i] and out_sqrt[i] = sqrt(x[i])
 * for n FP64 device values, computed as every FP64 factorisation of this
 * library computes them (hardware estimates refined in software, for normal
 * finite arguments, and sqrt(0) = 0), so that tests can compare them with
 * correctly rounded results. */
int psm_math_check(int64_t n, const double* x, double* out_rcp, double* out_sqrt, void* stream);

/* Schedule autotuner (no reference counterpart: the reference's schedule is
 * XLA's).  Times psm_filter_smoother (eager, on `stream`) for candidate
 * schedules whose workspace fits workspace_bytes: first chunk in {2, 4, 8, 16,
 * 32} states per lane x scan fan-in scan_k in {2, 4, 8}, with and without
 * PSM_SCAN_TEAM and PSM_SCAN_HALF, and with PSM_SCAN_TEAM | PSM_SCAN_QUAD
 * with and without PSM_SCAN_HALF (square-root form in float / double); then
 * the chunks at 5/8 .. 7/4 of the best chunk (those that are not powers of
 * two, at most 64) with the three best scan_k values.  Returns the fastest in
 * *best_chunk / *best_scan_k (and its time per pass in *best_ms, may be
 * NULL).  prob->chunk / prob->scan_k are ignored.  The outputs
 * (out_m, out_L, filt_m, filt_L) are overwritten; FP32 fallback counts
 * accumulate.  Synchronizes the stream.  Results are cached per (dtype, form,
 * nx, ny, n_seq, n_states, linearization) behind a mutex: a shape tuned before
 * returns without running. */
int psm_autotune(const psm_problem_t* prob, const psm_model_t* trans, const psm_model_t* obs,
                 const psm_sigma_t* sigma, const void* ys, const void* m0, const void* L0,
                 const void* nom_m, const void* nom_L, void* filt_m, void* filt_L, void* out_m,
                 void* out_L, void* workspace, size_t workspace_bytes, void* stream,
                 int32_t* best_chunk, int32_t* best_scan_k, float* best_ms);

/* Drop every cached hipGraph (call before freeing buffers they reference).
 * The cache keeps at most 64 graphs and drops the least recently used one
 * (after its last launch has finished) when a new one is captured.  Graphs of
 * a caller on a stream of its own are captured on and replayed on that
 * stream, so calls on different streams stay independent; callers on the
 * default (legacy / per-thread) stream share one side stream. */
void psm_clear_graphs(void);

/* Kernel timing for measurement: when enabled, each kernel launched by
 * psm_filter / psm_smoother / psm_filter_smoother (outside graph capture) is
 * bracketed by hipEvents on its stream.  psm_timing_report() waits for them
 * and writes {"kernel": [total_ms, launches], ...} into buf (returns the full
 * length), then forgets the records. */
void psm_timing_enable(int32_t enable);
int psm_timing_report(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* PSM_H_ */
