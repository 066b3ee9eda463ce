"""Benchmark of the parallel square-root iterated smoother on one MI355X.

Workload (one "step" = one iterated_smoothing call): bearings-only tracking
with the coordinated-turn model (nx=5, ny=2, dt=0.01, qc=0.01, qw=0.1, two
bearing sensors, noise std 0.5), extended linearization (IEKS), 10 iterations
from the nominal [-1,-1,6,4,2]/I as in the reference's runtime notebook, one
sequence of T=2^17 observations, FP64.  Inputs are synthetic (seeded
simulator) and resident in HBM before timing.

Printed value: smoothed step-iterations per second = T x n_iter x n_seq / time.
The same workload in FP32 is timed alongside (fp32_value), and the FP64 CPU
restatement of the algorithm is timed on a bounded sample (cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sqrt-parallel-smoothers_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

NX, NY = 5, 2
# algorithmic figures per step-iteration (one state through one filter + smoother pass), nx=5 ny=2
ALG_BYTES_WORDS = NY + 4 * (NX + NX * (NX + 1) // 2)     # 82 words
ALG_FLOPS = 3793                                       # work-optimal sqrt filter + smoother
# algorithmic work of the dominant kernels: flops per unit processed
# per state (phase kernels) or per chunk aggregate (scan kernels, 2 combines per item)
KERNEL_FLOPS = {
    "k_filter_reduce": ("state", 1600),     # one step absorbed into a chunk aggregate
    "k_filter_apply": ("state", 1110),      # one square-root Kalman predict + update
    "k_smoother_reduce": ("state", 2980),   # smoother element (2016) + aggregate combine (966)
    "k_smoother_apply": ("state", 966),     # one smoother application
    "k_linearize_states": ("state", 0),
    "k_fscan_up_l0": ("chunk", 2 * 5741),   # filtering operator, nx=5
    "k_sscan_up_l0": ("chunk", 2 * 966),    # smoothing operator, nx=5
}
# algorithmic HBM bytes of the same kernels per unit (FP64 words x 8): per state,
# or per chunk aggregate (86 words in + 86 out for the filter scan items, 56 + 56
# for the smoother's)
KERNEL_BYTES = {
    "k_filter_reduce": ("state", 8 * 73),
    "k_filter_apply": ("state", 8 * 103),
    "k_smoother_reduce": ("state", 8 * 141),
    "k_smoother_apply": ("state", 8 * 86),
    "k_linearize_states": ("state", 8 * 101),
    "k_fscan_up_l0": ("chunk", 8 * 172),
    "k_sscan_up_l0": ("chunk", 8 * 112),
}
PEAK_FP64 = 78.6e12
PEAK_FP32 = 157.3e12
PEAK_HBM = 8.0e12
# roofline of one step-iteration (nx=5) in step-it/s: HBM-bound for the
# extended FP64 pass, VALU-bound for the FP32 cubature pass
CEIL_EXT_FP64 = PEAK_HBM / (8 * ALG_BYTES_WORDS)              # 12.2 G
CEIL_CUB_FP32 = PEAK_FP32 / (ALG_FLOPS + 4514)               # 18.9 G
FP32_BAR = 1e-4   # smoothed-mean relative error of the FP32 path the metric requires


# rocprofv3 names of the kernels timed under a launch label (FP64, nx=5, ny=2);
# {wg}: lanes of the level-0 scan workgroup (32 with the half-wave option)
PMC_NAMES = {
    "k_fscan_up_l0": "psm::k_scan_up<psm::FAgg<double, 5>, double, true, {wg}>",
    "k_sscan_up_l0": "psm::k_scan_up<psm::SAgg<double, 5>, double, true, {wg}>",
    "k_fscan_up_l0_quad": "psm::k_scan_up_quad<psm::FAgg<double, 5>, double, {wg}>",
    "k_sscan_up_l0_quad": "psm::k_scan_up_quad<psm::SAgg<double, 5>, double, {wg}>",
    "k_filter_reduce": "psm::k_filter_reduce<double, 5, 2, true>",
    "k_filter_apply": "psm::k_filter_apply<double, 5, 2, true>",
    "k_smoother_reduce": "psm::k_smoother_reduce<double, 5, 2, true>",
    "k_smoother_apply": "psm::k_smoother_apply<double, 5, true>",
}
SCHED_KEYS = ("chunk", "scan_k", "team_upper_levels", "half_wave_level0", "quad_level0")


def measured_traffic(label, sched):
    """HBM bytes per (largest) dispatch of the kernel, from the FETCH_SIZE /
    WRITE_SIZE passes summarised in profiles/ (scaled by a copy kernel of known
    size); None when no such summary is present or when it was taken with
    another launch schedule than `sched`."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_bearings_ieks_T131072.json")
    if sched.get("quad_level0") and label + "_quad" in PMC_NAMES:
        label += "_quad"
    name = PMC_NAMES.get(label, "").format(wg=32 if sched.get("half_wave_level0") else 64)
    try:
        with open(path) as fh:
            d = json.load(fh)
        if any(d["schedule"][key] != sched[key] for key in SCHED_KEYS):
            return None
        k = d["kernels"].get(name)
    except (OSError, ValueError, KeyError):
        return None
    return None if k is None else k["hbm_bytes_max_scaled"]


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def make_inputs(T, seed=0):
    from parsmooth.data import bearings_data
    _, _, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0),
                             random_state=seed)
    return ys


def build_runner(T, ys, dtype, n_seq=1, method=None, sqrt=True):
    """Bearings-only IEKS/IPLS runner; square-root (MVNSqrt) or covariance
    (MVNStandard, sqrt=False) form.  The initial nominal is the notebook's
    [-1, -1, 6, 4, 2] with identity chol / covariance at every state."""
    import torch
    from parsmooth import FunctionalModel, MVNSqrt, MVNStandard
    from parsmooth.batch import IteratedSmoother
    from parsmooth.linearization import extended
    from parsmooth.models import make_parameters
    dev = torch.device("cuda", 0)
    Q, R, obs_f, trans_f = make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64)).to(dev, dtype)
    cls = MVNSqrt if sqrt else MVNStandard
    trans = FunctionalModel(trans_f, cls(t(np.zeros(5)), t(np.linalg.cholesky(Q) if sqrt else Q)))
    obs = FunctionalModel(obs_f, cls(t(np.zeros(2)), t(np.linalg.cholesky(R) if sqrt else R)))
    runner = IteratedSmoother([T] * n_seq, trans, obs, method or extended, dtype=dtype, sqrt=sqrt)
    m0 = np.tile(np.array([0.1, 0.2, 1.0, 0.0, 0.0]), (n_seq, 1))
    runner.set_data(t(np.tile(ys, (n_seq, 1))), t(m0), t(np.tile(np.eye(5), (n_seq, 1, 1))))
    M = runner.prob.n_states
    init = (t(np.tile(np.array([-1.0, -1.0, 6.0, 4.0, 2.0]), (M, 1))), t(np.tile(np.eye(5), (M, 1, 1))))
    return runner, init


def build_batch_runner(n_seq, lo, hi, seed, dtype, method):
    """Ragged batch: lengths U{lo..hi}, per-sequence prior mean ~ N([0.1,0.2,1,0,0], 0.1 I),
    chol I; initial nominal = the prior mean at every state, identity chol."""
    import torch
    from parsmooth import FunctionalModel, MVNSqrt
    from parsmooth.batch import IteratedSmoother
    from parsmooth.data import bearings_data_batch
    from parsmooth.models import make_parameters
    rng = np.random.default_rng(seed)
    lengths = rng.integers(lo, hi + 1, n_seq)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0]) + np.sqrt(0.1) * rng.standard_normal((n_seq, 5))
    _, obs = bearings_data_batch(lengths, m0[:, :4], 0.01, 0.5, (-1.5, 0.5), (1.0, 1.0), seed=seed + 1)
    dev = torch.device("cuda", 0)
    Q, R, obs_f, trans_f = make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64)).to(dev, dtype)
    trans = FunctionalModel(trans_f, MVNSqrt(t(np.zeros(5)), t(np.linalg.cholesky(Q))))
    obsm = FunctionalModel(obs_f, MVNSqrt(t(np.zeros(2)), t(np.linalg.cholesky(R))))
    runner = IteratedSmoother(lengths, trans, obsm, method, dtype=dtype)
    runner.set_data(t(np.concatenate(obs)), t(m0), t(np.tile(np.eye(5), (n_seq, 1, 1))))
    reps = lengths + 1
    init = (t(np.repeat(m0, reps, axis=0)), t(np.tile(np.eye(5), (int(reps.sum()), 1, 1))))
    return runner, init, int(lengths.sum())


def build_grad_objective(n_seq, T, n_iter, dtype, seed=5):
    """Pseudo-log-likelihood of cubature IPLS for n_seq bearings-only sequences of
    length T simulated at qc=0.01, qw=0.1, sigma=0.5, and its gradient in
    (log qc, log qw, log sigma, s1x, s1y, s2x, s2y)."""
    import torch
    from parsmooth import _native as N
    from parsmooth.data import bearings_data_batch
    from parsmooth.gradient import IteratedLoglikelihood, bearings_parameterisation
    from parsmooth.linearization import cubature
    rng = np.random.default_rng(seed)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0]) + np.sqrt(0.1) * rng.standard_normal((n_seq, 5))
    lengths = np.full(n_seq, T)
    # turn rate simulated as a random walk with the model's own intensity qw = 0.1
    # (increments of variance qw dt), positions by the exact turn
    truth, obs = bearings_data_batch(lengths, m0[:, :4], 0.01, 0.5, (-1.5, 0.5), (1.0, 1.0),
                                     q=np.sqrt(0.1 / 0.01), seed=seed + 1)
    L0 = np.tile(np.eye(5), (n_seq, 1, 1))
    build = bearings_parameterisation(0.01, m0, L0)
    # IPLS started at the simulated trajectories (chol 0.1 I), so that the
    # fixed point the gradient differentiates is reached in n_iter iterations
    init = (np.concatenate(truth).astype(np.float64), np.tile(0.1 * np.eye(5), (n_seq * (T + 1), 1, 1)))
    obj = IteratedLoglikelihood(build, N.MODEL_COORDINATED_TURN, N.MODEL_BEARINGS, 5, 2, lengths,
                                np.concatenate(obs), cubature, n_iter, init, dtype=dtype)
    theta = torch.tensor([np.log(0.01), np.log(0.1), np.log(0.5), -1.5, 0.5, 1.0, 1.0], dtype=torch.float64)
    return obj, theta


def time_gradient(obj, theta, steps, warmup, mode="reverse"):
    import torch
    for _ in range(warmup):
        obj.value_and_grad(theta, mode=mode)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ell, grad = obj.value_and_grad(theta, mode=mode)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, ell, grad


def time_calls(fn, steps, warmup):
    """The runtime notebook's protocol: warm-up calls, then `steps` timed calls,
    each synchronised; returns (mean, median) seconds."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return float(np.mean(ts)), float(np.median(ts))


def t_sweep(exps, n_iter, steps, warmup):
    """Bearings-only IEKS (extended, n_iter iterations) over T = 2^e, FP64 and FP32."""
    import torch
    res = []
    for e in exps:
        T = 1 << e
        ys = make_inputs(T)
        row = {"T": T}
        for name, dt_ in (("fp64", torch.float64), ("fp32", torch.float32)):
            r, init = build_runner(T, ys, dt_)
            mean, med = time_calls(lambda: r.run(n_iter, init, use_graph=True), steps, warmup)
            row[name] = {"mean_ms": mean * 1e3, "median_ms": med * 1e3, "step_its_per_s": T * n_iter / mean}
            del r, init
        res.append(row)
    torch.cuda.empty_cache()
    return res


def linear_cv_case(steps, warmup):
    """Linear-Gaussian constant-velocity model (nx=4, ny=2, dt=0.1, q=1, R=0.5^2 I),
    T=1024, FP64: one parallel filter + smoother pass (latency)."""
    import torch
    from parsmooth import FunctionalModel, MVNSqrt
    from parsmooth.batch import IteratedSmoother
    from parsmooth.data import cv_model, lgssm_data
    from parsmooth.linearization import extended
    from parsmooth.models import LinearFunction
    T = 1024
    F, Q, H, R = cv_model(0.1, 1.0, 0.5)
    rng = np.random.RandomState(0)
    _, ys = lgssm_data(rng.randn(4), F, H, R, Q, np.zeros(4), np.zeros(2), T, random_state=rng)
    dev = torch.device("cuda", 0)
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
    trans = FunctionalModel(LinearFunction(F), MVNSqrt(t(np.zeros(4)), t(np.linalg.cholesky(Q))))
    obs = FunctionalModel(LinearFunction(H), MVNSqrt(t(np.zeros(2)), t(np.linalg.cholesky(R))))
    r = IteratedSmoother([T], trans, obs, extended, dtype=torch.float64)
    r.set_data(t(ys), t(np.zeros((1, 4))), t(np.eye(4)[None]))
    nom = r.prob.default_nominal()
    out = r.prob.new_traj()
    mean, med = time_calls(lambda: r.pass_once(nom, out, use_graph=True), steps, warmup)
    return {"workload": "linear CV model nx=4 ny=2, T=1024, FP64, one filter+smoother pass (hipGraph)",
            "mean_us": mean * 1e6, "median_us": med * 1e6}


# Prefix lengths of the reference's FP32 robustness protocol
# (notebooks/robustness_100runs.py:48).
REF_FP32_NAN_T = [20, 30, 40, 50, 80, 100, 200, 300, 400, 500, 1000, 1500, 2000, 2500, 3000, 3500, 4000,
                  4500, 5000, 5500, 6000, 6500, 7000, 7500]


# Reference-published GPU times at T=4096 (notebooks/new_outputs/rtx-GPU-all-methods.csv
# and notebooks/outputs/GPU-all-methods, line 11): mean wall time per call of
# iterated_smoothing, 10 iterations, FP64.
REF_T4096_MS = {
    "extended_sqrt": {"rtx_3080_ti": 301.2, "tesla": 298.3},
    "cubature_sqrt": {"rtx_3080_ti": 375.1, "tesla": 272.5},
    "extended_cov": {"rtx_3080_ti": 68.5, "tesla": 74.4},
    "cubature_cov": {"rtx_3080_ti": 86.1, "tesla": 64.7},
}


def published_cases(steps, warmup):
    """The reference's own GPU benchmark cases at T=4096 (bearings-only, 10
    iterations, FP64, nominal [-1, -1, 6, 4, 2] / I; the notebook's protocol is
    warm-up then timed calls with a device sync) -- square-root and covariance
    form, extended and cubature -- next to its published times.  Each case runs
    the schedule psm_autotune picks for it, as the headline does."""
    import torch
    from parsmooth.linearization import cubature, extended
    T = 1 << 12
    ys = make_inputs(T)
    res = {}
    for mname, method in (("extended", extended), ("cubature", cubature)):
        for form, sqrt in (("sqrt", True), ("cov", False)):
            r, init = build_runner(T, ys, torch.float64, method=method, sqrt=sqrt)
            tuned = r.autotune(init)
            mean, med = time_calls(lambda: r.run(10, init, use_graph=True), steps, warmup)
            key = f"{mname}_{form}"
            ref = REF_T4096_MS[key]
            res[key] = {"ours_ms": mean * 1e3, "ours_median_ms": med * 1e3, "reference_ms": ref,
                        "schedule": _sched(tuned, r),
                        "speedup_vs_rtx_3080_ti": ref["rtx_3080_ti"] / (mean * 1e3),
                        "speedup_vs_tesla": ref["tesla"] / (mean * 1e3)}
            del r, init
    torch.cuda.empty_cache()
    return res


def _ct_torch(dt):
    """Coordinated-turn transition in torch operations (a user-supplied model)."""
    import torch

    def f(x):
        w = x[..., 4]
        c, s = torch.cos(w * dt), torch.sin(w * dt)
        small = w.abs() < 1e-6
        a = torch.where(small, torch.full_like(w, dt), s / torch.where(small, torch.ones_like(w), w))
        bb = torch.where(small, torch.zeros_like(w), (c - 1) / torch.where(small, torch.ones_like(w), w))
        return torch.stack([x[..., 0] + a * x[..., 2] - bb * x[..., 3], x[..., 1] + bb * x[..., 2] + a * x[..., 3],
                            c * x[..., 2] + s * x[..., 3], -s * x[..., 2] + c * x[..., 3], w], -1)
    return f


def _bearings_torch(s1, s2):
    import torch

    def f(x):
        return torch.stack([torch.atan2(x[..., 1] - s1[1], x[..., 0] - s1[0]),
                            torch.atan2(x[..., 1] - s2[1], x[..., 0] - s2[0])], -1)
    return f


def widening_legs(steps, warmup):
    """The rows beyond the smoothing path, timed through their public API:
    the pathwise sampler, conditional-moments models, torch-callable model
    functions (against the same model built into the kernels), the sequential
    passes (parallel=False) and the L-BFGS-B parameter-estimation driver.  FP64."""
    import math
    import torch
    import parsmooth
    from parsmooth import ConditionalMomentsModel, FunctionalModel, MVNSqrt
    from parsmooth import _native as N
    from parsmooth.estimation import fit
    from parsmooth.gradient import IteratedLoglikelihood
    from parsmooth.linearization import cubature, extended
    from parsmooth.models import make_parameters
    dev = torch.device("cuda", 0)
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
    res = {}
    T = 1 << 14
    ys = make_inputs(T)
    Q, R, obs_f, trans_f = make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    cQ, cR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    x0 = MVNSqrt(t([0.1, 0.2, 1.0, 0.0, 0.0]), t(np.eye(5)))
    init = MVNSqrt(t(np.tile([-1.0, -1.0, 6.0, 4.0, 2.0], (T + 1, 1))), t(np.tile(np.eye(5), (T + 1, 1, 1))))
    crit = lambda i, *_: i < 10

    # built-in vs torch-callable models, extended IEKS, 10 iterations
    builtin = (FunctionalModel(trans_f, MVNSqrt(t(np.zeros(5)), t(cQ))),
               FunctionalModel(obs_f, MVNSqrt(t(np.zeros(2)), t(cR))))
    torchf = (FunctionalModel(_ct_torch(0.01), MVNSqrt(t(np.zeros(5)), t(cQ))),
              FunctionalModel(_bearings_torch((-1.5, 0.5), (1.0, 1.0)), MVNSqrt(t(np.zeros(2)), t(cR))))
    row = {"workload": f"bearings-only IEKS, T={T}, 10 iterations, FP64, parsmooth.iterated_smoothing"}
    for name, (tm, om) in (("builtin_ms", builtin), ("torch_callable_ms", torchf)):
        mean, _ = time_calls(lambda: parsmooth.iterated_smoothing(t(ys), x0, tm, om, extended, init,
                                                                  criterion=crit), steps, warmup)
        row[name] = mean * 1e3
    res["torch_callable_models"] = row

    # pathwise sampler: 64 samples of the smoothed bearings trajectory
    S = 64
    nom = parsmooth.iterated_smoothing(t(ys), x0, *builtin, extended, init, criterion=lambda i, *_: i < 3)
    filt = parsmooth.filtering(t(ys), x0, *builtin, extended, nom)
    mean, _ = time_calls(lambda: parsmooth.sampling(0, S, builtin[0], filt, extended, nom), steps, warmup)
    res["sampler"] = {"workload": f"bearings-only, T={T}, {S} samples, extended, FP64, parsmooth.sampling",
                      "ms": mean * 1e3, "sampled_states_per_s": (T + 1) * S / mean}

    # sequential passes (parallel=False, one lane per sequence; without a
    # nominal the extended Kalman filter + RTS smoother) against the parallel
    # filter-smoother pass: one long sequence, and a batch of short ones
    from parsmooth._engine import Models, Problem, filter_smoother_pass, seq_filter_pass, seq_smoother_pass
    from parsmooth.linearization._methods import resolve
    row = {"workload": f"bearings-only filter + smoother, extended, FP64: one sequence of T={T} "
                       "(parsmooth.filter_smoother), and 16384 sequences of T=64 in one batch"}
    mean, _ = time_calls(lambda: parsmooth.filter_smoother(t(ys), x0, *builtin, extended, None, parallel=False),
                         steps, warmup)
    row["one_sequence_sequential_ms"] = mean * 1e3
    mean, _ = time_calls(lambda: parsmooth.filter_smoother(t(ys), x0, *builtin, extended, init), steps, warmup)
    row["one_sequence_parallel_ms"] = mean * 1e3
    nb, Tb = 16384, 64
    prob = Problem(5, 2, [Tb] * nb, torch.float64, True, dev)
    mods = Models(prob, *builtin, resolve(extended))
    yb = t(np.tile(ys[:Tb], (nb, 1)))
    m0b = t(np.tile([0.1, 0.2, 1.0, 0.0, 0.0], (nb, 1)))
    L0b = t(np.tile(np.eye(5), (nb, 1, 1)))
    nomb = prob.default_nominal()
    fb_, ob_ = prob.new_traj(), prob.new_traj()

    def seq_batch():
        seq_filter_pass(prob, mods, yb, m0b, L0b, None, None, False, out=fb_)
        seq_smoother_pass(prob, mods, fb_[0], fb_[1], None, None, out=ob_)

    mean, _ = time_calls(seq_batch, steps, warmup)
    row["batch_sequential_steps_per_s"] = nb * Tb / mean
    mean, _ = time_calls(lambda: filter_smoother_pass(prob, mods, yb, m0b, L0b, nomb[0], nomb[1], fb_, ob_),
                         steps, warmup)
    row["batch_parallel_steps_per_s"] = nb * Tb / mean
    # the batch runner's autotuner makes the same choice by timing both at a nominal
    from parsmooth.batch import IteratedSmoother
    rb = IteratedSmoother([Tb] * nb, *builtin, extended)
    rb.set_data(yb, m0b, L0b)
    rb.autotune()
    row["batch_runner_autotune"] = rb.tuned
    res["sequential_passes"] = row
    del prob, mods, yb, nomb, fb_, ob_, rb

    # conditional-moments models: population model, cubature IPLS
    rng = np.random.RandomState(3)
    x = np.empty(T + 1)
    x[0] = 1.0
    for k in range(T):
        x[k + 1] = math.log(44.7) + x[k] - math.exp(x[k]) + 0.3 * rng.randn()
    yp = rng.poisson(10.0 * np.exp(x[1:]))[:, None].astype(np.float64)
    tm = ConditionalMomentsModel(lambda z: math.log(44.7) + z - torch.exp(z),
                                 lambda z: torch.full((1, 1), 0.3, dtype=z.dtype, device=z.device))
    om = ConditionalMomentsModel(lambda z: 10.0 * torch.exp(z), lambda z: torch.sqrt(10.0 * torch.exp(z)).reshape(1, 1))
    px0 = MVNSqrt(t([1.0]), t([[1.0]]))
    mean, _ = time_calls(lambda: parsmooth.iterated_smoothing(t(yp), px0, tm, om, cubature,
                                                              criterion=lambda i, *_: i < 10), steps, warmup)
    res["conditional_moments"] = {"workload": f"population model (Ricker + Poisson), T={T}, cubature IPLS, "
                                              "10 iterations, FP64", "ms": mean * 1e3,
                                  "step_its_per_s": T * 10 / mean}

    # parameter estimation: L-BFGS-B on (log q, log r) of a linear model
    Fm, Hm, Tp = np.array([[1.0, 0.1], [0.0, 1.0]]), np.array([[1.0, 0.0]]), 4096

    def build(theta):
        q, r = torch.exp(theta[0]), torch.exp(theta[1])
        return {"trans_params": torch.tensor(Fm.reshape(-1), dtype=theta.dtype),
                "trans_q_mean": torch.zeros(2, dtype=theta.dtype),
                "trans_q_chol": torch.sqrt(q) * torch.eye(2, dtype=theta.dtype),
                "obs_params": torch.tensor(Hm.reshape(-1), dtype=theta.dtype),
                "obs_q_mean": torch.zeros(1, dtype=theta.dtype), "obs_q_chol": r.reshape(1, 1),
                "m0": torch.zeros(1, 2, dtype=theta.dtype), "L0": torch.eye(2, dtype=theta.dtype)[None]}
    from parsmooth.data import lgssm_data
    _, yl = lgssm_data(np.zeros(2), Fm, Hm, np.array([[0.09]]), 0.05 * np.eye(2), np.zeros(2), np.zeros(1), Tp,
                       random_state=np.random.RandomState(0))
    obj = IteratedLoglikelihood(build, N.MODEL_LINEAR, N.MODEL_LINEAR, 2, 1, [Tp], np.asarray(yl, np.float64),
                                extended, 2, (np.zeros((Tp + 1, 2)), np.tile(np.eye(2), (Tp + 1, 1, 1))))
    fit(obj, np.array([0.0, 0.0]), bounds=[(-8.0, 3.0), (-8.0, 3.0)])   # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fit(obj, np.array([0.0, 0.0]), bounds=[(-8.0, 3.0), (-8.0, 3.0)])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res["parameter_estimation"] = {"workload": f"linear model nx=2 ny=1, T={Tp}, L-BFGS-B on (log q, log r), "
                                               "value and gradient on the device, FP64",
                                   "ms_per_fit": dt * 1e3, "evaluations": int(out.nfev),
                                   "converged": bool(out.success)}
    torch.cuda.empty_cache()
    return res


def fp32_stability(Ts, n_runs=100, n_iter=20):
    """The reference's FP32 robustness protocol (notebooks/robustness_100runs.py:
    100 bearings-only runs, each a prefix of length T of one T=8000 simulation;
    m0 = [-4, -1, 2, 7, 3], P0 = I; nominal [-1, -1, 6, 4, 2] with I at every
    state; criterion i < 20; log-likelihood at the final nominal) in FP32,
    square-root and covariance form, extended and cubature, on the reference's
    own 100 runs (tests/golden/robustness_runs.npz, truncated to the
    protocol's longest prefix).  All runs of one T go into one ragged batch;
    the log-likelihood accumulates in FP64.  Reports per T: the % of runs with a
    non-finite FP32 log-likelihood next to the reference's (its FP32 tables),
    the % of runs whose FP64 iteration (same kernels) diverges -- final smoothed
    mean beyond 1e3 in magnitude or not finite -- and, on the runs finite in
    both, the median relative difference of our FP32 log-likelihood from the
    reference's FP32 one and from our FP64 one."""
    import torch
    from parsmooth import FunctionalModel, MVNSqrt, MVNStandard
    from parsmooth.batch import IteratedSmoother
    from parsmooth.linearization import cubature, extended
    from parsmooth.models import make_parameters
    dev = torch.device("cuda", 0)
    fx = np.load(os.path.join(ROOT, "tests", "golden", "robustness_runs.npz"), allow_pickle=False)
    runs = fx["ys"][:n_runs]
    fixture_T = list(fx["Ts"])
    Q, R, obs_f, trans_f = make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    m0 = np.array([-4.0, -1.0, 2.0, 7.0, 3.0])
    res = {"protocol": "the reference's 100 runs x 20 iterations, FP32, % runs with non-finite log-likelihood",
           "T": list(Ts), "ours": {}, "reference": {}, "fp64_diverged": {}, "ell_vs_reference_fp32": {},
           "ell_vs_ours_fp64": {}}
    for method_name, method in (("extended", extended), ("cubature", cubature)):
        for form, sqrt in (("sqrt", True), ("std", False)):
            cls = MVNSqrt if sqrt else MVNStandard
            key = f"{method_name}_{form}"
            ours, div64, vref, v64 = [], [], [], []
            for T in Ts:
                ells = {}
                for dt_ in (torch.float32, torch.float64):
                    tt = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64)).to(dev, dt_)   # noqa: E731
                    trans = FunctionalModel(trans_f, cls(tt(np.zeros(5)), tt(np.linalg.cholesky(Q) if sqrt else Q)))
                    obsm = FunctionalModel(obs_f, cls(tt(np.zeros(2)), tt(np.linalg.cholesky(R) if sqrt else R)))
                    r = IteratedSmoother([T] * n_runs, trans, obsm, method, dtype=dt_, sqrt=sqrt)
                    r.set_data(tt(runs[:, :T].reshape(-1, 2)), tt(np.tile(m0, (n_runs, 1))),
                               tt(np.tile(np.eye(5), (n_runs, 1, 1))))
                    M = r.prob.n_states
                    init = (tt(np.tile(np.array([-1.0, -1.0, 6.0, 4.0, 2.0]), (M, 1))),
                            tt(np.tile(np.eye(5), (M, 1, 1))))
                    mean, _, ell = r.run(n_iter, init, use_graph=False, return_loglikelihood=True)
                    ells[dt_] = ell.double().cpu().numpy()
                    if dt_ == torch.float64:
                        big = [not bool(torch.isfinite(x).all()) or float(x.abs().max()) > 1e3 for x in r.split(mean)]
                        div64.append(float(100.0 * np.mean(big)))
                    del r, init
                e32, e64 = ells[torch.float32], ells[torch.float64]
                ours.append(float(100.0 * np.mean(~np.isfinite(e32))))
                ref_row = fx[f"ell_{key}"][fixture_T.index(T)].astype(np.float64) if T in fixture_T else None
                if ref_row is not None:
                    ok = np.isfinite(e32) & np.isfinite(ref_row)
                    vref.append(float(np.median(np.abs(e32[ok] - ref_row[ok]) / np.abs(ref_row[ok]))) if ok.any()
                                else None)
                else:
                    vref.append(None)
                ok = np.isfinite(e32) & np.isfinite(e64)
                v64.append(float(np.median(np.abs(e32[ok] - e64[ok]) / np.abs(e64[ok]))) if ok.any() else None)
            res["ours"][key] = ours
            res["fp64_diverged"][key] = div64
            res["ell_vs_reference_fp32"][key] = vref
            res["ell_vs_ours_fp64"][key] = v64
            res["reference"][key] = [float(100.0 * np.mean(np.isnan(fx[f"ell_{key}"][fixture_T.index(T)])))
                                     if T in fixture_T else None for T in Ts]
    torch.cuda.empty_cache()
    return res


def fp32_cubature_errors(exps, n_iter):
    """FP32 square-root cubature IPLS vs the FP64 run of the same kernels (which
    match the FP64 oracle to 1e-9 on the tested sizes) on bearings-only data,
    per T: one pass at a common nominal and after n_iter iterations; max over
    states of the mean error and of the L L^T error, relative to the largest
    magnitude.  FP64 fallback recomputations per pass are reported too."""
    import torch
    from parsmooth.linearization import cubature
    res = []
    for e in exps:
        T = 1 << e
        ys = make_inputs(T, seed=1)
        r64, init64 = build_runner(T, ys, torch.float64, method=cubature)
        r32, _ = build_runner(T, ys, torch.float32, method=cubature)
        # common nominal: the FP64 iterate after n_iter iterations
        m64, L64 = r64.run(n_iter, init64, use_graph=False)
        nom = (m64.clone(), L64.clone())
        a64 = r64.prob.new_traj()
        a32 = r32.prob.new_traj()
        r64.pass_once(nom, a64)
        r32.prob.fallback_counter.zero_()
        r32.pass_once((nom[0].float(), nom[1].float()), a32)
        fb = r32.fallback_count()
        it32 = r32.run(n_iter, (init64[0].float(), init64[1].float()), use_graph=False)
        torch.cuda.synchronize()

        def rel(a, b):
            a, b = a.double(), b.double()
            return float((a - b).abs().max() / max(1.0, float(b.abs().max())))

        def gram(L):
            L = L.double()
            return L @ L.transpose(-1, -2)
        res.append({"T": T, "pass_mean": rel(a32[0], a64[0]), "pass_cov": rel(gram(a32[1]), gram(a64[1])),
                    "iterated_mean": rel(it32[0], m64), "iterated_cov": rel(gram(it32[1]), gram(L64)),
                    "fallbacks_per_pass": fb})
        del r64, r32, init64, nom, a64, a32
        torch.cuda.empty_cache()
    return res


def perturbation_growth(runner, init, n_iter, seed=0):
    """How the FP64 iteration itself amplifies FP32-sized perturbations: after
    k = 1..n_iter iterations, the relative difference between the FP64 run from
    the nominal and the FP64 run from the nominal perturbed by 6e-8 relative
    noise (about one FP32 rounding).  An iteration that amplifies this past the
    FP32 bar cannot meet it in FP32 arithmetic whatever the kernels do."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    noise = torch.randn(init[0].shape, generator=g, dtype=torch.float64).to(init[0].device)
    pert = (init[0] * (1.0 + 6e-8 * noise), init[1])
    out = []
    for k in range(1, n_iter + 1):
        a = runner.run(k, init, use_graph=False)[0].clone()
        b = runner.run(k, pert, use_graph=False)[0]
        out.append(float((a - b).abs().max() / max(1.0, float(a.abs().max()))))
    return out


def smoothing_fp32_vs_fp64(n_seq, T, n_iter, steps, warmup, seed=5):
    """FP32 against FP64 square-root cubature IPLS on the gradient workload
    (n_seq coordinated-turn sequences of length T simulated at qc=0.01, qw=0.1,
    sigma=0.5, n_iter iterations from the simulated trajectories): smoothed
    step-iterations per second of each, and the FP32 error (smoothed means,
    relative to max(1, |x|) per sequence; covariances L L^T), over all
    sequences and over those whose FP64 iteration reached a fixed point
    (max |f(x*) - x*| < 1e-6)."""
    import torch
    from parsmooth import FunctionalModel, MVNSqrt
    from parsmooth.batch import IteratedSmoother
    from parsmooth.data import bearings_data_batch
    from parsmooth.linearization import cubature
    from parsmooth.models import make_parameters
    rng = np.random.default_rng(seed)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0]) + np.sqrt(0.1) * rng.standard_normal((n_seq, 5))
    lengths = np.full(n_seq, T)
    truth, obs = bearings_data_batch(lengths, m0[:, :4], 0.01, 0.5, (-1.5, 0.5), (1.0, 1.0),
                                     q=np.sqrt(0.1 / 0.01), seed=seed + 1)
    Q, R, obs_f, trans_f = make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    dev = torch.device("cuda", 0)
    res, outs = {}, {}
    for name, dt_ in (("fp64", torch.float64), ("fp32", torch.float32)):
        t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64)).to(dev, dt_)   # noqa: E731
        trans = FunctionalModel(trans_f, MVNSqrt(t(np.zeros(5)), t(np.linalg.cholesky(Q))))
        obsm = FunctionalModel(obs_f, MVNSqrt(t(np.zeros(2)), t(np.linalg.cholesky(R))))
        r = IteratedSmoother(lengths, trans, obsm, cubature, dtype=dt_)
        r.set_data(t(np.concatenate(obs)), t(m0), t(np.tile(np.eye(5), (n_seq, 1, 1))))
        init = (t(np.concatenate(truth)), t(np.tile(0.1 * np.eye(5), (r.prob.n_states, 1, 1))))
        r.autotune(init)
        dt, out, _ = time_runner(r, init, n_iter, steps, warmup, True)
        res[f"{name}_step_its_per_s"] = n_seq * T * n_iter / dt
        res[f"{name}_ms"] = dt * 1e3
        outs[name] = [x.double().cpu().numpy() for x in out]
        offs = r.prob.state_offsets_host
        if name == "fp64":
            nxt = r.prob.new_traj()
            r.pass_once((out[0].clone(), out[1].clone()), nxt)
            d = (nxt[0] - out[0]).abs().amax(1).cpu().numpy()
            conv = np.maximum.reduceat(d, offs[:-1]) < 1e-6
            res["sequences_at_fixed_point"] = int(conv.sum())
        del r, init
        torch.cuda.empty_cache()
    m64, L64 = outs["fp64"]
    m32, L32 = outs["fp32"]
    em, ec = [], []
    for s_ in range(n_seq):
        a = slice(offs[s_], offs[s_ + 1])
        em.append(np.abs(m32[a] - m64[a]).max() / max(1.0, np.abs(m64[a]).max()))
        g32 = L32[a] @ np.swapaxes(L32[a], 1, 2)
        g64 = L64[a] @ np.swapaxes(L64[a], 1, 2)
        ec.append(np.abs(g32 - g64).max() / max(1.0, np.abs(g64).max()))
    em, ec = np.array(em), np.array(ec)
    res["fp32_over_fp64"] = res["fp32_step_its_per_s"] / res["fp64_step_its_per_s"]
    res["fp32_rel_err_mean"] = {"all_max": float(em.max()), "all_median": float(np.median(em)),
                                "fixed_point_max": float(em[conv].max()) if conv.any() else None,
                                "fraction_within_bar": float(np.mean(em <= FP32_BAR))}
    res["fp32_rel_err_cov"] = {"all_max": float(ec.max()), "fixed_point_max": float(ec[conv].max()) if conv.any() else None}
    res["workload"] = (f"{n_seq} bearings-only sequences x T={T}, square-root cubature IPLS, {n_iter} iterations from "
                       "the simulated trajectories (chol 0.1 I), one whole-loop hipGraph per run, autotuned schedule")
    return res


def _sched(tuned, runner):
    """The launch schedule a runner uses: chunk, scan fan-in, team upper levels,
    half-wave level 0, quad-cooperative level 0 (scan_k 0: the library default,
    fan-in 2 with team levels)."""
    k = int(runner.prob.struct.scan_k)
    team = bool(k & 0x100) or (k & 0x1FF) == 0
    return {"chunk": runner.chunk_size(), "scan_k": (k & 0xFF) or (2 if team else 4), "team_upper_levels": team,
            "half_wave_level0": bool(k & 0x200), "quad_level0": bool(k & 0x400),
            "autotuned": tuned is not None, "tuned_ms_per_pass": None if tuned is None else tuned[2]}


def time_runner(runner, init, n_iter, steps, warmup, use_graph, kernel_timing=False):
    import torch
    from parsmooth import _native as N
    for _ in range(warmup):
        runner.run(n_iter, init, use_graph=use_graph)
    torch.cuda.synchronize()
    if kernel_timing:
        N.timing_report()
        N.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = runner.run(n_iter, init, use_graph=use_graph)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    rep = None
    if kernel_timing:
        N.timing_enable(False)
        rep = N.timing_report()
    return dt, out, rep


def cpu_info():
    """(usable cores, CPU model) of this host."""
    cores = len(os.sched_getaffinity(0))
    model = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return cores, model


def cpu_baseline_batch(n_seq, n_iter=5, seed=0, threads=0):
    """The C FP64 restatement of cubature IPLS on a bounded sample of the batch
    workload (the first n_seq sequences of the 4096-sequence batch), one OpenMP
    thread per sequence at a time across the host's cores."""
    from oracle import psm_oracle as O
    from oracle import psm_oracle_c as C
    from parsmooth.data import bearings_data_batch
    rng = np.random.default_rng(seed)
    lengths = rng.integers(256, 8193, 4096)[:n_seq]
    m0 = (np.array([0.1, 0.2, 1.0, 0.0, 0.0]) + np.sqrt(0.1) * rng.standard_normal((4096, 5)))[:n_seq]
    _, obs = bearings_data_batch(lengths, m0[:, :4], 0.01, 0.5, (-1.5, 0.5), (1.0, 1.0), seed=seed + 1)
    Q, R = O.bearings_parameters(0.01, 0.1, 0.5, 0.01)
    tab = O.cubature_table(5)
    reps = lengths + 1
    t0 = time.perf_counter()
    _, _, _, used = C.iterated_smoother_batch(lengths, np.concatenate(obs), m0, np.tile(np.eye(5), (n_seq, 1, 1)),
                                              np.repeat(m0, reps, axis=0), None, ("ct", 0.01),
                                              ("bearings", ((-1.5, 0.5), (1.0, 1.0))), np.zeros(5),
                                              np.linalg.cholesky(Q), np.zeros(2), np.linalg.cholesky(R), n_iter,
                                              sigma=(tab.xi, tab.wm, tab.wc), threads=threads)
    dt = time.perf_counter() - t0
    return int(lengths.sum()) * n_iter / dt, dt, used


def cpu_baseline(T_sample, ys, n_iter, repeats=1):
    """FP64 C restatement of the algorithm (sequential square-root filter and
    smoother at each nominal, one core) on the same inputs; falls back to the
    numpy restatement (parallel formulation) when the C library is not built."""
    from oracle import psm_oracle as O
    Q, R = O.bearings_parameters(0.01, 0.1, 0.5, 0.01)
    cQ, cR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    y = ys[:T_sample]
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
    init_m = np.tile(np.array([-1.0, -1.0, 6.0, 4.0, 2.0]), (T_sample + 1, 1))
    try:
        from oracle import psm_oracle_c as C
        C.lib()
        t0 = time.perf_counter()
        for _ in range(repeats):
            C.iterated_smoother(y, m0, np.eye(5), init_m, ("ct", 0.01), ("bearings", ((-1.5, 0.5), (1.0, 1.0))),
                                np.zeros(5), cQ, np.zeros(2), cR, n_iter)
        dt = time.perf_counter() - t0
        desc = f"C FP64 restatement (sequential sqrt filter/smoother, 1 thread), T={T_sample}, {n_iter} iterations x {repeats}"
    except RuntimeError:
        to = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNSqrt(np.zeros(5), cQ))
        oo = O.FunctionalModel(O.Bearings((-1.5, 0.5), (1.0, 1.0)), O.MVNSqrt(np.zeros(2), cR))
        init = O.MVNSqrt(init_m, np.tile(np.eye(5), (T_sample + 1, 1, 1)))
        t0 = time.perf_counter()
        for _ in range(repeats):
            O.iterated_smoothing(y, O.MVNSqrt(m0, np.eye(5)), to, oo, O.extended, init, True,
                                 criterion=lambda i, *_: i < n_iter)
        dt = time.perf_counter() - t0
        desc = f"numpy FP64 restatement (parallel formulation), T={T_sample}, {n_iter} iterations x {repeats}"
    return T_sample * n_iter * repeats / dt, dt, desc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--T", type=int, default=1 << 17)
    ap.add_argument("--n-iter", type=int, default=10)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--cpu-repeats", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-seqs", type=int, default=4096)
    ap.add_argument("--grad-seqs", type=int, default=256)
    ap.add_argument("--grad-T", type=int, default=16384)
    ap.add_argument("--grad-iter", type=int, default=10)
    ap.add_argument("--autotune", type=int, default=1,
                    help="pick the launch schedule with psm_autotune before timing (1) or keep the default (0)")
    ap.add_argument("--schedule", type=str, default="",
                    help="fixed schedule 'chunk,scan_k,team[,half[,quad]]' for the T=2^17 legs (implies --autotune 0)")
    ap.add_argument("--sweep", type=str, default="10,12,14,16",
                    help="comma-separated log2 T of the IEKS T-sweep ('' to skip)")
    ap.add_argument("--fp32-errors", type=str, default="14,16,18",
                    help="comma-separated log2 T of the FP32 cubature error report ('' to skip)")
    ap.add_argument("--stability", type=str, default="80,100,1000,7500",
                    help="T values of the FP32 robustness protocol ('all': the reference's 24; '': skip)")
    ap.add_argument("--smoothing-seqs", type=int, default=256,
                    help="sequences of the FP32-vs-FP64 cubature IPLS smoothing leg (0 to skip)")
    ap.add_argument("--cpu-batch-seqs", type=int, default=384,
                    help="sequences of the all-core CPU baseline sample of the batch workload (0 to skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device copy first (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the library has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.copy_calibration:
        # a plain device copy of known size, for scaling HBM byte counters
        nbytes = 256 << 20
        src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    T, n_iter = args.T, args.n_iter
    ys = make_inputs(T)

    # FP64 (headline), with live per-kernel HIP-event timing in a separate pass
    log(f"FP64 IEKS T={T} n_iter={n_iter}")
    fixed = [int(v) for v in args.schedule.split(",")] if args.schedule else None
    if fixed:
        args.autotune = 0

    def apply_fixed(r):
        if fixed:
            r.prob.set_schedule(fixed[0], fixed[1] | (0x100 if fixed[2] else 0)
                                | (0x200 if len(fixed) > 3 and fixed[3] else 0)
                                | (0x400 if len(fixed) > 4 and fixed[4] else 0))

    r64, init64 = build_runner(T, ys, torch.float64)
    apply_fixed(r64)
    sched64 = r64.autotune(init64) if args.autotune else None
    dt64, out64, _ = time_runner(r64, init64, n_iter, args.steps, args.warmup, bool(args.graph))
    dtk, _, rep = time_runner(r64, init64, n_iter, max(1, args.steps // 2), 0, False, kernel_timing=True)
    mean64 = out64[0].double().cpu().numpy()

    # FP32
    log("FP32 IEKS")
    r32, init32 = build_runner(T, ys, torch.float32)
    apply_fixed(r32)
    sched32 = r32.autotune(init32) if args.autotune else None
    r32.prob.fallback_counter.zero_()
    dt32, out32, _ = time_runner(r32, init32, n_iter, args.steps, args.warmup, bool(args.graph))
    fb32 = r32.fallback_count() / max(1, args.steps + args.warmup)
    mean32 = out32[0].double().cpu().numpy()
    sc = max(1.0, np.abs(mean64).max())
    rel32_end = float(np.abs(mean32 - mean64).max() / sc)
    # one filter + smoother pass of each precision at the same (FP64 final) nominal
    nom64 = (out64[0].clone(), out64[1].clone())
    p64 = r64.prob.new_traj()
    p32 = r32.prob.new_traj()
    r64.pass_once(nom64, p64)
    r32.pass_once((nom64[0].float(), nom64[1].float()), p32)
    torch.cuda.synchronize()
    d1 = np.abs(p32[0].double().cpu().numpy() - p64[0].cpu().numpy()).max(1) / sc
    rel32_pass = {"max": float(d1.max()), "median": float(np.median(d1)), "p99": float(np.quantile(d1, 0.99))}
    log("FP64 growth of FP32-sized perturbations")
    growth = perturbation_growth(r64, init64, n_iter)

    batch = None
    if args.batch_seqs > 0:
        from parsmooth.linearization import cubature
        log(f"batch leg: {args.batch_seqs} sequences")
        rb, ib, nsteps = build_batch_runner(args.batch_seqs, 256, 8192, 0, torch.float32, cubature)
        schedb = rb.autotune(ib) if args.autotune else None
        dtb, _, repb = time_runner(rb, ib, 5, max(1, args.steps // 2), 1, bool(args.graph))
        rb.prob.fallback_counter.zero_()
        _, _, repb = time_runner(rb, ib, 5, 1, 0, False, kernel_timing=True)
        fb_batch = rb.fallback_count() / 5.0
        batch = {"workload": f"{args.batch_seqs} bearings-only sequences, lengths U[256, 8192], FP32 sqrt "
                             "cubature IPLS, 5 iterations, one launch sequence per pass",
                 "steps": nsteps, "ms_per_step": dtb * 1e3, "value": nsteps * 5 / dtb,
                 "schedule": _sched(schedb, rb), "fp64_fallbacks_per_pass": fb_batch,
                 "unit": "step-it/s", "kernel_ms": {k: v[0] / v[1] for k, v in repb.items()}}
        del rb, ib
        torch.cuda.empty_cache()

    grad = None
    if args.grad_seqs > 0:
        legs = {}
        for name, dt_ in (("fp32", torch.float32), ("fp64", torch.float64)):
            log(f"gradient leg {name}: {args.grad_seqs} x T={args.grad_T}")
            obj, theta = build_grad_objective(args.grad_seqs, args.grad_T, args.grad_iter, dt_)
            for mode in ("reverse", "forward"):
                obj.prob.fallback_counter.zero_()
                obj.tprob.fallback_counter.zero_()
                tg, ell_g, g_g = time_gradient(obj, theta, 1, 1 if name == "fp32" else 0, mode)
                fbk = int(obj.prob.fallback_counter.item()) + int(obj.tprob.fallback_counter.item())
                legs[name, mode] = (tg, ell_g.double().cpu().numpy(), g_g.double().cpu().numpy(), fbk)
            if name == "fp64":
                conv = obj.fixed_point_residual(theta) < 1e-6
            del obj
            torch.cuda.empty_cache()
        nst = args.grad_seqs * args.grad_T
        ref = legs["fp64", "reverse"][2][conv].sum(0)
        modes = {}
        for mode, how in (("reverse", "N + 3 adjoint passes (FP64 adjoint arithmetic; FP32 objectives on "
                                      "an FP64 twin of the FP32-rounded problem)"),
                          ("forward", "P (N + 3) tangent passes in the objective's precision")):
            t32, t64 = legs["fp32", mode][0], legs["fp64", mode][0]
            g32, g64 = legs["fp32", mode][2][conv].sum(0), legs["fp64", mode][2][conv].sum(0)
            e32, e64 = legs["fp32", mode][1][conv].sum(), legs["fp64", mode][1][conv].sum()
            modes[mode] = {"how": how,
                           "fp32_grad_evals_per_s": args.grad_seqs / t32,
                           "fp64_grad_evals_per_s": args.grad_seqs / t64,
                           "fp32_ms": t32 * 1e3, "fp64_ms": t64 * 1e3, "fp32_over_fp64": t64 / t32,
                           "fp32_grad_rel_err_vs_fp64": float(np.abs(g32 - g64).max() / np.abs(g64).max()),
                           "fp64_grad_rel_err_vs_fp64_reverse": float(np.abs(g64 - ref).max() / np.abs(ref).max()),
                           "fp32_ell_rel_err_vs_fp64": float(abs(e32 - e64) / abs(e64)),
                           "fp32_fp64_fallbacks": legs["fp32", mode][3]}
        grad = {"workload": f"{args.grad_seqs} bearings-only sequences x T={args.grad_T}, cubature IPLS, "
                            f"n_iter={args.grad_iter} from the simulated trajectories, gradient of the summed "
                            "pseudo-log-likelihood in 7 parameters (implicit fixed-point rule)",
                "reverse": modes["reverse"], "forward": modes["forward"],
                "reverse_over_forward_fp64": legs["fp64", "forward"][0] / legs["fp64", "reverse"][0],
                "fp32_smoothed_step_its_per_s": nst * args.grad_iter / legs["fp32", "reverse"][0],
                "fp64_sequences_at_fixed_point": int(conv.sum()),
                "fixed_point_criterion": "max |f(x*) - x*| over the state means < 1e-6 after n_iter iterations",
                "errors_over": "the sequences at a fixed point (summed gradient, summed ell)",
                "grad_fp64": ref.tolist()}

    smoothing = None
    if args.smoothing_seqs > 0:
        log(f"smoothing leg: {args.smoothing_seqs} x T={args.grad_T}, FP32 vs FP64 cubature IPLS")
        smoothing = smoothing_fp32_vs_fp64(args.smoothing_seqs, args.grad_T, args.grad_iter, max(1, args.steps // 2), 1)

    sweep = lin = errs = stab = pub = wide = None
    if args.sweep:
        log(f"T sweep {args.sweep}")
        sweep = t_sweep([int(x) for x in args.sweep.split(",")], n_iter, args.steps, args.warmup)
        lin = linear_cv_case(max(args.steps, 20), args.warmup)
        log("published reference cases, T=4096")
        pub = published_cases(max(args.steps, 10), args.warmup)
        log("sampler, conditional moments, torch-callable models, sequential passes, parameter estimation")
        wide = widening_legs(args.steps, 1)
    if args.fp32_errors:
        log(f"FP32 cubature errors {args.fp32_errors}")
        errs = fp32_cubature_errors([int(x) for x in args.fp32_errors.split(",")], 5)
    if args.stability:
        log(f"FP32 stability protocol T={args.stability}")
        Ts = REF_FP32_NAN_T if args.stability == "all" else [int(x) for x in args.stability.split(",")]
        stab = fp32_stability(Ts)

    value = T * n_iter / dt64
    # dominant kernel by total time; its algorithmic bytes and flops per launch
    dom, (tot_ms, launches) = max(rep.items(), key=lambda kv: kv[1][0])
    per_launch_s = tot_ms / launches / 1e3
    kind, fl = KERNEL_FLOPS.get(dom, ("state", 0))
    _, by = KERNEL_BYTES.get(dom, ("state", 0))
    units = (T + 1) if kind == "state" else r64.prob.n_states // r64.chunk_size()
    achieved_gbs = by * units / per_launch_s / 1e9
    valu = fl * units / per_launch_s / 1e12
    roof = {"bound": "hbm", "achieved": achieved_gbs, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": achieved_gbs / (PEAK_HBM / 1e9), "traffic": measured_traffic(dom, _sched(sched64, r64)),
            "kernel": dom, "kernel_avg_us": per_launch_s * 1e6,
            "alg_bytes_per_launch": by * units, "units_per_launch": units,
            "valu_fp64": {"achieved_tflops": valu, "peak_tflops": PEAK_FP64 / 1e12, "frac": valu / (PEAK_FP64 / 1e12)},
            "whole_pass": {"step_its_per_s": value, "ceiling_step_its_per_s": CEIL_EXT_FP64,
                           "frac": value / CEIL_EXT_FP64,
                           "ceiling": "HBM: 82 FP64 words per step-iteration at 8 TB/s"},
            "note": "no MFMA: <=10x10 square-root blocks; the FP64 VALU fraction of the same kernel is in valu_fp64"}
    cores, cpu_model = cpu_info()
    cpu = None
    if not args.no_cpu:
        log("CPU baseline")
        v, dtc, desc = cpu_baseline(min(args.cpu_sample, T), ys, n_iter, args.cpu_repeats)
        cpu = {"value": v, "unit": "step-it/s", "cores": 1, "kind": "port",
               "sample": f"{desc}, {dtc:.1f} s", "host_cores": cores, "cpu_model": cpu_model,
               "label": "CPU restatement of the reference algorithm (not the reference's JAX)"}
        if args.cpu_batch_seqs > 0:
            threads = min(cores, 16)
            vb, dtb_c, used = cpu_baseline_batch(args.cpu_batch_seqs, threads=threads)
            cpu["all_cores_batch"] = {
                "value": vb, "unit": "step-it/s", "cores": used, "kind": "port",
                "sample": f"C FP64 restatement of square-root cubature IPLS (sequential filter/smoother per "
                          f"sequence, OpenMP across sequences), first {args.cpu_batch_seqs} sequences of the "
                          f"4096-sequence batch, 5 iterations, {dtb_c:.1f} s",
                "gpu_fp32_batch_value": batch["value"] if batch else None}
    fp32_ok = rel32_end <= FP32_BAR
    out = {
        "metric": "smoothed steps/s (FP64 sqrt IEKS, bearings-only CT, 10 iterations)",
        "value": value, "unit": "step-it/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt64 * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded bearings-only simulator)",
        "config": {"workload": "bearings-only coordinated turn, extended IEKS, n_iter=10, one sequence",
                   "T": T, "nx": NX, "ny": NY, "n_seq": 1, "n_iter": n_iter, "graph": bool(args.graph),
                   "schedule": _sched(sched64, r64)},
        "fp32_schedule": _sched(sched32, r32),
        # the FP32 rate counts only when its iterate meets the bar on this workload
        "fp32_value": T * n_iter / dt32 if fp32_ok else None,
        "fp32_bar_met": fp32_ok, "fp32_value_unqualified": T * n_iter / dt32,
        "fp32_ms_per_step": dt32 * 1e3, "fp32_over_fp64": dt64 / dt32,
        "fp32_rel_err_single_pass": rel32_pass, "fp32_rel_err_after_iterations": rel32_end,
        "fp64_growth_of_fp32_sized_perturbations": growth,
        "smoothing_fp32_vs_fp64": smoothing,
        "fp32_fp64_fallbacks_per_step": fb32,
        "alg_bytes_per_step_it": ALG_BYTES_WORDS * 8, "alg_flops_per_step_it": ALG_FLOPS,
        "kernel_ms": {k: v[0] / v[1] for k, v in rep.items()},
        "batch_fp32_cubature": batch,
        "gradient": grad,
        "t_sweep_ieks": sweep,
        "linear_cv_T1024": lin,
        "fp32_cubature_error_per_T": errs,
        "fp32_stability": stab,
        "reference_published_T4096": pub,
        "beyond_the_smoothing_path": wide,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
