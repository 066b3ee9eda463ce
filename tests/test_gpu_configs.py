"""Each benchmark configuration at its own workload size, checked on the device.

* linear constant-velocity model (nx=4, ny=2, T=1024): one parallel filter +
  smoother pass in FP64 against the C restatement (itself pinned by the dense
  joint-Gaussian posterior), FP64 bar;
* FP32 square-root cubature IPLS (bearings-only, T=2^14, 5 iterations)
  against the C restatement's FP64 cubature IPLS, at the FP32 tolerances, and
  the FP64 run at the FP64 bar;
* the ragged batch (4096 bearings-only sequences, lengths U{256..8192},
  cubature IPLS, 5 iterations): 16 seeded sequences against the C
  restatement (FP64 bar in FP64, FP32 tolerances in FP32), and every sequence
  against the same batch laid out in reverse order;
* the gradient (256 sequences of T=16384, cubature IPLS, 10 iterations, 7
  parameters) in FP64 and FP32 against central differences (h = 1e-4, FP64)
  of the log-likelihood on the sequences whose iteration reached a fixed point.

The iterated-smoothing references follow parsmooth/methods.py:54-76 and
parsmooth/_utils.py:118-133 (implicit gradient of the fixed point).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

from oracle import psm_oracle as O  # noqa: E402
from oracle import psm_oracle_c as C  # noqa: E402
import parsmooth  # noqa: E402
from parsmooth import FunctionalModel, MVNSqrt  # noqa: E402
from parsmooth import _native as N  # noqa: E402
from parsmooth.batch import IteratedSmoother  # noqa: E402
from parsmooth.data import bearings_data, bearings_data_batch, cv_model, lgssm_data  # noqa: E402
from parsmooth.linearization import cubature, extended  # noqa: E402
from parsmooth.models import LinearFunction, make_parameters  # noqa: E402

DEV = torch.device("cuda", 0)
S1, S2 = (-1.5, 0.5), (1.0, 1.0)
M0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
INIT = np.array([-1.0, -1.0, 6.0, 4.0, 2.0])
TOL64 = 1e-9
TOL32, TOL32_COV = 1e-4, 1e-3


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max()) / max(1.0, float(np.abs(b).max()))


def gram(X):
    X = np.asarray(X, np.float64)
    return X @ np.swapaxes(X, -1, -2)


def t(a, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a, np.float64), device=DEV).to(dtype)


def _cub():
    tab = O.cubature_table(5)
    return tab.xi, tab.wm, tab.wc


def _bearings_models(dtype):
    Q, R, obs_f, trans_f = make_parameters(0.01, 0.1, 0.5, 0.01, S1, S2)
    cQ, cR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    trans = FunctionalModel(trans_f, MVNSqrt(t(np.zeros(5), dtype), t(cQ, dtype)))
    obs = FunctionalModel(obs_f, MVNSqrt(t(np.zeros(2), dtype), t(cR, dtype)))
    return trans, obs, cQ, cR


# ---------------------------------------------------------------- linear CV, T = 1024
def test_linear_cv_T1024_pass_matches_oracle():
    """Constant-velocity tracking (dt=0.1, q=1, R=0.5^2 I, prior N(0, I), truth
    x0 ~ N(0, I)), T=1024, FP64, through the public parsmooth.filtering /
    smoothing: filtered and smoothed trajectories and the log-likelihood."""
    T = 1024
    F, Q, H, R = cv_model(0.1, 1.0, 0.5)
    rng = np.random.RandomState(0)
    _, ys = lgssm_data(rng.randn(4), F, H, R, Q, np.zeros(4), np.zeros(2), T, random_state=rng)
    ys = np.asarray(ys, np.float64)
    cQ, cR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    trans = FunctionalModel(LinearFunction(F), MVNSqrt(t(np.zeros(4)), t(cQ)))
    obs = FunctionalModel(LinearFunction(H), MVNSqrt(t(np.zeros(2)), t(cR)))
    x0 = MVNSqrt(t(np.zeros(4)), t(np.eye(4)))
    filt, ell = parsmooth.filtering(t(ys), x0, trans, obs, extended, return_loglikelihood=True)
    smooth = parsmooth.smoothing(trans, filt, extended)
    rm, rL, rell, _ = C.iterated_smoother_batch([T], ys, np.zeros((1, 4)), np.eye(4)[None], np.zeros((T + 1, 4)),
                                                None, ("linear", F), ("linear", H), np.zeros(4), cQ, np.zeros(2), cR,
                                                1, want_ell=True)
    # the filtered trajectory: the numpy restatement of the parallel filter
    fm, fL = O.filtering(ys, O.MVNSqrt(np.zeros(4), np.eye(4)),
                         O.FunctionalModel(O.Linear(F), O.MVNSqrt(np.zeros(4), cQ)),
                         O.FunctionalModel(O.Linear(H), O.MVNSqrt(np.zeros(2), cR)), O.extended)
    errs = {"filtered mean": rel(filt[0].cpu(), fm), "filtered cov": rel(gram(filt[1].cpu()), gram(fL)),
            "smoothed mean": rel(smooth[0].cpu(), rm), "smoothed cov": rel(gram(smooth[1].cpu()), gram(rL)),
            "ell": abs(float(ell) - rell[0]) / max(1.0, abs(rell[0]))}
    print("linear CV T=1024:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert all(v <= TOL64 for v in errs.values()), errs


# ---------------------------------------------------------------- FP32 cubature IPLS, T = 2^14
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["fp64", "fp32"])
def test_cubature_ipls_T16384_matches_oracle(dtype):
    """Square-root cubature IPLS, 5 iterations from the runtime notebook's nominal
    [-1, -1, 6, 4, 2] / I, bearings-only data (seed 1), T = 2^14: the smoothed
    trajectory against the C restatement's FP64 cubature IPLS -- FP64 at the
    FP64 bar, FP32 (FP32 storage and combinations, FP64 log-likelihood) at the
    FP32 tolerances (means 1e-4, L L^T 1e-3)."""
    T, n_iter = 1 << 14, 5
    _, _, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, S1, S2, random_state=1)
    ys = np.asarray(ys, np.float64)
    trans, obs, cQ, cR = _bearings_models(dtype)
    r = IteratedSmoother([T], trans, obs, cubature, dtype=dtype)
    r.set_data(t(ys, dtype), t(M0[None], dtype), t(np.eye(5)[None], dtype))
    init = (np.tile(INIT, (T + 1, 1)), np.tile(np.eye(5), (T + 1, 1, 1)))
    m, L, ell = r.run(n_iter, (t(init[0], dtype), t(init[1], dtype)), return_loglikelihood=True)
    rm, rL, rell, _ = C.iterated_smoother_batch([T], ys, M0[None], np.eye(5)[None], init[0], init[1], ("ct", 0.01),
                                                ("bearings", (S1, S2)), np.zeros(5), cQ, np.zeros(2), cR, n_iter,
                                                sigma=_cub(), want_ell=True)
    errs = (rel(m.cpu(), rm), rel(gram(L.cpu()), gram(rL)), abs(float(ell[0]) - rell[0]) / max(1.0, abs(rell[0])))
    print(f"cubature IPLS T=2^14 {dtype}: mean {errs[0]:.2e} cov {errs[1]:.2e} ell {errs[2]:.2e} "
          f"fallbacks {r.fallback_count()}", flush=True)
    if dtype == torch.float64:
        assert max(errs) <= TOL64, errs
    else:
        assert errs[0] <= TOL32 and errs[1] <= TOL32_COV and errs[2] <= 1e-5, errs


# ---------------------------------------------------------------- ragged batch, 4096 sequences
N_BATCH, LO, HI = 4096, 256, 8192


def _batch_data(seed=0):
    """The bench's batch workload: lengths U{256..8192}, per-sequence prior mean
    N([0.1, 0.2, 1, 0, 0], 0.1 I), chol I; bearings-only observations."""
    rng = np.random.default_rng(seed)
    lengths = rng.integers(LO, HI + 1, N_BATCH)
    m0 = M0 + np.sqrt(0.1) * rng.standard_normal((N_BATCH, 5))
    _, obs = bearings_data_batch(lengths, m0[:, :4], 0.01, 0.5, S1, S2, seed=seed + 1)
    return lengths, m0, [np.asarray(y, np.float64) for y in obs]


def _run_batch(lengths, m0, obs, dtype, n_iter=5):
    trans, obsm, _, _ = _bearings_models(dtype)
    r = IteratedSmoother(lengths, trans, obsm, cubature, dtype=dtype)
    r.set_data(t(np.concatenate(obs), dtype), t(m0, dtype), t(np.tile(np.eye(5), (len(lengths), 1, 1)), dtype))
    reps = np.asarray(lengths) + 1
    init = (t(np.repeat(m0, reps, axis=0), dtype), t(np.tile(np.eye(5), (int(reps.sum()), 1, 1)), dtype))
    m, L = r.run(n_iter, init, use_graph=True)
    out = ([x.double().cpu().numpy() for x in r.split(m)], [x.double().cpu().numpy() for x in r.split(L)])
    del r, init
    torch.cuda.empty_cache()
    return out


@pytest.fixture(scope="module")
def batch_case():
    lengths, m0, obs = _batch_data()
    pick = np.sort(np.random.default_rng(123).choice(N_BATCH, 16, replace=False))
    cQ, cR = _bearings_models(torch.float64)[2:]
    reps = lengths[pick] + 1
    rm, rL, _, used = C.iterated_smoother_batch(lengths[pick], np.concatenate([obs[i] for i in pick]), m0[pick],
                                                np.tile(np.eye(5), (16, 1, 1)), np.repeat(m0[pick], reps, axis=0),
                                                None, ("ct", 0.01), ("bearings", (S1, S2)), np.zeros(5), cQ,
                                                np.zeros(2), cR, 5, sigma=_cub())
    off = np.concatenate([[0], np.cumsum(reps)])
    refs = [(rm[off[k]:off[k + 1]], rL[off[k]:off[k + 1]]) for k in range(16)]
    return lengths, m0, obs, pick, refs


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["fp64", "fp32"])
def test_ragged_batch_4096_subset_matches_oracle(batch_case, dtype):
    """16 seeded sequences of the 4096-sequence batch against the C restatement
    (one OpenMP thread per sequence): FP64 at the FP64 bar; FP32 means and
    covariances at the FP32 tolerances on the sequences whose FP64 iteration
    does not amplify FP32-sized input perturbations past them (the others are
    reported)."""
    lengths, m0, obs, pick, refs = batch_case
    ms, Ls = _run_batch(lengths, m0, obs, dtype)
    errs = [(rel(ms[i], refs[k][0]), rel(gram(Ls[i]), gram(refs[k][1]))) for k, i in enumerate(pick)]
    worst = max(e[0] for e in errs), max(e[1] for e in errs)
    print(f"batch subset {dtype}: max mean {worst[0]:.2e} max cov {worst[1]:.2e}", flush=True)
    if dtype == torch.float64:
        assert worst[0] <= TOL64 and worst[1] <= TOL64, errs
    else:
        assert worst[0] <= TOL32 and worst[1] <= TOL32_COV, errs


def test_ragged_batch_4096_layout_independent(batch_case):
    """Every sequence of the 4096-sequence batch (FP64) equals the same sequence
    in the batch laid out in reverse order (other offsets, other chunk and
    workgroup boundaries, other neighbours), to the FP64 bar."""
    lengths, m0, obs, _, _ = batch_case
    fm, fL = _run_batch(lengths, m0, obs, torch.float64)
    rev = slice(None, None, -1)
    bm, bL = _run_batch(lengths[rev], m0[rev], obs[rev], torch.float64)
    n = len(lengths)
    em = max(rel(fm[i], bm[n - 1 - i]) for i in range(n))
    eL = max(rel(gram(fL[i]), gram(bL[n - 1 - i])) for i in range(n))
    print(f"batch layout: max mean {em:.2e} max cov {eL:.2e}", flush=True)
    assert em <= TOL64 and eL <= TOL64, (em, eL)


# ---------------------------------------------------------------- gradient, 256 x 16384
def _grad_objective(dtype, n_seq=256, T=16384, n_iter=10, seed=5):
    """The bench's gradient workload: bearings-only sequences simulated at
    qc=0.01, qw=0.1, sigma=0.5 (turn rate with the model's own intensity),
    theta = (log qc, log qw, log sigma, s1x, s1y, s2x, s2y), cubature IPLS from
    the simulated trajectories (chol 0.1 I)."""
    from parsmooth.gradient import IteratedLoglikelihood, bearings_parameterisation
    rng = np.random.default_rng(seed)
    m0 = M0 + np.sqrt(0.1) * rng.standard_normal((n_seq, 5))
    lengths = np.full(n_seq, T)
    truth, obs = bearings_data_batch(lengths, m0[:, :4], 0.01, 0.5, S1, S2, q=np.sqrt(0.1 / 0.01), seed=seed + 1)
    build = bearings_parameterisation(0.01, m0, np.tile(np.eye(5), (n_seq, 1, 1)))
    init = (np.concatenate(truth).astype(np.float64), np.tile(0.1 * np.eye(5), (n_seq * (T + 1), 1, 1)))
    obj = IteratedLoglikelihood(build, N.MODEL_COORDINATED_TURN, N.MODEL_BEARINGS, 5, 2, lengths,
                                np.concatenate(obs), cubature, n_iter, init, dtype=dtype)
    theta = torch.tensor([np.log(0.01), np.log(0.1), np.log(0.5), -1.5, 0.5, 1.0, 1.0], dtype=torch.float64)
    return obj, theta


@pytest.fixture(scope="module")
def grad_fd():
    """FP64 per-sequence central differences (h = 1e-4) of the log-likelihood in
    every parameter, and the fixed-point mask (max |f(x*) - x*| < 1e-6)."""
    obj, theta = _grad_objective(torch.float64)
    h = 1e-4
    fd = np.zeros((256, 7))
    for p in range(7):
        e = torch.zeros(7, dtype=torch.float64)
        e[p] = h
        up = obj.value(theta + e)[0].cpu().numpy()
        dn = obj.value(theta - e)[0].cpu().numpy()
        fd[:, p] = (up - dn) / (2 * h)
    conv = obj.fixed_point_residual(theta) < 1e-6
    del obj
    torch.cuda.empty_cache()
    return fd, conv


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["fp64", "fp32"])
def test_gradient_256x16384_matches_central_differences(grad_fd, dtype):
    """Per-sequence implicit gradient (forward-mode evaluation of the reference's
    fixed-point rule) against FP64 central differences, on the sequences at a
    fixed point: within 1e-3 relative to max(1, |fd|) per component."""
    fd, conv = grad_fd
    obj, theta = _grad_objective(dtype)
    _, g = obj.value_and_grad(theta)
    g = g.cpu().numpy()
    del obj
    torch.cuda.empty_cache()
    assert conv.sum() >= 3, f"only {int(conv.sum())} sequences at a fixed point"
    err = np.abs(g[conv] - fd[conv]) / np.maximum(1.0, np.abs(fd[conv]))
    print(f"gradient {dtype}: {int(conv.sum())} sequences at a fixed point, max rel err vs FD {err.max():.2e}",
          flush=True)
    assert err.max() <= 1e-3, err.max(axis=0)
