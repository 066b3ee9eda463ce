"""The bench workload at its full size (bearings-only IEKS, T = 2^17, FP64 and
FP32) against the C restatement of the sequential square-root filter and
smoother (oracle/psm_oracle_c.c), which finishes one pass in about a second."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

from oracle import psm_oracle_c as C  # noqa: E402
from parsmooth import FunctionalModel, MVNSqrt, MVNStandard  # noqa: E402
from parsmooth import _native as N  # noqa: E402
from parsmooth.batch import IteratedSmoother  # noqa: E402
from parsmooth.data import bearings_data  # noqa: E402
from parsmooth.linearization import cubature, extended  # noqa: E402
from parsmooth.models import make_parameters  # noqa: E402

DEV = torch.device("cuda", 0)
T = 1 << 17
S1, S2 = (-1.5, 0.5), (1.0, 1.0)
M0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
INIT = np.array([-1.0, -1.0, 6.0, 4.0, 2.0])


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max()) / max(1.0, float(np.abs(b).max()))


def _problem():
    _, _, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, S1, S2, random_state=0)
    Q, R, obs_f, trans_f = make_parameters(0.01, 0.1, 0.5, 0.01, S1, S2)
    return np.asarray(ys, np.float64), np.linalg.cholesky(Q), np.linalg.cholesky(R), obs_f, trans_f


# the schedule the headline bench line runs (autotuned on MI355X): chunk 5,
# scan fan-in 2, team-cooperative upper levels, half-wave level 0
HEADLINE_SCHEDULE = (5, 2 | N.SCAN_TEAM | N.SCAN_HALF)


def _device(dtype, n_iter, sqrt=True, method=extended, schedule=None):
    ys, cQ, cR, obs_f, trans_f = _problem()
    t = lambda a: torch.as_tensor(np.asarray(a, np.float64), device=DEV).to(dtype)   # noqa: E731
    mvn = MVNSqrt if sqrt else MVNStandard
    q, r_ = (cQ, cR) if sqrt else (cQ @ cQ.T, cR @ cR.T)
    r = IteratedSmoother([T], FunctionalModel(trans_f, mvn(t(np.zeros(5)), t(q))),
                         FunctionalModel(obs_f, mvn(t(np.zeros(2)), t(r_))), method, dtype=dtype, sqrt=sqrt)
    r.set_data(t(ys), t(M0[None]), t(np.eye(5)[None]))
    if schedule is not None:
        r.prob.set_schedule(*schedule)
        assert r.chunk_size() == schedule[0]
    init = (t(np.tile(INIT, (T + 1, 1))), t(np.tile(np.eye(5), (T + 1, 1, 1))))
    m, L, ell = r.run(n_iter, init, return_loglikelihood=True)
    return m.double().cpu().numpy(), L.double().cpu().numpy(), float(ell[0])


def _oracle(n_iter):
    ys, cQ, cR, _, _ = _problem()
    return C.iterated_smoother(ys, M0, np.eye(5), np.tile(INIT, (T + 1, 1)), ("ct", 0.01), ("bearings", (S1, S2)),
                               np.zeros(5), cQ, np.zeros(2), cR, n_iter, want_ell=True)


def gram(X):
    return X @ np.swapaxes(X, -1, -2)


@pytest.mark.parametrize("n_iter,schedule", [(1, None), (10, None), (10, HEADLINE_SCHEDULE)],
                         ids=["1-default", "10-default", "10-headline"])
def test_full_size_fp64_matches_c_oracle(n_iter, schedule):
    """One pass from the notebook's nominal, and the whole bench workload (10
    iterations, with the default schedule and with the headline's autotuned
    one): smoothed means, covariances and the log-likelihood at the final
    nominal within the FP64 bar, 1e-9 relative to max(1, |x|)."""
    rm, rL, rell = _oracle(n_iter)
    m, L, ell = _device(torch.float64, n_iter, schedule=schedule)
    errs = (rel(m, rm), rel(gram(L), gram(rL)), abs(ell - rell) / max(1.0, abs(rell)))
    print(f"FP64 n_iter {n_iter} schedule {schedule}: mean {errs[0]:.2e} cov {errs[1]:.2e} ell {errs[2]:.2e}",
          flush=True)
    assert errs[0] <= 1e-9 and errs[1] <= 1e-9 and errs[2] <= 1e-9, errs


def test_full_size_fp32_pass_matches_c_oracle():
    """One FP32 pass at full size against the FP64 restatement (means 1e-4,
    covariances 1e-3 -- L L^T of an FP32 factor doubles the rounding --
    log-likelihood 1e-5; it accumulates in FP64)."""
    rm, rL, rell = _oracle(1)
    m, L, ell = _device(torch.float32, 1)
    errs = (rel(m, rm), rel(gram(L), gram(rL)), abs(ell - rell) / max(1.0, abs(rell)))
    print(f"FP32 one pass: mean {errs[0]:.2e} cov {errs[1]:.2e} ell {errs[2]:.2e}", flush=True)
    assert errs[0] <= 1e-4 and errs[1] <= 1e-3 and errs[2] <= 1e-5, errs


@pytest.mark.parametrize("method", [extended, cubature], ids=["extended", "cubature"])
def test_full_size_covariance_form_matches_square_root_form(method):
    """The covariance form (MVNStandard) of one full-size FP64 pass equals the
    square-root form's (pinned to the C restatement above for extended)."""
    m_s, L_s, ell_s = _device(torch.float64, 1, method=method)
    m_c, P_c, ell_c = _device(torch.float64, 1, sqrt=False, method=method)
    errs = (rel(m_c, m_s), rel(P_c, gram(L_s)), abs(ell_c - ell_s) / max(1.0, abs(ell_s)))
    print(f"covariance vs square-root form: mean {errs[0]:.2e} cov {errs[1]:.2e} ell {errs[2]:.2e}", flush=True)
    assert errs[0] <= 1e-9 and errs[1] <= 1e-9 and errs[2] <= 1e-9, errs


def test_full_size_sequential_pass_matches_c_oracle():
    """parallel=False at full size: one lane runs the sequential filter and
    smoother through all 2^17 steps at the notebook's nominal; the C
    restatement runs the same recursion (FP64 bar)."""
    import parsmooth
    ys, cQ, cR, obs_f, trans_f = _problem()
    rm, rL, rell = _oracle(1)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float64), device=DEV)   # noqa: E731
    trans = FunctionalModel(trans_f, MVNSqrt(t(np.zeros(5)), t(cQ)))
    obs = FunctionalModel(obs_f, MVNSqrt(t(np.zeros(2)), t(cR)))
    nom = MVNSqrt(t(np.tile(INIT, (T + 1, 1))), t(np.tile(np.eye(5), (T + 1, 1, 1))))
    x0 = MVNSqrt(t(M0), t(np.eye(5)))
    out = parsmooth.filter_smoother(t(ys), x0, trans, obs, extended, nom, parallel=False)
    m, L = out[0].cpu().numpy(), out[1].cpu().numpy()
    errs = (rel(m, rm), rel(gram(L), gram(rL)))
    print(f"sequential pass: mean {errs[0]:.2e} cov {errs[1]:.2e}", flush=True)
    assert errs[0] <= 1e-9 and errs[1] <= 1e-9, errs


def test_full_size_sample_mean():
    """64 pathwise samples of the full-size smoothing distribution (seeded,
    filter, smoother and sampler at one nominal):
    at 99.9 % of the states every coordinate of the sample mean lies within 4
    standard errors of the smoothed mean, and at every state within 7."""
    import parsmooth
    ys, cQ, cR, obs_f, trans_f = _problem()
    t = lambda a: torch.as_tensor(np.asarray(a, np.float64), device=DEV)   # noqa: E731
    trans = FunctionalModel(trans_f, MVNSqrt(t(np.zeros(5)), t(cQ)))
    obs = FunctionalModel(obs_f, MVNSqrt(t(np.zeros(2)), t(cR)))
    x0 = MVNSqrt(t(M0), t(np.eye(5)))
    nom = MVNSqrt(t(np.tile(INIT, (T + 1, 1))), t(np.tile(np.eye(5), (T + 1, 1, 1))))
    filt = parsmooth.filtering(t(ys), x0, trans, obs, extended, nom)
    smooth = parsmooth.smoothing(trans, filt, extended, nom)
    S = 64
    # the samples' mean is the smoother's at the sampler's own nominal: pass the same one
    samples = parsmooth.sampling(11, S, trans, filt, extended, nom).cpu().numpy()
    sm, sL = smooth[0].cpu().numpy(), smooth[1].cpu().numpy()
    se = np.sqrt(np.einsum("kij,kij->ki", sL, sL) / S)           # sqrt(diag(L L^T) / S)
    z = np.abs(samples.mean(axis=1) - sm) / np.maximum(se, 1e-300)
    frac = float(np.mean(np.all(z <= 4.0, axis=1)))
    print(f"sample mean: {frac:.5f} of states within 4 SE, max z {z.max():.2f}", flush=True)
    assert frac >= 0.999 and z.max() <= 7.0, (frac, float(z.max()))
