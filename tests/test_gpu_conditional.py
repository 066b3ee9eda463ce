"""ConditionalMomentsModel on the device (torch-evaluated conditional moments,
PSM_MODEL_CONDITIONAL) against the oracle's restatement of the reference's
conditional linearizations."""
import math

import numpy as np
import pytest
import torch

from oracle import psm_oracle as O

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

import parsmooth  # noqa: E402
from parsmooth import ConditionalMomentsModel, MVNSqrt, MVNStandard  # noqa: E402
from parsmooth.linearization import cubature, extended, gauss_hermite, unscented  # noqa: E402

from test_oracle import _CondConst, _PoissonFactor, _PoissonMean, _PopMean  # noqa: E402

DEV = torch.device("cuda", 0)
t64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=DEV)
npy = lambda t: t.detach().double().cpu().numpy()
METHODS = {"extended": extended, "cubature": cubature, "gauss_hermite": gauss_hermite, "unscented": unscented}
OMETH = {"extended": O.extended, "cubature": O.cubature, "gauss_hermite": O.gauss_hermite,
         "unscented": O.unscented}


def close(a, b, tol, what):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    err = float(np.abs(a - b).max()) / max(1.0, float(np.abs(b).max()))
    assert err <= tol, f"{what}: rel err {err:.3e} > {tol:.1e}"


def pop_models(sqrt):
    """Population (Ricker) transition with noise 0.3 and Poisson(10 exp x)
    observation (notebooks/population_model.py, tests/test_linearization.py:45-64)."""
    tm = ConditionalMomentsModel(lambda x: math.log(44.7) + x - torch.exp(x),
                                 lambda x: torch.full((1, 1), 0.3 if sqrt else 0.09, dtype=x.dtype, device=x.device))
    om = ConditionalMomentsModel(lambda x: 10.0 * torch.exp(x),
                                 (lambda x: torch.sqrt(10.0 * torch.exp(x)).reshape(1, 1)) if sqrt
                                 else (lambda x: (10.0 * torch.exp(x)).reshape(1, 1)))
    to = O.ConditionalMomentsModel(_PopMean(), _CondConst(np.array([[0.3 if sqrt else 0.09]])))
    oo = O.ConditionalMomentsModel(_PoissonMean(10.0), _PoissonFactor(10.0, sqrt))
    return (tm, om), (to, oo)


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_conditional_linearize_matches_oracle(method, sqrt):
    rng = np.random.RandomState(0)
    m = rng.randn(50, 1) * 0.5 + 1.0
    L = 0.2 + 0.3 * rng.rand(50, 1, 1)
    (tm, om), (to, oo) = pop_models(sqrt)
    x = MVNSqrt(t64(m), t64(L)) if sqrt else MVNStandard(t64(m), t64(L * L))
    xo = O.MVNSqrt(m, L) if sqrt else O.MVNStandard(m, L * L)
    for dev_model, o_model in ((tm, to), (om, oo)):
        got = METHODS[method](dev_model, x)
        ref = OMETH[method](o_model, xo)
        for u, v, name in zip(got, ref, "FCb"):
            close(npy(u), v, 1e-10, f"{method} {name}")


@pytest.mark.parametrize("method", ["extended", "cubature"])
@pytest.mark.parametrize("sqrt", [True, False])
def test_population_model_filter_smoother_matches_oracle(method, sqrt):
    rng = np.random.RandomState(3)
    T = 120
    x = np.empty(T + 1)
    x[0] = 1.0
    for k in range(T):
        x[k + 1] = math.log(44.7) + x[k] - math.exp(x[k]) + 0.3 * rng.randn()
    ys = rng.poisson(10.0 * np.exp(x[1:]))[:, None].astype(np.float64)
    (tm, om), (to, oo) = pop_models(sqrt)
    nom = (x[:, None] + 0.1 * rng.randn(T + 1, 1), np.full((T + 1, 1, 1), 0.5))
    if sqrt:
        x0d, x0o = MVNSqrt(t64([1.0]), t64([[1.0]])), O.MVNSqrt(np.array([1.0]), np.array([[1.0]]))
        nomd, nomo = MVNSqrt(t64(nom[0]), t64(nom[1])), O.MVNSqrt(*nom)
    else:
        x0d, x0o = MVNStandard(t64([1.0]), t64([[1.0]])), O.MVNStandard(np.array([1.0]), np.array([[1.0]]))
        nomd, nomo = MVNStandard(t64(nom[0]), t64(nom[1] ** 2)), O.MVNStandard(nom[0], nom[1] ** 2)
    ref = O.filter_smoother(ys, x0o, to, oo, OMETH[method], nomo, True)
    got = parsmooth.filter_smoother(t64(ys), x0d, tm, om, METHODS[method], nomd)
    close(npy(got[0]), ref[0], 1e-9, "mean")
    close(npy(got[1]), ref[1], 1e-9, "chol/cov")
    # and the iterated smoother with the default nominal runs end to end
    res = parsmooth.iterated_smoothing(t64(ys), x0d, tm, om, METHODS[method], criterion=lambda i, *_: i < 5)
    assert torch.isfinite(res[0]).all()


# ------------------------------------------- get_conditional_model on the device
from parsmooth import FunctionalModel  # noqa: E402
from parsmooth.linearization import get_conditional_model  # noqa: E402
from parsmooth.models import LinearFunction  # noqa: E402

from test_conditional_model import lin_xq_torch, mult_noise_torch  # noqa: E402
from test_oracle import _LinXQ, _MultNoise, _rand_lower  # noqa: E402


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_get_conditional_model_linear_on_device(dim, seed, method, sqrt):
    """tests/test_linearization.py:238-283 through psm_linearize: the outer
    linearization of the conditional model of a x + b q + c recovers a, the
    mean and (b chol_q)(b chol_q)^T, and equals the oracle's."""
    np.random.seed(seed)
    a, b, c = np.random.randn(dim, dim), np.random.randn(dim, dim), np.random.randn(dim)
    m_x, m_q = np.random.randn(dim), np.random.randn(dim)
    chol_x, chol_q = _rand_lower(dim), _rand_lower(dim)
    if sqrt:
        x_d, q_d = MVNSqrt(t64(m_x), t64(chol_x)), MVNSqrt(t64(m_q), t64(chol_q))
        x_o, q_o = O.MVNSqrt(m_x, chol_x), O.MVNSqrt(m_q, chol_q)
    else:
        P, Pq = chol_x @ chol_x.T, chol_q @ chol_q.T
        x_d, q_d = MVNStandard(t64(m_x), t64(P)), MVNStandard(t64(m_q), t64(Pq))
        x_o, q_o = O.MVNStandard(m_x, P), O.MVNStandard(m_q, Pq)
    model = get_conditional_model(lin_xq_torch(a, b, c), q_d, METHODS[method])
    F, C, rem = (npy(v) for v in METHODS[method](model, x_d))
    ref = OMETH[method](O.get_conditional_model(_LinXQ(a, b, c), q_o, OMETH[method]), x_o)
    for u, v, name in zip((F, C, rem), ref, "FCb"):
        close(u, v, 1e-10, f"{method} {name} vs oracle")
    Q = C @ C.T if sqrt else C
    xp = np.random.randn(dim)
    close(F, a, 1e-9, "F = a")
    close(F @ xp + rem, a @ xp + b @ m_q + c, 1e-9, "mean")
    close(Q, (b @ chol_q) @ (b @ chol_q).T, 1e-9, "Q")


def _mult_noise_problem(T, seed, sqrt, method):
    """2-D state through f(x, q) = g(x) exp(0.2 q) + 0.1 q, q ~ N(0, 0.3^2 I),
    observed as x0 + 0.5 x1 with noise 0.2."""
    rng = np.random.RandomState(seed)
    f = _MultNoise()
    sq = 0.3
    x = np.empty((T + 1, 2))
    x[0] = [0.5, -0.2]
    for k in range(T):
        x[k + 1] = f(x[k], sq * rng.randn(2))
    H = np.array([[1.0, 0.5]])
    ys = x[1:] @ H.T + 0.2 * rng.randn(T, 1)
    m0, L0 = np.array([0.5, -0.2]), 0.5 * np.eye(2)
    Lq, Lr = sq * np.eye(2), np.array([[0.2]])
    if sqrt:
        q_d, q_o = MVNSqrt(t64(np.zeros(2)), t64(Lq)), O.MVNSqrt(np.zeros(2), Lq)
        obs_d = FunctionalModel(LinearFunction(H), MVNSqrt(t64(np.zeros(1)), t64(Lr)))
        obs_o = O.FunctionalModel(O.Linear(H), O.MVNSqrt(np.zeros(1), Lr))
        x0d, x0o = MVNSqrt(t64(m0), t64(L0)), O.MVNSqrt(m0, L0)
    else:
        q_d, q_o = MVNStandard(t64(np.zeros(2)), t64(Lq @ Lq.T)), O.MVNStandard(np.zeros(2), Lq @ Lq.T)
        obs_d = FunctionalModel(LinearFunction(H), MVNStandard(t64(np.zeros(1)), t64(Lr @ Lr.T)))
        obs_o = O.FunctionalModel(O.Linear(H), O.MVNStandard(np.zeros(1), Lr @ Lr.T))
        x0d, x0o = MVNStandard(t64(m0), t64(L0 @ L0.T)), O.MVNStandard(m0, L0 @ L0.T)
    trans_d = get_conditional_model(mult_noise_torch, q_d, METHODS[method])
    trans_o = O.get_conditional_model(_MultNoise(), q_o, OMETH[method])
    return ys, x, (x0d, trans_d, obs_d), (x0o, trans_o, obs_o), rng


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_noise_nonlinear_filter_smoother_matches_oracle(method, sqrt):
    """A transition non-linear in its noise, through get_conditional_model:
    the filter (with log-likelihood) and smoother at a fixed nominal trajectory
    and three iterations of the iterated smoother match the oracle to 1e-9."""
    T = 60
    ys, x, (x0d, td, od), (x0o, to, oo), rng = _mult_noise_problem(T, 5, sqrt, method)
    nm = x + 0.05 * rng.randn(T + 1, 2)
    nL = np.tile(0.3 * np.eye(2), (T + 1, 1, 1))
    if sqrt:
        nomd, nomo = MVNSqrt(t64(nm), t64(nL)), O.MVNSqrt(nm, nL)
    else:
        nomd, nomo = MVNStandard(t64(nm), t64(nL @ nL)), O.MVNStandard(nm, nL @ nL)
    filt, ell = parsmooth.filtering(t64(ys), x0d, td, od, METHODS[method], nomd, return_loglikelihood=True)
    rf, rell = O.filtering(ys, x0o, to, oo, OMETH[method], nomo, return_loglikelihood=True)
    close(npy(filt[0]), rf[0], 1e-9, "filtered mean")
    close(npy(filt[1]), rf[1], 1e-9, "filtered chol/cov")
    close(float(ell), rell, 1e-9, "log-likelihood")
    got = parsmooth.filter_smoother(t64(ys), x0d, td, od, METHODS[method], nomd)
    ref = O.filter_smoother(ys, x0o, to, oo, OMETH[method], nomo, True)
    close(npy(got[0]), ref[0], 1e-9, "smoothed mean")
    close(npy(got[1]), ref[1], 1e-9, "smoothed chol/cov")
    crit = lambda i, *_: i < 3
    got = parsmooth.iterated_smoothing(t64(ys), x0d, td, od, METHODS[method], nomd, criterion=crit)
    ref = O.iterated_smoothing(ys, x0o, to, oo, OMETH[method], nomo, criterion=crit)
    close(npy(got[0]), ref[0], 1e-9, "iterated mean")
    close(npy(got[1]), ref[1], 1e-9, "iterated chol/cov")


def test_get_conditional_model_with_device_constants():
    """f keeps a constant on the device while q comes as numpy arrays: the
    dimension probe finds the device, and the device linearization of the
    resulting model equals the oracle's."""
    rng = np.random.RandomState(5)
    a, b = rng.randn(2, 2), 0.5 * np.eye(2)
    A = t64(a)
    f = lambda x, q: x @ A.T + 0.5 * q   # noqa: E731
    cq = np.array([[0.3, 0.0], [0.1, 0.2]])
    model = get_conditional_model(f, MVNSqrt(np.zeros(2), cq), cubature)
    m, L = rng.randn(7, 2), np.tile(np.diag([0.4, 0.5]), (7, 1, 1))
    got = cubature(model, MVNSqrt(t64(m), t64(L)))
    ref = O.cubature(O.get_conditional_model(_LinXQ(a, b, np.zeros(2)), O.MVNSqrt(np.zeros(2), cq), O.cubature),
                     O.MVNSqrt(m, L))
    for u, v, name in zip(got, ref, "FCb"):
        close(npy(u), v, 1e-10, f"device-constant model {name}")


@pytest.mark.parametrize("sqrt", [True, False])
def test_population_helpers_match_oracle(sqrt):
    """parsmooth.models.population_parameters on parsmooth.data.population_data
    (notebooks/population_model.py): filter and smoother at a nominal, and three
    IPLS iterations, against the oracle's population model."""
    from parsmooth.data import population_data
    from parsmooth.models import population_parameters
    Q, lam, T = np.array([[0.09]]), 10.0, 80
    states, ys = population_data(np.array([1.0]), T, Q, lam, random_state=4)
    mt, ct, mo, co, cht, cho = population_parameters(lam, Q)
    tm = ConditionalMomentsModel(mt, cht if sqrt else ct)
    om = ConditionalMomentsModel(mo, cho if sqrt else co)
    to = O.ConditionalMomentsModel(_PopMean(), _CondConst(np.array([[0.3 if sqrt else 0.09]])))
    oo = O.ConditionalMomentsModel(_PoissonMean(lam), _PoissonFactor(lam, sqrt))
    rng = np.random.RandomState(1)
    nm, nL = states + 0.1 * rng.randn(T + 1, 1), np.full((T + 1, 1, 1), 0.5)
    if sqrt:
        x0d, x0o = MVNSqrt(t64([1.0]), t64([[1.0]])), O.MVNSqrt(np.array([1.0]), np.array([[1.0]]))
        nomd, nomo = MVNSqrt(t64(nm), t64(nL)), O.MVNSqrt(nm, nL)
    else:
        x0d, x0o = MVNStandard(t64([1.0]), t64([[1.0]])), O.MVNStandard(np.array([1.0]), np.array([[1.0]]))
        nomd, nomo = MVNStandard(t64(nm), t64(nL ** 2)), O.MVNStandard(nm, nL ** 2)
    got = parsmooth.filter_smoother(t64(ys), x0d, tm, om, cubature, nomd)
    ref = O.filter_smoother(ys, x0o, to, oo, O.cubature, nomo, True)
    close(npy(got[0]), ref[0], 1e-9, "smoothed mean")
    close(npy(got[1]), ref[1], 1e-9, "smoothed chol/cov")
    crit = lambda i, *_: i < 3
    got = parsmooth.iterated_smoothing(t64(ys), x0d, tm, om, cubature, nomd, criterion=crit)
    ref = O.iterated_smoothing(ys, x0o, to, oo, O.cubature, nomo, criterion=crit)
    close(npy(got[0]), ref[0], 1e-9, "IPLS mean")
