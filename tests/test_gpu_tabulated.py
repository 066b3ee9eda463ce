"""Torch-callable model functions (PSM_MODEL_TABULATED): the host evaluates the
function at the linearization points, the kernels do the rest.  Checked
against the built-in device models and against the FP64 oracle on a model
with no built-in counterpart."""
import numpy as np
import pytest
import torch

from oracle import psm_oracle as O

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

import parsmooth  # noqa: E402
from parsmooth import FunctionalModel, MVNSqrt, MVNStandard  # noqa: E402
from parsmooth import models as M  # noqa: E402
from parsmooth.linearization import cubature, extended, gauss_hermite, unscented  # noqa: E402

DEV = torch.device("cuda", 0)
t64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=DEV)
npy = lambda t: t.detach().double().cpu().numpy()
METHODS = {"extended": extended, "cubature": cubature, "gauss_hermite": gauss_hermite, "unscented": unscented}
OMETH = {"extended": O.extended, "cubature": O.cubature}


def close(a, b, tol, what):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    err = float(np.abs(a - b).max()) / max(1.0, float(np.abs(b).max()))
    assert err <= tol, f"{what}: rel err {err:.3e} > {tol:.1e}"


def ct_torch(dt):
    """Coordinated turn written with torch ops (tests/bearings/bearings_utils.py:7-44)."""
    def f(x):
        w = x[..., 4]
        c, s = torch.cos(w * dt), torch.sin(w * dt)
        small = w.abs() < 1e-6
        a = torch.where(small, torch.full_like(w, dt), s / torch.where(small, torch.ones_like(w), w))
        bb = torch.where(small, torch.zeros_like(w), (c - 1) / torch.where(small, torch.ones_like(w), w))
        return torch.stack([x[..., 0] + a * x[..., 2] - bb * x[..., 3], x[..., 1] + bb * x[..., 2] + a * x[..., 3],
                            c * x[..., 2] + s * x[..., 3], -s * x[..., 2] + c * x[..., 3], w], -1)
    return f


def bearings_torch(s1, s2):
    def f(x):
        return torch.stack([torch.atan2(x[..., 1] - s1[1], x[..., 0] - s1[0]),
                            torch.atan2(x[..., 1] - s2[1], x[..., 0] - s2[0])], -1)
    return f


@pytest.mark.parametrize("method", ["extended", "cubature", "gauss_hermite", "unscented"])
def test_linearize_torch_callable_matches_builtin(method):
    rng = np.random.RandomState(0)
    m = rng.randn(40, 5) * 0.5 + np.array([0.1, 0.2, 1.0, 0.0, 0.5])
    L = np.tril(0.1 * rng.randn(40, 5, 5)) + 0.3 * np.eye(5)
    Q, R, obs_f, trans_f = M.make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    q = MVNSqrt(t64(np.zeros(5)), t64(np.linalg.cholesky(Q)))
    r = MVNSqrt(t64(np.zeros(2)), t64(np.linalg.cholesky(R)))
    x = MVNSqrt(t64(m), t64(L))
    for dev_f, torch_f, noise in ((trans_f, ct_torch(0.01), q), (obs_f, bearings_torch((-1.5, 0.5), (1.0, 1.0)), r)):
        a = METHODS[method](FunctionalModel(dev_f, noise), x)
        b = METHODS[method](FunctionalModel(torch_f, noise), x)
        # F and b to rounding; the noise factor of the coordinated-turn regression
        # comes out of rank-one downdates that cancel (chol Q ~ 6e-5 against a
        # sigma-point spread ~ 0.3), which magnifies the last-bit differences
        # between torch's and the kernels' sin / cos / atan2 to ~1e-6
        for u, v, name, tol in zip(a, b, "FCb", (1e-12, 1e-5, 1e-12)):
            close(npy(v), npy(u), tol if method != "extended" else 1e-12, f"{method} {name}")


def test_fp32_torch_callable_small_turn_rate():
    """An FP32 problem evaluates torch callables in FP64 (and rounds the table),
    like the kernels' own models: the coordinated-turn Jacobian's turn-rate
    column, which cancels in FP32 at small w, matches the FP64 oracle entry by
    entry to 1e-5 of its size, and the regression noise of a cubature
    linearization matches it too."""
    dt = float(np.float32(0.01))
    Q, _, _, _ = M.make_parameters(0.01, 0.1, 0.5, dt, (-1.5, 0.5), (1.0, 1.0))
    cQ = np.asarray(np.linalg.cholesky(Q), np.float32).astype(np.float64)
    rng = np.random.RandomState(13)
    ws = np.array([1e-2, -1.7e-3, 4e-4, -1e-4, 1e-5, 2.1e-6])
    m = np.zeros((len(ws), 5))
    m[:, :4] = 5.0 * rng.randn(len(ws), 4)
    m[:, 4] = ws
    m = np.asarray(m, np.float32).astype(np.float64)
    L = np.repeat(0.3 * np.eye(5)[None], len(ws), 0)
    t32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device=DEV)
    model = FunctionalModel(ct_torch(dt), MVNSqrt(t32(np.zeros(5)), t32(cQ)))
    om = O.FunctionalModel(O.CoordinatedTurn(dt), O.MVNSqrt(np.zeros(5), cQ))
    F, _, _ = [npy(g) for g in extended(model, MVNSqrt(t32(m), t32(L)))]
    Fo, _, _ = O.extended(om, O.MVNSqrt(m, L))
    assert np.all(np.abs(F[:, :, 4] - Fo[:, :, 4]) <= 1e-5 * np.abs(Fo[:, :, 4]) + 1e-12)
    _, C, _ = [npy(g) for g in cubature(model, MVNSqrt(t32(m), t32(L)))]
    _, Co, _ = O.cubature(om, O.MVNSqrt(m, L))
    dg, do = np.abs(np.diagonal(C, axis1=1, axis2=2)), np.abs(np.diagonal(Co, axis1=1, axis2=2))
    assert np.all(np.abs(dg - do) <= 1e-3 * do + 1e-9), (dg, do)


@pytest.mark.parametrize("method", ["extended", "cubature"])
def test_iterated_smoothing_with_torch_callables(method):
    """The drop-in call with plain callables in FunctionalModel gives the same
    smoother as the built-in models."""
    from parsmooth.data import bearings_data
    T = 500
    _, _, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0), random_state=3)
    Q, R, obs_f, trans_f = M.make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    q = MVNSqrt(t64(np.zeros(5)), t64(np.linalg.cholesky(Q)))
    r = MVNSqrt(t64(np.zeros(2)), t64(np.linalg.cholesky(R)))
    x0 = MVNSqrt(t64([0.1, 0.2, 1.0, 0.0, 0.0]), t64(np.eye(5)))
    crit = lambda i, *_: i < 3
    ref, ell_r = parsmooth.iterated_smoothing(t64(ys), x0, FunctionalModel(trans_f, q), FunctionalModel(obs_f, r),
                                              METHODS[method], criterion=crit, return_loglikelihood=True)
    got, ell_g = parsmooth.iterated_smoothing(t64(ys), x0, FunctionalModel(ct_torch(0.01), q),
                                              FunctionalModel(bearings_torch((-1.5, 0.5), (1.0, 1.0)), r),
                                              METHODS[method], criterion=crit, return_loglikelihood=True)
    close(npy(got.mean), npy(ref.mean), 1e-10, "mean")
    close(npy(got.chol), npy(ref.chol), 1e-10, "chol")
    assert abs(float(ell_g) - float(ell_r)) <= 1e-10 * abs(float(ell_r))


class _OracleWarp:
    """x -> (x0 + 0.1 sin x1, x1 + 0.05 x0^2), with its Jacobian, for the oracle."""

    def __call__(self, x):
        x = np.asarray(x, np.float64)
        return np.stack([x[..., 0] + 0.1 * np.sin(x[..., 1]), x[..., 1] + 0.05 * x[..., 0] ** 2], -1)

    def jac(self, x):
        x = np.asarray(x, np.float64)
        J = np.zeros(x.shape[:-1] + (2, 2))
        J[..., 0, 0] = 1.0
        J[..., 0, 1] = 0.1 * np.cos(x[..., 1])
        J[..., 1, 0] = 0.1 * x[..., 0]
        J[..., 1, 1] = 1.0
        return J


def _warp_torch(x):
    return torch.stack([x[..., 0] + 0.1 * torch.sin(x[..., 1]), x[..., 1] + 0.05 * x[..., 0] ** 2], -1)


@pytest.mark.parametrize("method", ["extended", "cubature"])
@pytest.mark.parametrize("sqrt", [True, False])
def test_filter_smoother_custom_model_vs_oracle(method, sqrt):
    """A model with no built-in counterpart, both forms, against the oracle."""
    rng = np.random.RandomState(5)
    T = 200
    ys = np.cumsum(0.1 * rng.randn(T, 1), 0)
    Qc, Rc = 0.1 * np.eye(2), np.array([[0.3]])
    H = np.array([[1.0, 0.5]])
    nom = (0.2 * rng.randn(T + 1, 2), np.repeat(0.5 * np.eye(2)[None], T + 1, 0))
    to = O.FunctionalModel(_OracleWarp(), O.MVNSqrt(np.zeros(2), Qc))
    oo = O.FunctionalModel(O.Linear(H), O.MVNSqrt(np.zeros(1), Rc))
    x0 = O.MVNSqrt(np.zeros(2), np.eye(2))
    if not sqrt:
        to = O.FunctionalModel(to.function, O.MVNStandard(np.zeros(2), Qc @ Qc.T))
        oo = O.FunctionalModel(oo.function, O.MVNStandard(np.zeros(1), Rc @ Rc.T))
        x0 = O.MVNStandard(np.zeros(2), np.eye(2))
        nom = (nom[0], nom[1] @ np.swapaxes(nom[1], -1, -2))
    NomC = O.MVNSqrt if sqrt else O.MVNStandard
    ref = O.filter_smoother(ys, x0, to, oo, OMETH[method], NomC(*nom), True)
    Dv = MVNSqrt if sqrt else MVNStandard
    td = FunctionalModel(_warp_torch, Dv(t64(np.zeros(2)), t64(to.mvn[1])))
    od = FunctionalModel(M.LinearFunction(H), Dv(t64(np.zeros(1)), t64(oo.mvn[1])))
    got = parsmooth.filter_smoother(t64(ys), Dv(t64(x0[0]), t64(x0[1])), td, od, METHODS[method],
                                    Dv(t64(nom[0]), t64(nom[1])))
    close(npy(got[0]), ref[0], 1e-9, "mean")
    G = npy(got[1])
    if sqrt:
        G = G @ np.swapaxes(G, -1, -2)
        Rf = ref[1] @ np.swapaxes(ref[1], -1, -2)
    else:
        Rf = ref[1]
    close(G, Rf, 1e-9, "cov")
