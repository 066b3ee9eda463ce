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
