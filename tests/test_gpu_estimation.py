"""Parameter estimation driver: L-BFGS-B on the device gradient reaches the
maximum-likelihood estimate found by a derivative-free search on the FP64
oracle's log-likelihood."""
import numpy as np
import pytest
import torch
from scipy.optimize import minimize

from oracle import psm_oracle as O

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

from parsmooth import _native as N  # noqa: E402
from parsmooth.estimation import fit  # noqa: E402
from parsmooth.gradient import IteratedLoglikelihood  # noqa: E402
from parsmooth.linearization import extended  # noqa: E402

F = np.array([[1.0, 0.1], [0.0, 1.0]])
H = np.array([[1.0, 0.0]])


def _build(theta):
    """theta = (log q, log r): Q = q I, R = r^2."""
    q, r = torch.exp(theta[0]), torch.exp(theta[1])
    return {"trans_params": torch.tensor(F.reshape(-1), dtype=theta.dtype),
            "trans_q_mean": torch.zeros(2, dtype=theta.dtype),
            "trans_q_chol": torch.sqrt(q) * torch.eye(2, dtype=theta.dtype),
            "obs_params": torch.tensor(H.reshape(-1), dtype=theta.dtype),
            "obs_q_mean": torch.zeros(1, dtype=theta.dtype), "obs_q_chol": r.reshape(1, 1),
            "m0": torch.zeros(1, 2, dtype=theta.dtype), "L0": torch.eye(2, dtype=theta.dtype)[None]}


def _oracle_nll(theta, ys):
    q, r = np.exp(theta)
    tm = O.FunctionalModel(O.Linear(F), O.MVNSqrt(np.zeros(2), np.sqrt(q) * np.eye(2)))
    om = O.FunctionalModel(O.Linear(H), O.MVNSqrt(np.zeros(1), np.array([[r]])))
    _, ell = O.filtering(ys, O.MVNSqrt(np.zeros(2), np.eye(2)), tm, om, O.extended, None, True,
                         return_loglikelihood=True)
    return -ell


def test_lbfgsb_reaches_the_oracle_mle():
    T = 400
    _, ys = O.lgssm_data(np.zeros(2), F, H, np.array([[0.3 ** 2]]), 0.05 * np.eye(2), np.zeros(2), np.zeros(1), T,
                         random_state=np.random.RandomState(0))
    ys = ys.astype(np.float64)
    obj = IteratedLoglikelihood(_build, N.MODEL_LINEAR, N.MODEL_LINEAR, 2, 1, [T], ys, extended, 2,
                                (np.zeros((T + 1, 2)), np.tile(np.eye(2), (T + 1, 1, 1))))
    res = fit(obj, np.array([0.0, 0.0]), bounds=[(-8.0, 3.0), (-8.0, 3.0)])
    ref = minimize(_oracle_nll, np.array([0.0, 0.0]), args=(ys,), method="Nelder-Mead",
                   options={"xatol": 1e-7, "fatol": 1e-10, "maxiter": 2000})
    assert res.success
    np.testing.assert_allclose(res.x, ref.x, atol=1e-4)
    assert abs(res.fun - ref.fun) <= 1e-7 * abs(ref.fun)
