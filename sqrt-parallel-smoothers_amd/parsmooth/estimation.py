"""Parameter estimation by maximising the pseudo-log-likelihood of the
iterated smoother, as in the reference's parameter-estimation notebooks
(notebooks/experiment_bearing_only_parameter_estimation.ipynb: jax
``value_and_grad`` of the negative log-likelihood handed to
``scipy.optimize.minimize(..., jac=True, method="L-BFGS-B", bounds=...)``).

Here every evaluation is one ``IteratedLoglikelihood.value_and_grad`` on the
device (value passes plus the tangent passes of the implicit gradient); the
optimiser runs on the host.
"""
from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np
import torch
from scipy.optimize import OptimizeResult, minimize

from .gradient import IteratedLoglikelihood


def negative_loglikelihood(objective: IteratedLoglikelihood):
    """theta (numpy) -> (-sum ell, -sum grad) in float64, scipy's jac=True form."""
    def fun(theta):
        ell, grad = objective.value_and_grad(torch.as_tensor(np.asarray(theta, dtype=np.float64)))
        return float(-ell.sum().item()), -grad.sum(0).double().cpu().numpy()
    return fun


def fit(objective: IteratedLoglikelihood, theta0, method: str = "L-BFGS-B",
        bounds: Optional[Sequence[Tuple[float, float]]] = None, tol: Optional[float] = None,
        options: Optional[dict] = None) -> OptimizeResult:
    """Maximum pseudo-likelihood estimate of theta; returns scipy's result
    (``x`` the estimate, ``fun`` the negative log-likelihood there)."""
    return minimize(negative_loglikelihood(objective), np.asarray(theta0, dtype=np.float64), jac=True,
                    method=method, bounds=bounds, tol=tol, options=options)
