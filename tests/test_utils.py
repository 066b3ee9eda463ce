"""parsmooth._utils host helpers (CPU): fixed_point and its implicit gradient
(tests/test_math_utils.py:55-66 of the reference), the trajectory shifts, and
the reference's module paths."""
import numpy as np
import pytest
import torch

from parsmooth import MVNSqrt
from parsmooth._utils import fixed_point, none_or_concat, none_or_shift


def _my_fun(a, b, x0, n=500):
    f = lambda x: (a * x[0] + b[0],)
    return fixed_point(f, x0, lambda i, *_: i < n)


def test_fixed_point_value_and_gradient():
    """The reference's case: x = 0.7 x + 0.5 from 1, 500 applications; the
    gradient against central differences of the whole map and the closed form
    of the converged fixed point (x* = b / (1 - a))."""
    a = torch.tensor(0.7, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    x0 = (torch.tensor(1.0, dtype=torch.float64),)
    out = _my_fun(a, (b,), x0)
    assert float(out[0]) == pytest.approx(0.5 / 0.3, rel=1e-7, abs=1e-7)
    ga, gb = torch.autograd.grad(out[0], (a, b))
    assert float(ga) == pytest.approx(0.5 / 0.3 ** 2, rel=1e-9)
    assert float(gb) == pytest.approx(1.0 / 0.3, rel=1e-9)
    h = 1e-6
    with torch.no_grad():
        fd_a = (_my_fun(a + h, (b,), x0)[0] - _my_fun(a - h, (b,), x0)[0]) / (2 * h)
        fd_b = (_my_fun(a, (b + h,), x0)[0] - _my_fun(a, (b - h,), x0)[0]) / (2 * h)
    assert float(ga) == pytest.approx(float(fd_a), rel=1e-6)
    assert float(gb) == pytest.approx(float(fd_b), rel=1e-6)
    assert torch.autograd.gradcheck(lambda aa, bb: _my_fun(aa, (bb,), x0)[0], (a, b))


def test_fixed_point_iteration_counts():
    """criterion i < n applies f n times; the reverse pass iterates
    u <- x_bar + J^T u n + 1 times from u = x_bar (_utils.py:115-126), which at
    n = 3 gives u = 1 + a + a^2 + a^3 + a^4 and d/db = u, d/da = x_3 u."""
    calls = []
    a = torch.tensor(0.6, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)

    def f(x):
        calls.append(1)
        return (a * x[0] + b,)

    x0 = (torch.tensor(2.0, dtype=torch.float64),)
    out = fixed_point(f, x0, lambda i, *_: i < 3)
    assert len(calls) == 3 + 1  # three iterations, one more application for the gradient graph
    x = 2.0
    for _ in range(3):
        x = 0.6 * x + 0.5
    assert float(out[0]) == pytest.approx(x, rel=1e-15)
    ga, gb = torch.autograd.grad(out[0], (a, b))
    u = sum(0.6 ** k for k in range(5))
    assert float(gb) == pytest.approx(u, rel=1e-14)
    assert float(ga) == pytest.approx(x * u, rel=1e-14)


def test_fixed_point_initial_value_gets_no_gradient():
    a = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    x0 = torch.tensor([1.0, 2.0], dtype=torch.float64, requires_grad=True)
    out = fixed_point(lambda x: a * x + 1.0, x0, lambda i, *_: i < 20)
    ga, gx0 = torch.autograd.grad(out.sum(), (a, x0), allow_unused=True)
    assert gx0 is None or float(gx0.abs().max()) == 0.0
    assert torch.isfinite(ga)


def test_fixed_point_without_parameters_and_under_no_grad():
    out = fixed_point(lambda x: 0.5 * x + 1.0, torch.tensor(0.0, dtype=torch.float64), lambda i, *_: i < 60)
    assert float(out) == pytest.approx(2.0, rel=1e-12)
    a = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    with torch.no_grad():
        out = fixed_point(lambda x: a * x + 1.0, torch.tensor(0.0, dtype=torch.float64), lambda i, *_: i < 60)
    assert not out.requires_grad


def test_none_or_shift_and_concat():
    """parsmooth/_utils.py:84-98 on an MVNSqrt trajectory."""
    T, n = 6, 3
    traj = MVNSqrt(torch.arange(T * n, dtype=torch.float64).reshape(T, n),
                   torch.arange(T * n * n, dtype=torch.float64).reshape(T, n, n))
    assert none_or_shift(None, 1) is None
    head = none_or_shift(traj, -1)
    tail = none_or_shift(traj, 1)
    assert isinstance(head, MVNSqrt) and head.mean.shape == (T - 1, n)
    assert torch.equal(head.mean, traj.mean[:-1]) and torch.equal(tail.chol, traj.chol[1:])
    x0 = MVNSqrt(-torch.ones(n, dtype=torch.float64), -torch.ones(n, n, dtype=torch.float64))
    front = none_or_concat(tail, x0, 1)
    back = none_or_concat(head, x0, -1)
    assert none_or_concat(None, x0) is None and none_or_concat(traj, None) is None
    assert front.mean.shape == (T, n) and torch.equal(front.mean[0], x0.mean)
    assert torch.equal(front.mean[1:], traj.mean[1:])
    assert torch.equal(back.chol[-1], x0.chol) and torch.equal(back.chol[:-1], traj.chol[:-1])
    arr = none_or_concat(np.zeros((2, 2)), np.ones(2), 1)
    assert arr.shape == (3, 2) and np.all(arr[0] == 1.0)


def test_reference_module_paths():
    """The module paths the reference's tests and users import resolve to this
    package's functions."""
    from parsmooth.linearization import cubature, extended, gauss_hermite, unscented
    from parsmooth.linearization._cubature import linearize as c
    from parsmooth.linearization._extended import linearize as e
    from parsmooth.linearization._gh import linearize as g
    from parsmooth.linearization._unscented import linearize as u
    assert (c, e, g, u) == (cubature, extended, gauss_hermite, unscented)
    from parsmooth._utils import _cholesky_update, cholesky_update_many, mvn_loglikelihood, tria  # noqa: F401
    from parsmooth.parallel._filtering import _sqrt_associative_params_one, _standard_associative_params_one  # noqa: F401
    from parsmooth.parallel._operators import sqrt_filtering_operator, sqrt_smoothing_operator  # noqa: F401
    from parsmooth.parallel._smoothing import _sqrt_associative_params, _standard_associative_params  # noqa: F401
    from parsmooth.sequential._filtering import _sqrt_predict, _sqrt_update, _standard_predict, _standard_update  # noqa: F401
    from parsmooth.sequential._smoothing import _sqrt_smooth, _standard_smooth  # noqa: F401
    from parsmooth.methods import iterated_smoothing, sampling  # noqa: F401
