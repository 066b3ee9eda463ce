"""parsmooth._utils on the device (psm_tria, psm_cholesky_update,
psm_mvn_loglikelihood) against the oracle and closed forms, after the
reference's tests/test_math_utils.py:15-52."""
import numpy as np
import pytest
import torch
from scipy.stats import multivariate_normal

from oracle import psm_oracle as O

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

from parsmooth import _native as N  # noqa: E402
from parsmooth._utils import _cholesky_update, cholesky_update_many, mvn_loglikelihood, tria  # noqa: E402

DEV = torch.device("cuda", 0)
npy = lambda t: t.detach().cpu().numpy().astype(np.float64)


def _lower(rng, n):
    L = rng.rand(n, n)
    L[np.triu_indices(n, 1)] = 0.0
    return L


@pytest.mark.parametrize("multiplier", [1.0, -0.1])
@pytest.mark.parametrize("seed", [0, 42, 666])
@pytest.mark.parametrize("dim_x", [2, 3, 10, 11])
def test_cholesky_update(multiplier, seed, dim_x):
    """tests/test_math_utils.py:15-33: L L^T + mult v v^T (rtol 1e-4, the
    reference's), and the oracle's recurrence to 1e-10."""
    rng = np.random.RandomState(seed)
    cholQ = _lower(rng, dim_x)
    v = rng.randn(dim_x)
    expected = cholQ @ cholQ.T + multiplier * np.outer(v, v)
    if np.linalg.eigvalsh(expected).min() <= 1e-6:
        pytest.skip("random vectors do not result in a positive definite matrix.")
    got = npy(_cholesky_update(torch.as_tensor(cholQ, device=DEV), torch.as_tensor(v, device=DEV), multiplier))
    np.testing.assert_allclose(got @ got.T, expected, rtol=1e-4)
    np.testing.assert_allclose(got, O.cholesky_update(cholQ, v, multiplier), rtol=1e-10, atol=1e-12)
    assert np.all(np.triu(got, 1) == 0.0)


@pytest.mark.parametrize("multiplier", [1.0, -0.1])
@pytest.mark.parametrize("seed", [0, 1, 2, 42, 666])
@pytest.mark.parametrize("dim_x", [2, 3, 10, 11])
def test_cholesky_update_many(multiplier, seed, dim_x):
    """tests/test_math_utils.py:36-52: three vectors in turn."""
    rng = np.random.RandomState(seed)
    B = 3
    cholQ = _lower(rng, dim_x)
    v = rng.rand(B, dim_x)
    expected = cholQ @ cholQ.T + multiplier * sum(np.outer(v[k], v[k]) for k in range(B))
    if np.linalg.eigvalsh(expected).min() <= 1e-6:
        pytest.skip("random vectors do not result in a positive definite matrix.")
    got = npy(cholesky_update_many(torch.as_tensor(cholQ, device=DEV), torch.as_tensor(v, device=DEV), multiplier))
    np.testing.assert_allclose(got @ got.T, expected, rtol=1e-4)
    np.testing.assert_allclose(got, O.cholesky_update_many(cholQ, v, multiplier), rtol=1e-10, atol=1e-12)


def test_cholesky_update_batched_fp32_and_indefinite():
    """A batch of 257 factors (more than one workgroup) in FP32 against the
    oracle; a downdate that leaves the matrix indefinite gives the oracle's
    zeros instead of NaNs (_utils.py:80)."""
    rng = np.random.RandomState(7)
    n, B = 5, 257
    L = np.stack([_lower(rng, n) + np.eye(n) for _ in range(B)])
    V = 0.3 * rng.randn(B, 2, n)
    got = npy(cholesky_update_many(torch.as_tensor(L, dtype=torch.float32, device=DEV),
                                   torch.as_tensor(V, dtype=torch.float32, device=DEV), 1.0))
    L32, V32 = L.astype(np.float32).astype(np.float64), V.astype(np.float32).astype(np.float64)
    ref = np.stack([O.cholesky_update_many(L32[b], V32[b], 1.0) for b in range(B)])
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5)
    Lb = np.eye(3)
    vb = np.array([2.0, 0.0, 0.0])
    got = npy(_cholesky_update(torch.as_tensor(Lb, device=DEV), torch.as_tensor(vb, device=DEV), -1.0))
    ref = O.cholesky_update(Lb, vb, -1.0)
    assert np.all(np.isfinite(got))
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("shape", [(5, 10), (10, 5), (7, 7), (1, 3), (3, 1), (4, 4, 9), (2, 3, 8, 6)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_tria(shape, dtype):
    """L L^T = A A^T, L lower with a non-negative diagonal and the oracle's
    tria (QR with the signs normalised)."""
    rng = np.random.RandomState(sum(shape))
    A = rng.randn(*shape)
    if dtype == torch.float32:
        A = A.astype(np.float32).astype(np.float64)
    got = npy(tria(torch.as_tensor(A, dtype=dtype, device=DEV)))
    r, c = shape[-2], shape[-1]
    assert got.shape == shape[:-2] + (r, min(r, c))
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    AA = A @ np.swapaxes(A, -1, -2)
    scale = max(1.0, np.abs(AA).max())
    assert np.abs(got @ np.swapaxes(got, -1, -2) - AA).max() <= tol * scale * 10
    assert np.all(np.triu(got[..., :min(r, c), :], 1) == 0.0)
    assert np.all(np.diagonal(got, axis1=-2, axis2=-1) >= 0.0)
    ref = O.tria(A)
    assert np.abs(got - ref).max() <= tol * 100 * max(1.0, np.abs(ref).max())


def test_tria_rank_deficient_and_zero():
    """Zero rows and a zero matrix stay zero; a rank-one block keeps L L^T."""
    A = np.zeros((3, 4))
    A[0] = [1.0, 2.0, 0.0, -1.0]
    A[2] = 2.0 * A[0]
    got = npy(tria(torch.as_tensor(A, device=DEV)))
    np.testing.assert_allclose(got @ got.T, A @ A.T, atol=1e-14)
    assert np.all(got[1] == 0.0)
    z = npy(tria(torch.zeros(2, 2, 5, dtype=torch.float64, device=DEV)))
    assert np.all(z == 0.0)


def test_mvn_loglikelihood():
    """Against scipy's multivariate normal, batched and broadcast; FP32 inputs
    accumulate in FP64."""
    rng = np.random.RandomState(3)
    n, B = 4, 100
    L = np.stack([_lower(rng, n) + np.eye(n) for _ in range(B)])
    x = rng.randn(B, n)
    got = npy(mvn_loglikelihood(torch.as_tensor(x, device=DEV), torch.as_tensor(L, device=DEV)))
    ref = np.array([multivariate_normal(np.zeros(n), L[b] @ L[b].T).logpdf(x[b]) for b in range(B)])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(got, O.mvn_loglikelihood(x, L), rtol=1e-12, atol=1e-12)
    one = float(mvn_loglikelihood(torch.as_tensor(x[0], device=DEV), torch.as_tensor(L[0], device=DEV)))
    assert one == pytest.approx(ref[0], rel=1e-12)
    shared = npy(mvn_loglikelihood(torch.as_tensor(x, device=DEV), torch.as_tensor(L[0], device=DEV)))
    np.testing.assert_allclose(shared, multivariate_normal(np.zeros(n), L[0] @ L[0].T).logpdf(x), rtol=1e-12)
    x32, L32 = x.astype(np.float32), L.astype(np.float32)
    got32 = npy(mvn_loglikelihood(torch.as_tensor(x32, device=DEV), torch.as_tensor(L32, device=DEV)))
    np.testing.assert_allclose(got32, O.mvn_loglikelihood(x32.astype(np.float64), L32.astype(np.float64)),
                               rtol=1e-12, atol=1e-12)


def test_utils_reject_bad_arguments():
    lib = N.lib()
    assert lib.psm_tria(N.PSM_F64, 1, 0, 3, None, None) != 0
    assert lib.psm_cholesky_update(N.PSM_F64, 2, 3, 1, None, None, 1.0, None, None) != 0
    assert lib.psm_mvn_loglikelihood(7, 1, 2, None, None, None, None, None) != 0
    with pytest.raises(ValueError):
        cholesky_update_many(torch.eye(3, device=DEV), torch.ones(2, 4, device=DEV), 1.0)
    with pytest.raises(ValueError):
        mvn_loglikelihood(torch.ones(3, device=DEV), torch.eye(4, device=DEV))
