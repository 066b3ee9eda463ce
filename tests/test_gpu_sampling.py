"""Pathwise sampler (psm_sample) against the oracle's restatement of the
reference's sampler, with the same standard normals."""
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
from parsmooth.linearization import cubature, extended  # noqa: E402

DEV = torch.device("cuda", 0)
t64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=DEV)
npy = lambda t: t.detach().double().cpu().numpy()
gram = lambda L: L @ np.swapaxes(L, -1, -2)


def close(a, b, tol, what):
    err = float(np.abs(np.asarray(a) - np.asarray(b)).max()) / max(1.0, float(np.abs(b).max()))
    assert err <= tol, f"{what}: rel err {err:.3e} > {tol:.1e}"


@pytest.mark.parametrize("method", ["extended", "cubature"])
def test_sampling_bearings_matches_oracle(method):
    from parsmooth.data import bearings_data
    T, S = 700, 33   # several sampling chunks, a sample count off the wave size
    _, _, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0), random_state=4)
    ys = ys.astype(np.float64)
    Q, R, obs_f, trans_f = M.make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    cQ, cR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    to = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNSqrt(np.zeros(5), cQ))
    oo = O.FunctionalModel(O.Bearings((-1.5, 0.5), (1.0, 1.0)), O.MVNSqrt(np.zeros(2), cR))
    om = {"extended": O.extended, "cubature": O.cubature}[method]
    x0 = O.MVNSqrt(np.array([0.1, 0.2, 1.0, 0.0, 0.0]), np.eye(5))
    nom = O.iterated_smoothing(ys, x0, to, oo, om, None, True, criterion=lambda i, *_: i < 2)
    filt = O.filtering(ys, x0, to, oo, om, nom, True)
    eps = np.random.RandomState(0).randn(T + 1, S, 5)
    ref = O.sampling(eps, to, filt, om, nom)
    td = FunctionalModel(trans_f, MVNSqrt(t64(np.zeros(5)), t64(cQ)))
    got = parsmooth.sampling(t64(eps), S, td, MVNSqrt(t64(filt.mean), t64(filt.chol)),
                             {"extended": extended, "cubature": cubature}[method],
                             MVNSqrt(t64(nom.mean), t64(nom.chol)))
    assert got.shape == (T + 1, S, 5)
    # a sample uses the noise factor itself (not only its Gram matrix); the
    # cubature regression of the coordinated turn forms that factor by
    # cancelling rank-one downdates (chol Q ~ 6e-5 against the sigma-point
    # spread), where the kernel's and the oracle's FP64 operation orders part
    # at ~1e-8 relative
    close(npy(got), ref, 1e-9 if method == "extended" else 1e-7, "samples")


def test_sampling_covariance_form_linear_matches_oracle():
    np.random.seed(5)
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(3, 3)
    _, _, H, R, cholR, c, _ = O.get_system(3, 2)
    F = 0.9 * F / max(1.0, np.abs(np.linalg.eigvals(F)).max())
    T, S = 300, 64
    ys = np.random.RandomState(2).randn(T, 2)
    tq = O.FunctionalModel(O.Linear(F), O.MVNStandard(b, Q))
    to = O.FunctionalModel(O.Linear(H), O.MVNStandard(c, R))
    filt = O.filtering(ys, x0, tq, to, O.extended, None, True)
    smooth = O.smoothing(tq, filt, O.extended, None, True)
    eps = np.random.RandomState(3).randn(T + 1, S, 3)
    ref = O.sampling(eps, tq, filt, O.extended, smooth)
    td = FunctionalModel(M.LinearFunction(F), MVNStandard(t64(b), t64(Q)))
    got = parsmooth.sampling(t64(eps), S, td, MVNStandard(t64(filt.mean), t64(filt.cov)), extended,
                             MVNStandard(t64(smooth.mean), t64(smooth.cov)))
    close(npy(got), ref, 1e-9, "samples")


def test_sampling_moments_and_default_nominal():
    """Seeded sampling with the reference's default nominal (the smoothed
    trajectory): sample moments are the smoothed marginals."""
    np.random.seed(8)
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(2, 2)
    _, _, H, R, cholR, c, _ = O.get_system(2, 1)
    F = 0.9 * F / max(1.0, np.abs(np.linalg.eigvals(F)).max())
    T, S = 100, 20000
    ys = np.random.RandomState(4).randn(T, 1)
    tq = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cholQ))
    to = O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cholR))
    filt = O.filtering(ys, cx0, tq, to, O.extended, None, True)
    smooth = O.smoothing(tq, filt, O.extended, None, True)
    td = FunctionalModel(M.LinearFunction(F), MVNSqrt(t64(b), t64(cholQ)))
    xs = npy(parsmooth.sampling(1234, S, td, MVNSqrt(t64(filt.mean), t64(filt.chol)), extended))
    P = gram(smooth.chol)
    se = np.sqrt(np.diagonal(P, axis1=1, axis2=2) / S)
    assert np.all(np.abs(xs.mean(1) - smooth.mean) <= 5 * se)
    d = xs - xs.mean(1, keepdims=True)
    C = np.einsum("tsi,tsj->tij", d, d) / (S - 1)
    np.testing.assert_allclose(C, P, rtol=0.05, atol=0.05 * np.abs(P).max())


def test_sampling_key_forms_agree():
    """The key may be normals, a seed or a torch.Generator on any device: a CPU
    generator draws on the CPU and gives exactly the samples of the same normals
    passed as a tensor."""
    np.random.seed(9)
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(2, 2)
    _, _, H, R, cholR, c, _ = O.get_system(2, 1)
    F = 0.9 * F / max(1.0, np.abs(np.linalg.eigvals(F)).max())
    T, S = 50, 7
    ys = np.random.RandomState(5).randn(T, 1)
    tq = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cholQ))
    to = O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cholR))
    filt = O.filtering(ys, cx0, tq, to, O.extended, None, True)
    td = FunctionalModel(M.LinearFunction(F), MVNSqrt(t64(b), t64(cholQ)))
    ft = MVNSqrt(t64(filt.mean), t64(filt.chol))
    eps = torch.randn((T + 1, S, 2), generator=torch.Generator().manual_seed(11), dtype=torch.float64)
    a = npy(parsmooth.sampling(torch.Generator().manual_seed(11), S, td, ft, extended))
    b_ = npy(parsmooth.sampling(eps.to(DEV), S, td, ft, extended))
    np.testing.assert_array_equal(a, b_)
    s1 = npy(parsmooth.sampling(77, S, td, ft, extended))
    s2 = npy(parsmooth.sampling(77, S, td, ft, extended))
    np.testing.assert_array_equal(s1, s2)
