"""Parity of the HIP path with the FP64 oracle at the larger compiled sizes
(state dimension 6-8, observation dimension 4): operators, the linear
filter/smoother in both forms, every scan schedule family on a ragged batch,
iterated cubature and extended smoothing of a nonlinear model, the
Gauss-Hermite, unscented and cubature regressions, FP32, the sequential passes,
the pathwise sampler and the reverse-mode gradient.

The reference takes its sizes from the array shapes
(parsmooth/parallel/_operators.py:43-77, parallel/_filtering.py:13-61); the
kernels are compiled for nx 1..8 and ny 1..4.  Tolerances: FP64 1e-9 relative
to max(1, |reference|) for one pass (1e-8 after iterations, as in
test_gpu_parity), FP32 1e-4 for means and 1e-3 for covariances against the
oracle on the same FP32-rounded inputs.
"""
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
from parsmooth import _native as N  # noqa: E402
from parsmooth.linearization import cubature, extended  # noqa: E402
from parsmooth.parallel import _operators as ops  # noqa: E402

DEV = torch.device("cuda", 0)
TOL64 = 1e-9


def t64(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=DEV)


def t32(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=DEV)


def npy(x):
    return x.detach().cpu().numpy().astype(np.float64)


def r32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def gram(L):
    L = np.asarray(L)
    return L @ np.swapaxes(L, -1, -2)


def close(a, b, tol=TOL64, what=""):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    err = float(np.abs(a - b).max()) / max(1.0, float(np.abs(b).max()))
    assert err <= tol, f"{what}: rel err {err:.3e} > {tol:.1e}"


def rand_lower(n, rng):
    L = rng.rand(n, n)
    L[np.triu_indices(n, 1)] = 0.0
    return L


def _linear_problem(nx, ny, seed, T):
    """Stable linear-Gaussian system (F = 0.98 x an orthogonal matrix) and its data."""
    np.random.seed(seed)
    x0, cx0, _, Q, cholQ, b, _ = O.get_system(nx, nx)
    F = 0.98 * np.linalg.qr(np.random.randn(nx, nx))[0]
    _, _, H, R, cholR, c, _ = O.get_system(nx, ny)
    H = H + 0.3 * np.random.randn(ny, nx)
    _, ys = O.lgssm_data(x0.mean, F, H, R, Q, b, c, T, random_state=seed)
    return x0, cx0, F, Q, cholQ, b, H, R, cholR, c, np.asarray(ys, np.float64)


@pytest.mark.parametrize("dim", [6, 7, 8])
@pytest.mark.parametrize("sqrt", [True, False])
def test_operators_large(dim, sqrt):
    """Filtering and smoothing operators (parallel/_operators.py) at nx 6-8."""
    rng = np.random.RandomState(dim)
    B = 9

    def felem():
        U = np.stack([rand_lower(dim, rng) for _ in range(B)])
        Z = np.stack([rand_lower(dim, rng) for _ in range(B)])
        e = [rng.randn(B, dim, dim) / np.sqrt(dim), rng.randn(B, dim), U, rng.randn(B, dim), Z]
        if not sqrt:
            e[2], e[4] = gram(U), gram(Z)
        return e

    e1, e2 = felem(), felem()
    ref = (O.sqrt_filtering_operator if sqrt else O.standard_filtering_operator)(tuple(e1), tuple(e2))
    got = (ops.sqrt_filtering_operator if sqrt else ops.standard_filtering_operator)(
        tuple(t64(x) for x in e1), tuple(t64(x) for x in e2))
    got = [npy(x) for x in got]
    for k, name in ((0, "A"), (1, "b"), (3, "eta")):
        close(got[k], ref[k], what=f"filtering {name}")
    if sqrt:
        close(gram(got[2]), gram(ref[2]), what="UU^T")
        close(gram(got[4]), gram(ref[4]), what="ZZ^T")
    else:
        close(got[2], ref[2], what="C")
        close(got[4], ref[4], what="J")

    def selem():
        D = np.stack([rand_lower(dim, rng) for _ in range(B)])
        return [rng.randn(B, dim), rng.randn(B, dim, dim) / np.sqrt(dim), D if sqrt else gram(D)]

    s1, s2 = selem(), selem()
    ref = (O.sqrt_smoothing_operator if sqrt else O.standard_smoothing_operator)(tuple(s1), tuple(s2))
    got = (ops.sqrt_smoothing_operator if sqrt else ops.standard_smoothing_operator)(
        tuple(t64(x) for x in s1), tuple(t64(x) for x in s2))
    got = [npy(x) for x in got]
    close(got[0], ref[0], what="smoothing g")
    close(got[1], ref[1], what="smoothing E")
    close(gram(got[2]) if sqrt else got[2], gram(ref[2]) if sqrt else ref[2], what="smoothing D")


@pytest.mark.parametrize("nx,ny", [(6, 4), (7, 3), (8, 4), (8, 1), (2, 4), (5, 4)])
@pytest.mark.parametrize("sqrt", [True, False])
@pytest.mark.parametrize("T", [1, 10, 200])
def test_filter_smoother_linear_large(nx, ny, sqrt, T):
    x0, cx0, F, Q, cholQ, b, H, R, cholR, c, ys = _linear_problem(nx, ny, 1, T)
    if sqrt:
        xo = cx0
        to, oo = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cholQ)), O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cholR))
        cls, qn, rn = MVNSqrt, (b, cholQ), (c, cholR)
    else:
        xo = x0
        to, oo = O.FunctionalModel(O.Linear(F), O.MVNStandard(b, Q)), O.FunctionalModel(O.Linear(H), O.MVNStandard(c, R))
        cls, qn, rn = MVNStandard, (b, Q), (c, R)
    nom = np.random.RandomState(2).randn(T + 1, nx)
    nom_o = type(xo)(nom, np.repeat(np.eye(nx)[None], T + 1, 0))
    fo, ell_o = O.filtering(ys, xo, to, oo, O.extended, nom_o, True, return_loglikelihood=True)
    so = O.smoothing(to, fo, O.extended, nom_o, True)
    td = FunctionalModel(M.LinearFunction(F), cls(t64(qn[0]), t64(qn[1])))
    od = FunctionalModel(M.LinearFunction(H), cls(t64(rn[0]), t64(rn[1])))
    xd = cls(t64(xo[0]), t64(xo[1]))
    nom_d = cls(t64(nom_o[0]), t64(nom_o[1]))
    fd, ell_d = parsmooth.filtering(t64(ys), xd, td, od, extended, nom_d, return_loglikelihood=True)
    close(npy(fd[0]), fo[0], what="filtered mean")
    close(gram(npy(fd[1])) if sqrt else npy(fd[1]), gram(fo[1]) if sqrt else fo[1], what="filtered cov")
    assert abs(float(ell_d) - ell_o) <= TOL64 * max(1.0, abs(ell_o)), (float(ell_d), ell_o)
    sd = parsmooth.smoothing(td, fd, extended, nom_d)
    close(npy(sd[0]), so[0], what="smoothed mean")
    close(gram(npy(sd[1])) if sqrt else npy(sd[1]), gram(so[1]) if sqrt else so[1], what="smoothed cov")


_LENGTHS = [300, 17, 1, 900]


def _batch_case(nx, ny, seed):
    """A ragged batch of one linear model (4 sequences, the lengths above)."""
    x0, cx0, F, Q, cholQ, b, H, R, cholR, c, _ = _linear_problem(nx, ny, seed, 2)
    rng = np.random.RandomState(seed)
    m0 = x0.mean + 0.3 * rng.randn(len(_LENGTHS), nx)
    obs = [np.asarray(O.lgssm_data(m0[s], F, H, R, Q, b, c, T, random_state=seed + s)[1], np.float64)
           for s, T in enumerate(_LENGTHS)]
    return m0, cx0.chol, (F, cholQ, b, H, cholR, c), obs


@pytest.mark.parametrize("chunk", [2, 4, 5])
def test_schedules_large(chunk):
    """One filter-smoother pass of an nx = 8, ny = 4 ragged batch with the
    per-lane, half-wave, team-cooperative and quad-cooperative scan schedules,
    against the oracle per sequence."""
    from parsmooth.batch import IteratedSmoother
    nx, ny = 8, 4
    m0, L0, (F, cQ, b, H, cR, c), obs = _batch_case(nx, ny, 6)
    to = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cQ))
    oo = O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cR))
    td = FunctionalModel(M.LinearFunction(F), MVNSqrt(t64(b), t64(cQ)))
    od = FunctionalModel(M.LinearFunction(H), MVNSqrt(t64(c), t64(cR)))
    rng = np.random.RandomState(1)
    offs = np.concatenate([[0], np.cumsum(np.asarray(_LENGTHS) + 1)])
    nom = (np.repeat(m0, np.asarray(_LENGTHS) + 1, 0) + 0.05 * rng.randn(offs[-1], nx),
           np.repeat(np.eye(nx)[None] * 0.5, offs[-1], 0))
    refs = [O.filter_smoother(y, O.MVNSqrt(m0[s], L0), to, oo, O.extended,
                              O.MVNSqrt(nom[0][offs[s]:offs[s + 1]], nom[1][offs[s]:offs[s + 1]]), True)
            for s, y in enumerate(obs)]
    half, quad = N.SCAN_HALF, N.SCAN_TEAM | N.SCAN_QUAD
    for k in (2, 4, 2 | N.SCAN_TEAM, 2 | half, 2 | N.SCAN_TEAM | half, 2 | quad, 2 | quad | half, 4 | quad):
        r = IteratedSmoother(_LENGTHS, td, od, extended, dtype=torch.float64)
        r.prob.set_schedule(chunk, k)
        r.set_data(t64(np.concatenate(obs)), t64(m0), t64(np.repeat(L0[None], len(_LENGTHS), 0)))
        assert r.chunk_size() == chunk
        out = r.prob.new_traj()
        r.pass_once((t64(nom[0]), t64(nom[1])), out)
        got_m, got_L = r.split(out[0]), r.split(out[1])
        for s, ref in enumerate(refs):
            close(npy(got_m[s]), ref.mean, what=f"chunk {chunk} K {k:#x} seq {s} mean")
            close(gram(npy(got_L[s])), gram(ref.chol), what=f"chunk {chunk} K {k:#x} seq {s} cov")
        del r, out


@pytest.mark.parametrize("sqrt", [True, False])
def test_iterated_cubature_large(sqrt):
    """Three iterated cubature smoothing passes of x' = sin(x) + q at nx = 6 seen
    through a linear ny = 4 observation (methods.py:iterated_smoothing)."""
    nx, ny, T = 6, 4, 150
    rng = np.random.RandomState(11)
    cQ = 0.2 * np.eye(nx) + 0.05 * rand_lower(nx, rng)
    cR = 0.1 * np.eye(ny) + 0.02 * rand_lower(ny, rng)
    H = rng.randn(ny, nx) / np.sqrt(nx)
    x = rng.randn(nx)
    ys = []
    for _ in range(T):
        x = np.sin(x) + cQ @ rng.randn(nx)
        ys.append(H @ x + cR @ rng.randn(ny))
    ys = np.asarray(ys)
    m0, L0 = np.zeros(nx), np.eye(nx)
    if sqrt:
        to = O.FunctionalModel(O.Elementwise("sin"), O.MVNSqrt(np.zeros(nx), cQ))
        oo = O.FunctionalModel(O.Linear(H), O.MVNSqrt(np.zeros(ny), cR))
        x0o, cls, q, r = O.MVNSqrt(m0, L0), MVNSqrt, cQ, cR
    else:
        to = O.FunctionalModel(O.Elementwise("sin"), O.MVNStandard(np.zeros(nx), gram(cQ)))
        oo = O.FunctionalModel(O.Linear(H), O.MVNStandard(np.zeros(ny), gram(cR)))
        x0o, cls, q, r = O.MVNStandard(m0, L0), MVNStandard, gram(cQ), gram(cR)
    ref, ell_o = O.iterated_smoothing(ys, x0o, to, oo, O.cubature, None, True,
                                      criterion=lambda i, *_: i < 3, return_loglikelihood=True)
    td = FunctionalModel(M.Elementwise("sin", nx), cls(t64(np.zeros(nx)), t64(q)))
    od = FunctionalModel(M.LinearFunction(H), cls(t64(np.zeros(ny)), t64(r)))
    got, ell_d = parsmooth.iterated_smoothing(t64(ys), cls(t64(m0), t64(L0)), td, od, cubature,
                                              criterion=lambda i, *_: i < 3, return_loglikelihood=True)
    close(npy(got[0]), ref[0], tol=1e-8, what="mean")
    close(gram(npy(got[1])) if sqrt else npy(got[1]), gram(ref[1]) if sqrt else ref[1], tol=1e-8, what="cov")
    assert abs(float(ell_d) - ell_o) <= 1e-8 * max(1.0, abs(ell_o)), (float(ell_d), ell_o)


@pytest.mark.parametrize("k", [0, 2 | N.SCAN_TEAM | N.SCAN_QUAD | N.SCAN_HALF])
def test_fp32_large(k):
    """FP32 square-root filter-smoother pass at nx = 8, ny = 4 against the FP64
    oracle on the same FP32-rounded inputs."""
    from parsmooth.batch import IteratedSmoother
    nx, ny = 8, 4
    m0, L0, (F, cQ, b, H, cR, c), obs = _batch_case(nx, ny, 9)
    m0, L0, F, cQ, b, H, cR, c = (r32(a) for a in (m0, L0, F, cQ, b, H, cR, c))
    obs = [r32(y) for y in obs]
    to = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cQ))
    oo = O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cR))
    td = FunctionalModel(M.LinearFunction(t32(F)), MVNSqrt(t32(b), t32(cQ)))
    od = FunctionalModel(M.LinearFunction(t32(H)), MVNSqrt(t32(c), t32(cR)))
    rng = np.random.RandomState(2)
    offs = np.concatenate([[0], np.cumsum(np.asarray(_LENGTHS) + 1)])
    nom = (r32(np.repeat(m0, np.asarray(_LENGTHS) + 1, 0) + 0.05 * rng.randn(offs[-1], nx)),
           r32(np.repeat(np.eye(nx)[None] * 0.5, offs[-1], 0)))
    r = IteratedSmoother(_LENGTHS, td, od, extended, dtype=torch.float32)
    if k:
        r.prob.set_schedule(4, k)
    r.set_data(t32(np.concatenate(obs)), t32(m0), t32(np.repeat(L0[None], len(_LENGTHS), 0)))
    out = r.prob.new_traj()
    r.pass_once((t32(nom[0]), t32(nom[1])), out)
    got_m, got_L = r.split(out[0]), r.split(out[1])
    for s, y in enumerate(obs):
        ref = O.filter_smoother(y, O.MVNSqrt(m0[s], L0), to, oo, O.extended,
                                O.MVNSqrt(nom[0][offs[s]:offs[s + 1]], nom[1][offs[s]:offs[s + 1]]), True)
        close(npy(got_m[s]), ref.mean, tol=1e-4, what=f"fp32 seq {s} mean")
        close(gram(npy(got_L[s])), gram(ref.chol), tol=1e-3, what=f"fp32 seq {s} cov")


@pytest.mark.parametrize("sqrt", [True, False])
def test_sequential_large(sqrt):
    """The sequential filter and smoother (parallel=False) at nx = 7, ny = 4."""
    nx, ny, T = 7, 4, 120
    x0, cx0, F, Q, cholQ, b, H, R, cholR, c, ys = _linear_problem(nx, ny, 4, T)
    if sqrt:
        xo = cx0
        to, oo = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cholQ)), O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cholR))
        cls, qn, rn = MVNSqrt, (b, cholQ), (c, cholR)
    else:
        xo = x0
        to, oo = O.FunctionalModel(O.Linear(F), O.MVNStandard(b, Q)), O.FunctionalModel(O.Linear(H), O.MVNStandard(c, R))
        cls, qn, rn = MVNStandard, (b, Q), (c, R)
    fo = O.filtering(ys, xo, to, oo, O.extended, None, False)
    so = O.smoothing(to, fo, O.extended, None, False)
    td = FunctionalModel(M.LinearFunction(F), cls(t64(qn[0]), t64(qn[1])))
    od = FunctionalModel(M.LinearFunction(H), cls(t64(rn[0]), t64(rn[1])))
    fd = parsmooth.filtering(t64(ys), cls(t64(xo[0]), t64(xo[1])), td, od, extended, parallel=False)
    close(npy(fd[0]), fo[0], what="sequential filtered mean")
    close(gram(npy(fd[1])) if sqrt else npy(fd[1]), gram(fo[1]) if sqrt else fo[1], what="sequential filtered cov")
    sd = parsmooth.smoothing(td, fd, extended, parallel=False)
    close(npy(sd[0]), so[0], what="sequential smoothed mean")
    close(gram(npy(sd[1])) if sqrt else npy(sd[1]), gram(so[1]) if sqrt else so[1], what="sequential smoothed cov")


def test_sampling_large():
    """Pathwise sampler (methods.py:79-91) at nx = 8 with the oracle's normals."""
    nx, ny, T, S = 8, 4, 200, 40
    x0, cx0, F, Q, cholQ, b, H, R, cholR, c, ys = _linear_problem(nx, ny, 7, T)
    to = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cholQ))
    oo = O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cholR))
    filt = O.filtering(ys, cx0, to, oo, O.extended, None, True)
    smooth = O.smoothing(to, filt, O.extended, None, True)
    eps = np.random.RandomState(3).randn(T + 1, S, nx)
    ref = O.sampling(eps, to, filt, O.extended, smooth)
    td = FunctionalModel(M.LinearFunction(F), MVNSqrt(t64(b), t64(cholQ)))
    got = parsmooth.sampling(t64(eps), S, td, MVNSqrt(t64(filt.mean), t64(filt.chol)), extended,
                             MVNSqrt(t64(smooth.mean), t64(smooth.chol)))
    assert got.shape == (T + 1, S, nx)
    close(npy(got), ref, 1e-9, "samples")


def _grad_build(theta, nx=6, ny=4):
    """theta = (log q, log r, a): F = 0.9 I + a S (S the shift), Q = q I,
    R = r I, H = the first ny rows of I."""
    q, r, a = torch.exp(theta[0]), torch.exp(theta[1]), theta[2]
    eye = torch.eye(nx, dtype=theta.dtype)
    F = 0.9 * eye + a * torch.diag(torch.ones(nx - 1, dtype=theta.dtype), 1)
    return {"trans_params": F.reshape(-1), "trans_q_mean": torch.zeros(nx, dtype=theta.dtype),
            "trans_q_chol": torch.sqrt(q) * eye, "obs_params": eye[:ny].reshape(-1),
            "obs_q_mean": torch.zeros(ny, dtype=theta.dtype),
            "obs_q_chol": torch.sqrt(r) * torch.eye(ny, dtype=theta.dtype),
            "m0": torch.zeros(1, nx, dtype=theta.dtype), "L0": eye[None]}


def _oracle_grad_ell(theta, ys, nx=6, ny=4):
    v = {k: t.detach().numpy() for k, t in _grad_build(torch.as_tensor(theta)).items()}
    tm = O.FunctionalModel(O.Linear(v["trans_params"].reshape(nx, nx)), O.MVNSqrt(v["trans_q_mean"], v["trans_q_chol"]))
    om = O.FunctionalModel(O.Linear(v["obs_params"].reshape(ny, nx)), O.MVNSqrt(v["obs_q_mean"], v["obs_q_chol"]))
    _, ell = O.filtering(ys, O.MVNSqrt(v["m0"][0], v["L0"][0]), tm, om, O.extended, None, True,
                         return_loglikelihood=True)
    return ell


def test_reverse_gradient_large():
    """Reverse-mode gradient of the log-likelihood at nx = 6, ny = 4 against
    central differences of the oracle."""
    from parsmooth.gradient import IteratedLoglikelihood
    nx, ny, T = 6, 4, 150
    theta = np.array([np.log(0.05), np.log(0.3), 0.1])
    v = {k: t.numpy() for k, t in _grad_build(torch.as_tensor(theta)).items()}
    Fm, Hm = v["trans_params"].reshape(nx, nx), v["obs_params"].reshape(ny, nx)
    _, ys = O.lgssm_data(np.zeros(nx), Fm, Hm, v["obs_q_chol"] @ v["obs_q_chol"].T,
                         v["trans_q_chol"] @ v["trans_q_chol"].T, np.zeros(nx), np.zeros(ny), T,
                         random_state=np.random.RandomState(0))
    ys = np.asarray(ys, np.float64)
    obj = IteratedLoglikelihood(_grad_build, N.MODEL_LINEAR, N.MODEL_LINEAR, nx, ny, [T], ys, extended, 2,
                                (np.zeros((T + 1, nx)), np.tile(np.eye(nx), (T + 1, 1, 1))))
    ell, grad = obj.value_and_grad(torch.as_tensor(theta))
    ell_o = _oracle_grad_ell(theta, ys)
    assert float(ell[0]) == pytest.approx(ell_o, rel=1e-10)
    h = 1e-5
    fd = np.array([(_oracle_grad_ell(theta + h * e, ys) - _oracle_grad_ell(theta - h * e, ys)) / (2 * h)
                   for e in np.eye(3)])
    np.testing.assert_allclose(grad[0].cpu().numpy(), fd, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("sqrt", [True, False])
def test_iterated_extended_large(sqrt):
    """Three extended (IEKS) passes of x' = sin(x) + q at nx = 8 through a
    linear ny = 4 observation."""
    from parsmooth.linearization import extended as ext
    nx, ny, T = 8, 4, 120
    rng = np.random.RandomState(21)
    cQ = 0.2 * np.eye(nx) + 0.05 * rand_lower(nx, rng)
    cR = 0.1 * np.eye(ny) + 0.02 * rand_lower(ny, rng)
    H = rng.randn(ny, nx) / np.sqrt(nx)
    x = rng.randn(nx)
    ys = []
    for _ in range(T):
        x = np.sin(x) + cQ @ rng.randn(nx)
        ys.append(H @ x + cR @ rng.randn(ny))
    ys = np.asarray(ys)
    m0, L0 = np.zeros(nx), np.eye(nx)
    cls_o, cls = (O.MVNSqrt, MVNSqrt) if sqrt else (O.MVNStandard, MVNStandard)
    q, r = (cQ, cR) if sqrt else (gram(cQ), gram(cR))
    to = O.FunctionalModel(O.Elementwise("sin"), cls_o(np.zeros(nx), q))
    oo = O.FunctionalModel(O.Linear(H), cls_o(np.zeros(ny), r))
    ref = O.iterated_smoothing(ys, cls_o(m0, L0), to, oo, O.extended, None, True, criterion=lambda i, *_: i < 3)
    td = FunctionalModel(M.Elementwise("sin", nx), cls(t64(np.zeros(nx)), t64(q)))
    od = FunctionalModel(M.LinearFunction(H), cls(t64(np.zeros(ny)), t64(r)))
    got = parsmooth.iterated_smoothing(t64(ys), cls(t64(m0), t64(L0)), td, od, ext, criterion=lambda i, *_: i < 3)
    close(npy(got[0]), ref[0], tol=1e-8, what="mean")
    close(gram(npy(got[1])) if sqrt else npy(got[1]), gram(ref[1]) if sqrt else ref[1], tol=1e-8, what="cov")


@pytest.mark.parametrize("method,nx", [("gauss_hermite", 6), ("unscented", 7), ("cubature", 8)])
def test_sigma_point_linearize_large(method, nx):
    """The sigma-point regressions at the larger sizes (Gauss-Hermite order 3:
    729 points at nx = 6) against the oracle, on an elementwise model."""
    from parsmooth.linearization import cubature as cub, gauss_hermite, unscented
    dm = {"gauss_hermite": gauss_hermite, "unscented": unscented, "cubature": cub}[method]
    om = {"gauss_hermite": O.gauss_hermite, "unscented": O.unscented, "cubature": O.cubature}[method]
    rng = np.random.RandomState(nx)
    B = 4
    m = rng.randn(B, nx)
    L = np.stack([0.3 * rand_lower(nx, rng) + 0.1 * np.eye(nx) for _ in range(B)])
    qm, qc = rng.randn(nx), 0.1 * rand_lower(nx, rng) + 0.05 * np.eye(nx)
    ref = om(O.FunctionalModel(O.Elementwise("arctan"), O.MVNSqrt(qm, qc)), O.MVNSqrt(m, L))
    got = [npy(g) for g in dm(FunctionalModel(M.Elementwise("arctan", nx), MVNSqrt(t64(qm), t64(qc))),
                              MVNSqrt(t64(m), t64(L)))]
    close(got[0], ref[0], tol=1e-8, what=f"{method} F")
    close(got[2], ref[2], tol=1e-8, what=f"{method} b")
    close(gram(got[1]), gram(ref[1]), tol=1e-8, what=f"{method} Q")


@pytest.mark.parametrize("nx,ny", [(4, 2), (8, 4)])
def test_noise_factor_any_square_root(nx, ny):
    """Noise factors given as any square root of Q and R (a lower factor times an
    orthogonal matrix, so with entries above the diagonal): the linearized
    layout keeps lower factors, so the kernels triangularise such a factor
    first; the result matches the oracle, which takes any square root."""
    T = 50
    x0, cx0, F, Q, cholQ, b, H, R, cholR, c, ys = _linear_problem(nx, ny, 8, T)
    rng = np.random.RandomState(9)
    sQ = cholQ @ np.linalg.qr(rng.randn(nx, nx))[0]
    sR = cholR @ np.linalg.qr(rng.randn(ny, ny))[0]
    assert np.abs(np.triu(sQ, 1)).max() > 0.1 and np.allclose(sQ @ sQ.T, Q)
    to = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, sQ))
    oo = O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, sR))
    fo, ell_o = O.filtering(ys, cx0, to, oo, O.extended, None, True, return_loglikelihood=True)
    so = O.smoothing(to, fo, O.extended, None, True)
    td = FunctionalModel(M.LinearFunction(F), MVNSqrt(t64(b), t64(sQ)))
    od = FunctionalModel(M.LinearFunction(H), MVNSqrt(t64(c), t64(sR)))
    fd, ell_d = parsmooth.filtering(t64(ys), MVNSqrt(t64(cx0.mean), t64(cx0.chol)), td, od, extended,
                                    return_loglikelihood=True)
    close(npy(fd.mean), fo.mean, what="filtered mean")
    close(gram(npy(fd.chol)), gram(fo.chol), what="filtered cov")
    assert abs(float(ell_d) - ell_o) <= TOL64 * max(1.0, abs(ell_o))
    sd = parsmooth.smoothing(td, fd, extended)
    close(npy(sd.mean), so.mean, what="smoothed mean")
    close(gram(npy(sd.chol)), gram(so.chol), what="smoothed cov")
