"""Parity of the HIP path (through the C ABI via parsmooth) with the FP64 oracle.

FP64 tolerance: 1e-9 relative to max(1, |reference|) for means, gains and
covariances (L L^T); chol factors are compared directly as well (both sides
return the non-negative-diagonal factor).  FP32 tolerances are stated per test.
"""
import functools

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
from parsmooth.linearization import cubature, extended, gauss_hermite, unscented  # noqa: E402
from parsmooth.parallel import _operators as ops  # noqa: E402
from parsmooth.parallel._filtering import _sqrt_associative_params_one, _standard_associative_params_one  # noqa: E402
from parsmooth.parallel._smoothing import _sqrt_associative_params, _standard_associative_params  # noqa: E402

DEV = torch.device("cuda", 0)
TOL64 = 1e-9
TOL32 = 1e-4  # FP32 arithmetic against the FP64 oracle (relative to the largest magnitude): means
TOL32_COV = 1e-3  # covariances L L^T of FP32 chol factors (products double the rounding)


def t64(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float64), device=DEV)


def npy(x):
    return x.detach().cpu().numpy().astype(np.float64)


def close(a, b, tol=TOL64, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.abs(b).max()) if b.size else 1.0)
    err = float(np.abs(a - b).max()) / scale if a.size else 0.0
    assert err <= tol, f"{what}: rel err {err:.3e} > {tol:.1e}"


def gram(L):
    L = np.asarray(L)
    return L @ np.swapaxes(L, -1, -2)


def rand_lower(n, rng):
    L = rng.rand(n, n)
    L[np.triu_indices(n, 1)] = 0.0
    return L


OMETH = {"extended": O.extended, "cubature": O.cubature, "gauss_hermite": O.gauss_hermite,
         "unscented": O.unscented}
DMETH = {"extended": extended, "cubature": cubature, "gauss_hermite": gauss_hermite, "unscented": unscented}


# ---------------------------------------------------------------- operators
@pytest.mark.parametrize("dim", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("sqrt", [True, False])
def test_filtering_operator(dim, seed, sqrt):
    rng = np.random.RandomState(seed)
    B = 7
    def elem():
        U = np.stack([rand_lower(dim, rng) for _ in range(B)])
        Z = np.stack([rand_lower(dim, rng) for _ in range(B)])
        e = [rng.randn(B, dim, dim), rng.randn(B, dim), U, rng.randn(B, dim), Z]
        if not sqrt:
            e[2], e[4] = gram(U), gram(Z)
        return e
    e1, e2 = elem(), elem()
    ref = (O.sqrt_filtering_operator if sqrt else O.standard_filtering_operator)(tuple(e1), tuple(e2))
    got = (ops.sqrt_filtering_operator if sqrt else ops.standard_filtering_operator)(
        tuple(t64(x) for x in e1), tuple(t64(x) for x in e2))
    got = [npy(x) for x in got]
    for k, name in enumerate("A b U eta Z".split()):
        close(got[k], ref[k], what=name)
    if sqrt:
        close(gram(got[2]), gram(ref[2]), what="UU^T")
        close(gram(got[4]), gram(ref[4]), what="ZZ^T")


@pytest.mark.parametrize("dim", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("sqrt", [True, False])
def test_smoothing_operator(dim, seed, sqrt):
    rng = np.random.RandomState(seed)
    B = 7
    def elem():
        D = np.stack([rand_lower(dim, rng) for _ in range(B)])
        return [rng.randn(B, dim), rng.randn(B, dim, dim), D if sqrt else gram(D)]
    e1, e2 = elem(), elem()
    ref = (O.sqrt_smoothing_operator if sqrt else O.standard_smoothing_operator)(tuple(e1), tuple(e2))
    got = (ops.sqrt_smoothing_operator if sqrt else ops.standard_smoothing_operator)(
        tuple(t64(x) for x in e1), tuple(t64(x) for x in e2))
    got = [npy(x) for x in got]
    for k, name in enumerate("g E D".split()):
        close(got[k], ref[k], what=name)


# ---------------------------------------------------------------- elements
def _system(dim_x, dim_y, seed):
    np.random.seed(seed)
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    _, _, H, R, cholR, c, y = O.get_system(dim_x, dim_y)
    return x0, cx0, F, Q, cholQ, b, H, R, cholR, c, y


@pytest.mark.parametrize("dim_x", [1, 2, 3])
@pytest.mark.parametrize("dim_y", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("method", ["extended", "cubature"])
def test_filter_element(dim_x, dim_y, seed, method):
    """tests/test_parallel_filter.py:26-73, against the oracle."""
    x0, cx0, F, Q, cholQ, b, H, R, cholR, c, y = _system(dim_x, dim_y, seed)
    m_nom = np.random.randn(dim_x)
    for sqrt in (True, False):
        cls_o, cls_d = (O.MVNSqrt, MVNSqrt) if sqrt else (O.MVNStandard, MVNStandard)
        qo, ro = (O.MVNSqrt(b, cholQ), O.MVNSqrt(c, cholR)) if sqrt else (O.MVNStandard(b, Q), O.MVNStandard(c, R))
        qd = (MVNSqrt if sqrt else MVNStandard)(t64(qo[0]), t64(qo[1]))
        rd = (MVNSqrt if sqrt else MVNStandard)(t64(ro[0]), t64(ro[1]))
        nom_o = cls_o(m_nom, np.eye(dim_x))
        Ft, qt, bt = OMETH[method](O.FunctionalModel(O.Linear(F), qo), nom_o)
        Hh, rr, cc = OMETH[method](O.FunctionalModel(O.Linear(H), ro), nom_o)
        mx, Px = (cx0.mean, cx0.chol) if sqrt else (x0.mean, x0.cov)
        ref = (O.sqrt_filter_element if sqrt else O.std_filter_element)(Ft, qt, bt, Hh, rr, cc, mx, Px, y)
        fn = _sqrt_associative_params_one if sqrt else _standard_associative_params_one
        nom_d = cls_d(t64(m_nom), t64(np.eye(dim_x)))
        got, ssm = fn(DMETH[method], FunctionalModel(M.LinearFunction(F), qd),
                      FunctionalModel(M.LinearFunction(H), rd), nom_d, nom_d, t64(mx), t64(Px), t64(y))
        got = [npy(g) for g in got]
        close(got[0], ref[0], what="A")
        close(got[1], ref[1], what="b")
        close(got[3], ref[3], what="eta")
        if sqrt:
            close(gram(got[2]), gram(ref[2]), what="UU^T")
            close(gram(got[4]), gram(ref[4]), what="ZZ^T")
        else:
            close(got[2], ref[2], what="C")
            close(got[4], ref[4], what="J")


@pytest.mark.parametrize("dim_x", [1, 2, 3, 5])
@pytest.mark.parametrize("seed", [0, 42])
def test_smoother_element(dim_x, seed):
    np.random.seed(seed)
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    F = np.random.randn(dim_x, dim_x)
    nom = np.random.randn(dim_x)
    for sqrt in (True, False):
        q = O.MVNSqrt(b, cholQ) if sqrt else O.MVNStandard(b, Q)
        ref = (O.sqrt_smoother_element if sqrt else O.std_smoother_element)(
            F, q[1], b, x0.mean, cx0.chol if sqrt else x0.cov)
        cls = MVNSqrt if sqrt else MVNStandard
        fn = _sqrt_associative_params if sqrt else _standard_associative_params
        got = fn(extended, FunctionalModel(M.LinearFunction(F), cls(t64(b), t64(q[1]))),
                 cls(t64(nom), t64(np.eye(dim_x))), t64(x0.mean), t64(cx0.chol if sqrt else x0.cov))
        got = [npy(g) for g in got]
        close(got[0], ref[0], what="g")
        close(got[1], ref[1], what="E")
        close(got[2], ref[2], what="D")


# ---------------------------------------------------------------- linearization
def _model_pairs(rng):
    A = rng.randn(3, 5)
    return [
        ("linear", M.LinearFunction(A), O.Linear(A), 5, 3),
        ("ct", M.CoordinatedTurnTransition(0.01), O.CoordinatedTurn(0.01), 5, 5),
        ("bearings", M.BearingsObservation((-1.5, 0.5), (1.0, 1.0)), O.Bearings((-1.5, 0.5), (1.0, 1.0)), 5, 2),
        ("sin", M.Elementwise("sin", 3), O.Elementwise("sin"), 3, 3),
        ("cos", M.Elementwise("cos", 3), O.Elementwise("cos"), 3, 3),
        ("exp", M.Elementwise("exp", 2), O.Elementwise("exp"), 2, 2),
        ("arctan", M.Elementwise("arctan", 3), O.Elementwise("arctan"), 3, 3),
    ]


@pytest.mark.parametrize("method", list(OMETH))
@pytest.mark.parametrize("sqrt", [True, False])
def test_linearize(method, sqrt):
    rng = np.random.RandomState(3)
    for name, fd, fo, nin, nout in _model_pairs(rng):
        for w in (0.7, 1e-8):
            B = 5
            m = rng.randn(B, nin)
            if name == "ct":
                m[:, 4] = w
            L = np.stack([rand_lower(nin, rng) * 0.3 + 0.1 * np.eye(nin) for _ in range(B)])
            qm, qc = rng.randn(nout), rand_lower(nout, rng)
            if sqrt:
                xo, qo = O.MVNSqrt(m, L), O.MVNSqrt(qm, qc)
                xd, qd = MVNSqrt(t64(m), t64(L)), MVNSqrt(t64(qm), t64(qc))
            else:
                xo, qo = O.MVNStandard(m, gram(L)), O.MVNStandard(qm, gram(qc))
                xd, qd = MVNStandard(t64(m), t64(gram(L))), MVNStandard(t64(qm), t64(gram(qc)))
            ref = OMETH[method](O.FunctionalModel(fo, qo), xo)
            got = [npy(g) for g in DMETH[method](FunctionalModel(fd, qd), xd)]
            close(got[0], ref[0], tol=1e-8, what=f"{name} F")
            close(got[2], ref[2], tol=1e-8, what=f"{name} b")
            close(gram(got[1]) if sqrt else got[1], gram(ref[1]) if sqrt else ref[1], tol=1e-8,
                  what=f"{name} Q")


def test_gauss_hermite_order_partial():
    rng = np.random.RandomState(0)
    m, L = rng.randn(2, 3), np.stack([np.eye(3)] * 2)
    A = rng.randn(3, 3)
    for order in (2, 4):
        ref = O.gauss_hermite(O.FunctionalModel(O.Elementwise("sin"), O.MVNSqrt(np.zeros(3), np.eye(3))),
                              O.MVNSqrt(m, L), order=order)
        got = functools.partial(gauss_hermite, order=order)(
            FunctionalModel(M.Elementwise("sin", 3), MVNSqrt(t64(np.zeros(3)), t64(np.eye(3)))),
            MVNSqrt(t64(m), t64(L)))
        close(npy(got[0]), ref[0], tol=1e-9, what="F")


# ---------------------------------------------------------------- filter / smoother / iterated
def _linear_problem(dim_x, dim_y, seed, T):
    np.random.seed(seed)
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    # stable dynamics (spectral radius 0.98) so long sequences stay well scaled
    F = 0.98 * np.linalg.qr(np.random.randn(dim_x, dim_x))[0]
    _, _, H, R, cholR, c, _ = O.get_system(dim_x, dim_y)
    _, ys = O.lgssm_data(x0.mean, F, H, R, Q, b, c, T, random_state=seed)
    return x0, cx0, F, Q, cholQ, b, H, R, cholR, c, ys


@pytest.mark.parametrize("dim_x", [1, 2, 3, 4])
@pytest.mark.parametrize("dim_y", [1, 2, 3])
@pytest.mark.parametrize("sqrt", [True, False])
@pytest.mark.parametrize("T", [1, 10, 333])
def test_filter_smoother_linear(dim_x, dim_y, sqrt, T):
    x0, cx0, F, Q, cholQ, b, H, R, cholR, c, ys = _linear_problem(dim_x, dim_y, 1, T)
    if sqrt:
        xo, to, oo = cx0, O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cholQ)), O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cholR))
        cls, qn, rn = MVNSqrt, (b, cholQ), (c, cholR)
    else:
        xo, to, oo = x0, O.FunctionalModel(O.Linear(F), O.MVNStandard(b, Q)), O.FunctionalModel(O.Linear(H), O.MVNStandard(c, R))
        cls, qn, rn = MVNStandard, (b, Q), (c, R)
    nom = np.random.RandomState(2).randn(T + 1, dim_x)
    nom_o = type(xo)(nom, np.repeat(np.eye(dim_x)[None], T + 1, 0))
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
    if sqrt:
        close(npy(sd[1]), so[1], tol=1e-8, what="smoothed chol")


def _linear_models(F, b, q, H, c, r, sqrt):
    cls = MVNSqrt if sqrt else MVNStandard
    return (FunctionalModel(M.LinearFunction(F), cls(t64(b), t64(q))),
            FunctionalModel(M.LinearFunction(H), cls(t64(c), t64(r))), cls)


@pytest.mark.parametrize("dim_x,dim_y", [(1, 2), (3, 2), (3, 3)])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("sqrt", [True, False])
@pytest.mark.parametrize("method", list(DMETH))
def test_filter_no_information(dim_x, dim_y, seed, sqrt, method):
    """tests/test_sequential_filter.py:120-150 through the parallel filter:
    observation noise x1e12 (chol x1e6), the filtered means are the prediction
    F m + b iterated from the prior (rtol/atol 1e-3, the reference's)."""
    np.random.seed(seed)
    T = 3
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    _, _, H, R, cholR, c, _ = O.get_system(dim_x, dim_y)
    R, cholR = 1e12 * R, 1e6 * cholR
    _, ys = O.lgssm_data(x0.mean, F, H, R, Q, b, c, T, random_state=seed)
    td, od, cls = _linear_models(F, b, cholQ if sqrt else Q, H, c, cholR if sqrt else R, sqrt)
    prior = cls(t64(x0.mean), t64(cx0.chol if sqrt else x0.cov))
    f = parsmooth.filtering(t64(ys), prior, td, od, DMETH[method])
    expected = [x0.mean]
    for _ in range(T):
        expected.append(F @ expected[-1] + b)
    np.testing.assert_allclose(npy(f.mean), np.stack(expected), atol=1e-3, rtol=1e-3)


@pytest.mark.parametrize("dim", [1, 3])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("sqrt", [True, False])
@pytest.mark.parametrize("method", list(DMETH))
def test_filter_infinite_information(dim, seed, sqrt, method):
    """tests/test_sequential_filter.py:153-184 through the parallel filter: F = H
    = I, observation noise x1e-6 (chol x1e-3): the filtered means are y - c."""
    np.random.seed(seed)
    T = 3
    x0, cx0, _, Q, cholQ, b, _ = O.get_system(dim, dim)
    F = np.eye(dim)
    _, _, _, R, cholR, c, _ = O.get_system(dim, dim)
    H = np.eye(dim)
    R, cholR = 1e-6 * R, 1e-3 * cholR
    _, ys = O.lgssm_data(x0.mean, F, H, R, Q, b, c, T, random_state=seed)
    td, od, cls = _linear_models(F, b, cholQ if sqrt else Q, H, c, cholR if sqrt else R, sqrt)
    prior = cls(t64(x0.mean), t64(cx0.chol if sqrt else x0.cov))
    f = parsmooth.filtering(t64(ys), prior, td, od, DMETH[method])
    np.testing.assert_allclose(npy(f.mean)[1:], np.asarray(ys) - c, atol=1e-3, rtol=1e-3)


@pytest.mark.parametrize("dim_x", [1, 3])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("sqrt", [True, False])
@pytest.mark.parametrize("noise", ["none", "infinite"])
def test_smoother_step_noise_limits(dim_x, seed, sqrt, noise):
    """tests/test_sequential_smoother.py:67-120 through the parallel smoother:
    smoothing the two-state filtered trajectory [x_f, x_s] gives at state 0 the
    smoothing step from x_f given the smoothed x_s.  No process noise: the RTS
    step with gain F^+ (mean x_f + F^+ (m_s - F m_f - b), covariance
    P_f + F^+ (P_s - F P_f F^T) F^+T); infinite process noise: x_f unchanged
    (atol 1e-5, the reference's)."""
    np.random.seed(seed)
    xf, cxf, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    xs, cxs, *_ = O.get_system(dim_x, dim_x)
    if noise == "none":
        Q, cholQ = 0.0 * Q, 0.0 * cholQ
    else:
        Q, cholQ = 1e12 * Q, 1e6 * cholQ
    cls = MVNSqrt if sqrt else MVNStandard
    model = FunctionalModel(M.LinearFunction(F), cls(t64(b), t64(cholQ if sqrt else Q)))
    traj = cls(t64(np.stack([xf.mean, xs.mean])),
               t64(np.stack([cxf.chol, cxs.chol]) if sqrt else np.stack([xf.cov, xs.cov])))
    sm = parsmooth.smoothing(model, traj, extended)
    m0 = npy(sm.mean)[0]
    P0 = gram(npy(sm.chol)[0]) if sqrt else npy(sm.cov)[0]
    if noise == "none":
        G = np.linalg.pinv(F)
        ms = xf.mean + G @ (xs.mean - (F @ xf.mean + b))
        Ps = xf.cov + G @ (xs.cov - F @ xf.cov @ F.T) @ G.T
    else:
        ms, Ps = xf.mean, xf.cov
    np.testing.assert_allclose(m0, ms, atol=1e-5)
    np.testing.assert_allclose(P0, Ps, atol=1e-5)


def _bearings_setup(T, seed=0):
    Q, R, obs_f, trans_f = M.make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    from parsmooth.data import bearings_data
    _, _, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0),
                             random_state=seed)
    cQ, cR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    to = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNSqrt(np.zeros(5), cQ))
    oo = O.FunctionalModel(O.Bearings((-1.5, 0.5), (1.0, 1.0)), O.MVNSqrt(np.zeros(2), cR))
    td = FunctionalModel(trans_f, MVNSqrt(t64(np.zeros(5)), t64(cQ)))
    od = FunctionalModel(obs_f, MVNSqrt(t64(np.zeros(2)), t64(cR)))
    return ys, (to, oo), (td, od)


@pytest.mark.parametrize("method", ["extended", "cubature"])
def test_iterated_bearings_vs_oracle(method):
    T = 300
    ys, (to, oo), (td, od) = _bearings_setup(T)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
    ref, ell_o = O.iterated_smoothing(ys, O.MVNSqrt(m0, np.eye(5)), to, oo, OMETH[method], None, True,
                                      criterion=lambda i, *_: i < 3, return_loglikelihood=True)
    got, ell_d = parsmooth.iterated_smoothing(t64(ys), MVNSqrt(t64(m0), t64(np.eye(5))), td, od,
                                              DMETH[method], criterion=lambda i, *_: i < 3,
                                              return_loglikelihood=True)
    # iterated nonlinear smoothing amplifies rounding differences between passes
    close(npy(got.mean), ref.mean, tol=1e-8, what="mean")
    close(gram(npy(got.chol)), gram(ref.chol), tol=1e-8, what="cov")
    assert abs(float(ell_d) - ell_o) <= 1e-8 * abs(ell_o)


@pytest.mark.parametrize("method", ["extended", "cubature", "gauss_hermite", "unscented"])
def test_single_pass_bearings_1e9(method):
    """One filter_smoother pass at an identical nominal: 1e-9 relative (Gauss-Hermite:
    order 3, 243 points; unscented: alpha 1, beta 0, kappa 3 + n, 11 points)."""
    T = 2000
    ys, (to, oo), (td, od) = _bearings_setup(T, seed=3)
    rng = np.random.RandomState(0)
    nom_m = np.cumsum(0.01 * rng.randn(T + 1, 5), 0) + np.array([0.1, 0.2, 1.0, 0.0, 0.5])
    nom_c = np.repeat(np.eye(5)[None] * 0.1, T + 1, 0)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
    ref = O.filter_smoother(ys, O.MVNSqrt(m0, np.eye(5)), to, oo, OMETH[method], O.MVNSqrt(nom_m, nom_c), True)
    got = parsmooth.filter_smoother(t64(ys), MVNSqrt(t64(m0), t64(np.eye(5))), td, od, DMETH[method],
                                    MVNSqrt(t64(nom_m), t64(nom_c)))
    close(npy(got.mean), ref.mean, what="mean")
    close(gram(npy(got.chol)), gram(ref.chol), what="cov")


@pytest.mark.parametrize("method", ["extended", "cubature", "gauss_hermite", "unscented"])
def test_single_pass_bearings_standard_form(method):
    """Covariance form (MVNStandard in and out) on the nonlinear bearings model: one
    filter_smoother pass at an identical nominal against the oracle, 1e-9
    relative, plus the filter log-likelihood."""
    T = 500
    ys, _, _ = _bearings_setup(T, seed=4)
    Q, R, obs_f, trans_f = M.make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    to = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNStandard(np.zeros(5), Q))
    oo = O.FunctionalModel(O.Bearings((-1.5, 0.5), (1.0, 1.0)), O.MVNStandard(np.zeros(2), R))
    td = FunctionalModel(trans_f, MVNStandard(t64(np.zeros(5)), t64(Q)))
    od = FunctionalModel(obs_f, MVNStandard(t64(np.zeros(2)), t64(R)))
    rng = np.random.RandomState(1)
    nom_m = np.cumsum(0.01 * rng.randn(T + 1, 5), 0) + np.array([0.1, 0.2, 1.0, 0.0, 0.5])
    nom_P = np.repeat(np.eye(5)[None] * 0.1, T + 1, 0)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
    nom_o = O.MVNStandard(nom_m, nom_P)
    fo, ell_o = O.filtering(ys, O.MVNStandard(m0, np.eye(5)), to, oo, OMETH[method], nom_o, True,
                            return_loglikelihood=True)
    so = O.smoothing(to, fo, OMETH[method], nom_o, True)
    nom_d = MVNStandard(t64(nom_m), t64(nom_P))
    fd, ell_d = parsmooth.filtering(t64(ys), MVNStandard(t64(m0), t64(np.eye(5))), td, od, DMETH[method],
                                    nom_d, return_loglikelihood=True)
    sd = parsmooth.smoothing(td, fd, DMETH[method], nom_d)
    assert isinstance(sd, MVNStandard)
    close(npy(fd.mean), fo.mean, what="filtered mean")
    close(npy(fd.cov), fo.cov, what="filtered cov")
    close(npy(sd.mean), so.mean, what="smoothed mean")
    close(npy(sd.cov), so.cov, what="smoothed cov")
    assert abs(float(ell_d) - ell_o) <= TOL64 * max(1.0, abs(ell_o)), (float(ell_d), ell_o)


@pytest.mark.parametrize("dim_x,dim_y", [(2, 1), (4, 2)])
def test_fp32_standard_form_linear(dim_x, dim_y):
    """FP32 covariance form on a stable linear model, T=333: filtered and smoothed
    means and covariances against the FP64 oracle run on the same FP32-rounded
    inputs, within the FP32 tolerances."""
    r32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    x0, _, F, Q, _, b, H, R, _, c, ys = _linear_problem(dim_x, dim_y, 3, 333)
    F, Q, b, H, R, c, ys = map(r32, (F, Q, b, H, R, c, ys))
    m0, P0 = r32(x0.mean), r32(x0.cov)
    nom = r32(np.random.RandomState(5).randn(334, dim_x))
    nom_P = np.repeat(np.eye(dim_x)[None], 334, 0)
    to = O.FunctionalModel(O.Linear(F), O.MVNStandard(b, Q))
    oo = O.FunctionalModel(O.Linear(H), O.MVNStandard(c, R))
    fo = O.filtering(ys, O.MVNStandard(m0, P0), to, oo, O.extended, O.MVNStandard(nom, nom_P), True)
    so = O.smoothing(to, fo, O.extended, O.MVNStandard(nom, nom_P), True)
    t32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device=DEV)
    td = FunctionalModel(M.LinearFunction(F), MVNStandard(t32(b), t32(Q)))
    od = FunctionalModel(M.LinearFunction(H), MVNStandard(t32(c), t32(R)))
    nom_d = MVNStandard(t32(nom), t32(nom_P))
    fd = parsmooth.filtering(t32(ys), MVNStandard(t32(m0), t32(P0)), td, od, extended, nom_d)
    sd = parsmooth.smoothing(td, fd, extended, nom_d)
    assert sd.mean.dtype == torch.float32
    close(npy(fd.mean), fo.mean, tol=TOL32, what="fp32 std filtered mean")
    close(npy(fd.cov), fo.cov, tol=TOL32_COV, what="fp32 std filtered cov")
    close(npy(sd.mean), so.mean, tol=TOL32, what="fp32 std smoothed mean")
    close(npy(sd.cov), so.cov, tol=TOL32_COV, what="fp32 std smoothed cov")


def test_ieks_fixture(golden_dir):
    """tests/test_bearings_only.py:21-72 on the device (extended, 100 iterations)."""
    ys = np.load(f"{golden_dir}/bearings_ys.npy")
    exp = np.load(f"{golden_dir}/bearings_ieks.npz")
    _, (to, oo), (td, od) = _bearings_setup(10)
    got = parsmooth.iterated_smoothing(t64(ys), MVNSqrt(t64(np.array([-1.0, -1.0, 0, 0, 0])), t64(np.eye(5))),
                                       td, od, extended, criterion=lambda i, *_: i < 100)
    np.testing.assert_array_almost_equal(npy(got.mean)[1:], exp["arr_0"], decimal=3)
    np.testing.assert_array_almost_equal(gram(npy(got.chol))[1:], exp["arr_1"], decimal=3)


def test_default_criterion_stops_like_oracle():
    T = 200
    ys, (to, oo), (td, od) = _bearings_setup(T, seed=5)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
    n_o = []
    def crit_o(i, p, c):
        n_o.append(i)
        return O.default_criterion(i, p, c)
    n_d = []
    def crit_d(i, p, c):
        n_d.append(i)
        return parsmooth.methods._default_criterion(i, p, c)
    O.iterated_smoothing(ys, O.MVNSqrt(m0, np.eye(5)), to, oo, O.extended, None, True, criterion=crit_o)
    parsmooth.iterated_smoothing(t64(ys), MVNSqrt(t64(m0), t64(np.eye(5))), td, od, extended, criterion=crit_d)
    assert n_o == n_d


# ---------------------------------------------------------------- ragged batch
@pytest.mark.parametrize("method", ["extended", "cubature"])
def test_ragged_batch_matches_single_sequences(method):
    from parsmooth.batch import iterated_smoothing as batch_is
    lengths = [1, 7, 64, 5, 130, 2]
    rng = np.random.RandomState(11)
    Q, R, obs_f, trans_f = M.make_parameters(0.01, 0.1, 0.5, 0.01, (-1.5, 0.5), (1.0, 1.0))
    from parsmooth.data import bearings_data
    ys = [bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0),
                        random_state=k)[2] for k, T in enumerate(lengths)]
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0]) + 0.3 * rng.randn(len(lengths), 5)
    L0 = np.repeat(np.eye(5)[None], len(lengths), 0)
    cQ, cR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    td = FunctionalModel(trans_f, MVNSqrt(t64(np.zeros(5)), t64(cQ)))
    od = FunctionalModel(obs_f, MVNSqrt(t64(np.zeros(2)), t64(cR)))
    trajs, ell = batch_is(ys, MVNSqrt(t64(m0), t64(L0)), td, od, DMETH[method], n_iter=2,
                          return_loglikelihood=True)
    to = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNSqrt(np.zeros(5), cQ))
    oo = O.FunctionalModel(O.Bearings((-1.5, 0.5), (1.0, 1.0)), O.MVNSqrt(np.zeros(2), cR))
    ell = npy(ell)
    for s, y in enumerate(ys):
        ref, e = O.iterated_smoothing(y, O.MVNSqrt(m0[s], L0[s]), to, oo, OMETH[method], None, True,
                                      criterion=lambda i, *_: i < 2, return_loglikelihood=True)
        close(npy(trajs[s].mean), ref.mean, tol=1e-8, what=f"seq {s} mean")
        close(gram(npy(trajs[s].chol)), gram(ref.chol), tol=1e-8, what=f"seq {s} cov")
        assert abs(ell[s] - e) <= 1e-8 * max(1.0, abs(e))


def _schedule_runner(lengths, seed, chunk, scan_k, dtype=torch.float64):
    from parsmooth.batch import IteratedSmoother
    from parsmooth.data import bearings_data_batch
    rng = np.random.RandomState(seed)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0]) + 0.3 * rng.randn(len(lengths), 5)
    _, obs = bearings_data_batch(lengths, m0[:, :4], 0.01, 0.5, (-1.5, 0.5), (1.0, 1.0), seed=seed)
    _, _, (td, od) = _bearings_setup(2, seed=0)
    r = IteratedSmoother(lengths, td, od, extended, dtype=dtype)
    r.prob.set_schedule(chunk, scan_k)
    r.set_data(t64(np.concatenate(obs)), t64(m0), t64(np.repeat(np.eye(5)[None], len(lengths), 0)))
    M = r.prob.n_states
    nom = (np.repeat(m0, np.asarray(lengths) + 1, 0) + 0.05 * rng.randn(M, 5),
           np.repeat(np.eye(5)[None] * 0.3, M, 0))
    return r, obs, m0, nom


_SCHED_LENGTHS = [300, 17, 1, 2500, 64]


@functools.lru_cache(maxsize=1)
def _schedule_refs():
    """Oracle filter-smoother of the schedule-test batch (seed 3), per sequence."""
    r, obs, m0, nom = _schedule_runner(_SCHED_LENGTHS, 3, 4, 0)
    to, oo = _bearings_setup(2, seed=0)[1]
    o = r.state_offsets
    return [O.filter_smoother(np.asarray(y, np.float64), O.MVNSqrt(m0[s], np.eye(5)), to, oo, O.extended,
                              O.MVNSqrt(nom[0][o[s]:o[s + 1]], nom[1][o[s]:o[s + 1]]), True)
            for s, y in enumerate(obs)]


@pytest.mark.parametrize("chunk", [2, 3, 4, 5, 7, 8, 12, 16, 32])
def test_every_schedule_matches_oracle(chunk):
    """One filter-smoother pass over a ragged batch for each (chunk, scan fan-in,
    scan mode) schedule the autotuner may pick, against the oracle per sequence."""
    lengths = _SCHED_LENGTHS
    refs = _schedule_refs()
    top16 = 16 << N.SCAN_TOP_SHIFT  # top level may fold up to 16 items per lane
    half, quad = N.SCAN_HALF, N.SCAN_TEAM | N.SCAN_QUAD
    for k in (2, 4, 8, 2 | N.SCAN_TEAM, 8 | N.SCAN_TEAM, 2 | N.SCAN_TEAM | top16, 4 | top16,
              2 | half, 4 | N.SCAN_TEAM | half, 2 | N.SCAN_TEAM | half, 2 | quad, 4 | quad | half,
              2 | quad | half, 8 | N.SCAN_QUAD):
        r, obs, m0, nom = _schedule_runner(lengths, 3, chunk, k)
        assert r.chunk_size() == chunk
        out = r.prob.new_traj()
        r.pass_once((t64(nom[0]), t64(nom[1])), out)
        got_m, got_L = r.split(out[0]), r.split(out[1])
        for s, ref in enumerate(refs):
            close(npy(got_m[s]), ref.mean, what=f"chunk {chunk} K {k} seq {s} mean")
            close(gram(npy(got_L[s])), gram(ref.chol), what=f"chunk {chunk} K {k} seq {s} cov")


def test_deep_scan_schedules_agree():
    """T = 300000: two- to five-level scans (chunk 32/K 8 ... chunk 2/K 2, per-lane
    or team upper levels, top level folding up to K or 16 items per lane) give
    the same pass to FP64 rounding."""
    T = 300000
    outs = []
    top16 = 16 << N.SCAN_TOP_SHIFT
    for chunk, k in ((32, 8), (4, 4), (2, 2), (4, 4 | N.SCAN_TEAM), (2, 2 | N.SCAN_TEAM),
                     (8, 2 | N.SCAN_TEAM | top16), (2, 2 | top16), (4, 2 | N.SCAN_TEAM | N.SCAN_HALF),
                     (5, 2 | N.SCAN_TEAM | N.SCAN_QUAD | N.SCAN_HALF), (4, 4 | N.SCAN_TEAM | N.SCAN_QUAD)):
        r, _, _, nom = _schedule_runner([T], 4, chunk, k)
        out = r.prob.new_traj()
        r.pass_once((t64(nom[0]), t64(nom[1])), out)
        torch.cuda.synchronize()
        outs.append((npy(out[0]), gram(npy(out[1]))))
        del r, out
    for m, P in outs[1:]:
        close(m, outs[0][0], what="deep scan mean")
        close(P, outs[0][1], what="deep scan cov")


def test_autotune_result_matches_default_schedule():
    lengths = [4000, 123, 9000]
    r, _, _, nom = _schedule_runner(lengths, 5, 0, 0)
    ref = [npy(x) for x in r.run(3, (nom[0], nom[1]), use_graph=False)]
    chunk, k, ms = r.autotune((nom[0], nom[1]))
    assert 2 <= chunk <= 64 and k in r.prob.TUNE_SCAN_K and ms > 0
    assert r.chunk_size() == chunk
    got = [npy(x) for x in r.run(3, (nom[0], nom[1]), use_graph=True)]
    close(got[0], ref[0], tol=1e-8, what="autotuned mean")
    close(gram(got[1]), gram(ref[1]), tol=1e-8, what="autotuned cov")
    # a second call on the same shape is served from the cache
    assert r.autotune((nom[0], nom[1]))[:2] == (chunk, k)


def test_graph_replay_matches_eager():
    from parsmooth.batch import IteratedSmoother
    T = 500
    ys, _, (td, od) = _bearings_setup(T, seed=2)
    m0 = np.array([[0.1, 0.2, 1.0, 0.0, 0.0]])
    runner = IteratedSmoother([T], td, od, extended)
    runner.set_data(t64(ys), t64(m0), t64(np.eye(5)[None]))
    eager = [npy(x) for x in runner.run(4, use_graph=False)]
    for _ in range(3):
        graph = [npy(x) for x in runner.run(4, use_graph=True)]
        torch.cuda.synchronize()
        np.testing.assert_array_equal(graph[0], eager[0])
        np.testing.assert_array_equal(graph[1], eager[1])


# ---------------------------------------------------------------- FP32
def test_fp32_bearings_vs_fp64_oracle():
    """FP32 square-root IEKS, 5 iterations, T=1000: max relative error of the
    smoothed mean against the FP64 oracle below 1e-4 of the trajectory scale."""
    T = 1000
    ys, (to, oo), (td64, od64) = _bearings_setup(T, seed=7)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
    ref = O.iterated_smoothing(ys, O.MVNSqrt(m0, np.eye(5)), to, oo, O.extended, None, True,
                               criterion=lambda i, *_: i < 5)
    t32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device=DEV)
    td = FunctionalModel(td64.function, MVNSqrt(t32(npy(td64.mvn.mean)), t32(npy(td64.mvn.chol))))
    od = FunctionalModel(od64.function, MVNSqrt(t32(npy(od64.mvn.mean)), t32(npy(od64.mvn.chol))))
    got = parsmooth.iterated_smoothing(t32(ys), MVNSqrt(t32(m0), t32(np.eye(5))), td, od, extended,
                                       criterion=lambda i, *_: i < 5)
    assert got.mean.dtype == torch.float32
    close(npy(got.mean), ref.mean, tol=1e-4, what="fp32 mean")


def _ct_turn_rate_column(m, dt):
    """The turn-rate column of the coordinated-turn Jacobian from the Taylor
    series of d/dw sin(w dt)/w = dt^2 s'(x) and d/dw (cos(w dt) - 1)/w =
    -dt^2 g'(x), x = w dt (exact to FP64 rounding for |x| < 1e-2, where the
    closed forms (x cos x - sin x)/x^2 and (x sin x - 1 + cos x)/x^2 cancel)."""
    x = m[:, 4] * dt
    x2 = x * x
    ds = -x / 3 + x * x2 / 30 - x * x2 * x2 / 840
    dg = 0.5 - x2 / 8 + x2 * x2 / 144 - x2 * x2 * x2 / 5760
    da, db = dt * dt * ds, -dt * dt * dg
    col = np.zeros((m.shape[0], 5))
    col[:, 0] = da * m[:, 2] - db * m[:, 3]
    col[:, 1] = db * m[:, 2] + da * m[:, 3]
    col[:, 2] = -dt * np.sin(x) * m[:, 2] + dt * np.cos(x) * m[:, 3]
    col[:, 3] = -dt * np.cos(x) * m[:, 2] - dt * np.sin(x) * m[:, 3]
    col[:, 4] = 1.0
    return col


def test_fp32_extended_small_turn_rate():
    """FP32 extended linearization of the coordinated turn at small turn rates
    (w dt from 1e-4 down to 2e-8, above the model's 1e-6 branch on w): the
    turn-rate column of the Jacobian, e.g. d/dw (1 - cos(w dt)) / w, cancels in
    the closed form (in FP32 cos(w dt) rounds to 1 and the derivative doubles),
    so FP32 problems evaluate it from its Taylor series.  Each entry must match
    the exact derivative to 1e-5 of its own size, and the FP64 oracle (the
    reference's closed form, which itself loses up to ~1e-16 / (w dt)^2 of the
    value) to 1e-5 plus the oracle's own error; b = f(m) - F m to FP32 rounding
    of a large |m|."""
    dt = float(np.float32(0.01))
    Q, _, _, trans_f = M.make_parameters(0.01, 0.1, 0.5, dt, (-1.5, 0.5), (1.0, 1.0))
    cQ = np.asarray(np.linalg.cholesky(Q), np.float32).astype(np.float64)
    rng = np.random.RandomState(12)
    ws = np.array([1e-2, -1.7e-3, 4e-4, -1e-4, 1e-5, 2.1e-6])
    m = np.zeros((len(ws), 5))
    m[:, :4] = 3e3 * rng.randn(len(ws), 4)
    m[:, 4] = ws
    m = np.asarray(m, np.float32).astype(np.float64)
    L = np.repeat(np.eye(5)[None], len(ws), 0)
    t32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device=DEV)
    F, _, b = [npy(g) for g in extended(FunctionalModel(trans_f, MVNSqrt(t32(np.zeros(5)), t32(cQ))),
                                        MVNSqrt(t32(m), t32(L)))]
    Fo, _, bo = O.extended(O.FunctionalModel(O.CoordinatedTurn(dt), O.MVNSqrt(np.zeros(5), cQ)), O.MVNSqrt(m, L))
    col, colo, exact = F[:, :, 4], Fo[:, :, 4], _ct_turn_rate_column(m, dt)
    # the series agrees with the oracle's closed form where that one is accurate (w dt ~ 1e-2)
    mm = m.copy()
    mm[:, 4] = np.array([1.0, -2.0, 0.5, -0.7, 1.3, 0.9])
    assert np.allclose(_ct_turn_rate_column(mm, dt), O.CoordinatedTurn(dt).jac(mm)[:, :, 4], rtol=1e-10, atol=0)
    assert np.all(np.abs(col - exact) <= 1e-5 * np.abs(exact) + 1e-12), (col, exact)
    assert np.all(np.abs(col - colo) <= 1e-5 * np.abs(colo) + np.abs(exact - colo) + 1e-12), (col, colo)
    close(np.delete(F, 4, axis=2), np.delete(Fo, 4, axis=2), tol=1e-6, what="fp32 CT Jacobian")
    close(b, bo, tol=1e-6, what="fp32 CT offset")


@pytest.mark.parametrize("method", ["cubature", "unscented", "gauss_hermite"])
@pytest.mark.parametrize("scale", [1.0, 100.0, 1500.0])
def test_fp32_sigma_regression_small_noise(method, scale):
    """FP32 sigma-point linearization of the coordinated-turn transition, whose
    process noise (chol ~6e-5 on the positions) is tiny against the sigma-point
    spread (0.3): the regression noise factor chol(Phi + Q - F P F^T) against the
    FP64 oracle on the same FP32-rounded inputs.  The regression works from
    centred residuals with FP64 weights, so neither the cancellation of Phi and
    F P F^T nor weights rounded off one (an FP32 table leaves ~1e-8, more than
    Q itself) reaches the factor; positions up to 1500 (scale) check that the
    sigma points' increments are formed without rounding m + d (at 1500 an FP32
    position carries ~1e-4 of rounding, above chol Q)."""
    r32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    dt = float(np.float32(0.01))
    Q, _, _, trans_f = M.make_parameters(0.01, 0.1, 0.5, dt, (-1.5, 0.5), (1.0, 1.0))
    cQ = r32(np.linalg.cholesky(Q))
    rng = np.random.RandomState(11)
    B = 64
    m = np.array([0.1, 0.2, 1.0, 0.0, 0.5]) + rng.randn(B, 5)
    m[:, :2] *= scale
    m = r32(m)
    L = r32(np.stack([0.3 * np.eye(5) + 0.05 * np.tril(rng.randn(5, 5), -1) for _ in range(B)]))
    t32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device=DEV)
    got = [npy(g) for g in DMETH[method](FunctionalModel(trans_f, MVNSqrt(t32(np.zeros(5)), t32(cQ))),
                                         MVNSqrt(t32(m), t32(L)))]
    ref = OMETH[method](O.FunctionalModel(O.CoordinatedTurn(dt), O.MVNSqrt(np.zeros(5), cQ)), O.MVNSqrt(m, L))
    close(got[0], ref[0], tol=TOL32, what=f"fp32 {method} F")
    if scale == 1.0:
        close(got[2], ref[2], tol=TOL32, what=f"fp32 {method} b")
    # the linearized prediction F m + b (b = f(m) + ... - F m carries the FP32
    # rounding of F m, ~1e-7 of |m|, which the prediction cancels)
    pred = lambda F, b: np.einsum("bij,bj->bi", F, m) + b
    close(pred(got[0], got[2]), pred(ref[0], ref[2]), tol=TOL32, what=f"fp32 {method} F m + b")
    # the noise factor relative to its own scale (Q ~ 1e-9 .. 1e-3)
    Pg, Pr = gram(got[1]), gram(ref[1])
    err = float(np.abs(Pg - Pr).max() / np.abs(Pr).max())
    assert err <= 1e-3, f"fp32 {method} regression noise: rel err {err:.2e}"
    dg, dr = np.abs(np.diagonal(got[1], axis1=1, axis2=2)), np.abs(np.diagonal(ref[1], axis1=1, axis2=2))
    assert np.all(np.abs(dg - dr) <= 1e-3 * dr + 1e-9), "fp32 regression noise factor diagonal"


@pytest.mark.parametrize("k", [2, 2 | N.SCAN_TEAM, 2 | N.SCAN_TEAM | N.SCAN_HALF, 4 | N.SCAN_HALF,
                               2 | N.SCAN_TEAM | N.SCAN_QUAD | N.SCAN_HALF, 4 | N.SCAN_QUAD])
def test_fp32_schedules_match_oracle(k):
    """FP32 filter-smoother pass over a ragged batch with per-lane, team-cooperative
    and half-wave scan schedules, against the FP64 oracle run on the same
    FP32-rounded inputs (FP32 tolerances)."""
    from parsmooth.batch import IteratedSmoother
    from parsmooth.data import bearings_data_batch
    lengths = [2000, 33, 1, 700]
    rng = np.random.RandomState(8)
    r32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    m0 = r32(np.array([0.1, 0.2, 1.0, 0.0, 0.0]) + 0.3 * rng.randn(len(lengths), 5))
    _, obs = bearings_data_batch(lengths, m0[:, :4], 0.01, 0.5, (-1.5, 0.5), (1.0, 1.0), seed=8)
    obs = [r32(y) for y in obs]
    _, (to, oo), (td64, od64) = _bearings_setup(2, seed=0)
    t32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device=DEV)
    td = FunctionalModel(td64.function, MVNSqrt(t32(npy(td64.mvn.mean)), t32(npy(td64.mvn.chol))))
    od = FunctionalModel(od64.function, MVNSqrt(t32(npy(od64.mvn.mean)), t32(npy(od64.mvn.chol))))
    r = IteratedSmoother(lengths, td, od, extended, dtype=torch.float32)
    r.prob.set_schedule(4, k)
    r.set_data(t32(np.concatenate(obs)), t32(m0), t32(np.repeat(np.eye(5)[None], len(lengths), 0)))
    M = r.prob.n_states
    nom = (r32(np.repeat(m0, np.asarray(lengths) + 1, 0) + 0.05 * rng.randn(M, 5)),
           r32(np.repeat(np.eye(5)[None] * 0.3, M, 0)))
    out = r.prob.new_traj()
    r.pass_once((t32(nom[0]), t32(nom[1])), out)
    got_m, got_L = r.split(out[0]), r.split(out[1])
    o = r.state_offsets
    for s, y in enumerate(obs):
        ref = O.filter_smoother(y, O.MVNSqrt(m0[s], np.eye(5)), to, oo, O.extended,
                                O.MVNSqrt(nom[0][o[s]:o[s + 1]], nom[1][o[s]:o[s + 1]]), True)
        close(npy(got_m[s]), ref.mean, tol=TOL32, what=f"fp32 K {k:#x} seq {s} mean")
        close(gram(npy(got_L[s])), gram(ref.chol), tol=TOL32_COV, what=f"fp32 K {k:#x} seq {s} cov")


@pytest.mark.parametrize("scan_k", [0, 2 | N.SCAN_TEAM | N.SCAN_QUAD | N.SCAN_HALF])
def test_fp32_fallback_recomputes_in_fp64(scan_k):
    """FP32 square-root pass at an identical nominal.  With the re-triangularisation
    threshold forced to 1 every check fires and the FP64 recomputation path runs
    (counter > 0); with the check off nothing is counted.  Both results match the
    FP64 oracle evaluated on the same FP32-rounded inputs to the FP32 tolerance,
    and recomputing in FP64 does not make the result worse.  The default schedule
    and one with the quad-cooperative level 0."""
    from parsmooth import set_fp32_fallback
    from parsmooth.batch import IteratedSmoother
    T = 1500
    ys, _, (td64, od64) = _bearings_setup(T, seed=9)
    r32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    ys = r32(ys)
    m0 = r32([0.1, 0.2, 1.0, 0.0, 0.0])
    rng = np.random.RandomState(1)
    nom_m = r32(np.cumsum(0.01 * rng.randn(T + 1, 5), 0) + np.array([0.1, 0.2, 1.0, 0.0, 0.5]))
    nom_c = r32(np.repeat(np.eye(5)[None] * 0.1, T + 1, 0))
    cQ, cR = r32(npy(td64.mvn.chol)), r32(npy(od64.mvn.chol))
    to = O.FunctionalModel(O.CoordinatedTurn(float(np.float32(0.01))), O.MVNSqrt(np.zeros(5), cQ))
    oo = O.FunctionalModel(O.Bearings((-1.5, 0.5), (1.0, 1.0)), O.MVNSqrt(np.zeros(2), cR))
    ref = O.filter_smoother(ys, O.MVNSqrt(m0, np.eye(5)), to, oo, O.extended, O.MVNSqrt(nom_m, nom_c), True)
    t32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device=DEV)
    td = FunctionalModel(td64.function, MVNSqrt(t32(np.zeros(5)), t32(cQ)))
    od = FunctionalModel(od64.function, MVNSqrt(t32(np.zeros(2)), t32(cR)))
    errs, counts = [], []
    try:
        for ratio in (0.0, 1.0):
            set_fp32_fallback(ratio)
            r = IteratedSmoother([T], td, od, extended, dtype=torch.float32)
            r.prob.set_schedule(0, scan_k)
            r.set_data(t32(ys), t32(m0[None]), t32(np.eye(5)[None]))
            out = r.prob.new_traj()
            r.pass_once((t32(nom_m), t32(nom_c)), out)
            torch.cuda.synchronize()
            close(npy(out[0]), ref.mean, tol=TOL32, what=f"fp32 mean, fallback ratio {ratio}")
            close(gram(npy(out[1])), gram(ref.chol), tol=TOL32_COV, what=f"fp32 cov, fallback ratio {ratio}")
            errs.append(np.abs(npy(out[0]) - ref.mean).max())
            counts.append(r.fallback_count())
    finally:
        set_fp32_fallback(1e-4)
    print("fp32 mean errors (no fallback, all recomputed):", errs, "fallback counts:", counts)
    assert counts[0] == 0 and counts[1] > 0
    assert errs[1] <= 1.5 * errs[0]


def test_fp32_cubature_pass_matches_oracle():
    """FP32 cubature pass on the bearings model at a tight nominal (chol 0.1 I).
    In FP32 the rank-one downdates of the transition's statistical linear
    regression would cancel (chol Q ~ 6e-5 against a sigma-point spread ~ 0.1);
    the FP32 path runs the regression in FP64, and the pass matches the FP64
    oracle on the same FP32-rounded inputs."""
    from parsmooth.batch import IteratedSmoother
    from parsmooth.data import bearings_data
    T = 1000
    _, truth, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0),
                                 random_state=1)
    r32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    _, (to64, oo64), (td64, od64) = _bearings_setup(2)
    cQ, cR = r32(npy(td64.mvn.chol)), r32(npy(od64.mvn.chol))
    to = O.FunctionalModel(O.CoordinatedTurn(float(np.float32(0.01))), O.MVNSqrt(np.zeros(5), cQ))
    oo = O.FunctionalModel(O.Bearings((-1.5, 0.5), (1.0, 1.0)), O.MVNSqrt(np.zeros(2), cR))
    nom_m = r32(truth)
    nom_c = r32(np.repeat(0.1 * np.eye(5)[None], T + 1, 0))
    m0 = r32([0.1, 0.2, 1.0, 0.0, 0.0])
    ys = r32(ys)
    ref = O.filter_smoother(ys, O.MVNSqrt(m0, np.eye(5)), to, oo, O.cubature, O.MVNSqrt(nom_m, nom_c), True)
    t32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device=DEV)
    td = FunctionalModel(td64.function, MVNSqrt(t32(np.zeros(5)), t32(cQ)))
    od = FunctionalModel(od64.function, MVNSqrt(t32(np.zeros(2)), t32(cR)))
    r = IteratedSmoother([T], td, od, cubature, dtype=torch.float32)
    r.set_data(t32(ys), t32(m0[None]), t32(np.eye(5)[None]))
    out = r.prob.new_traj()
    r.pass_once((t32(nom_m), t32(nom_c)), out)
    torch.cuda.synchronize()
    assert torch.isfinite(out[0]).all() and torch.isfinite(out[1]).all()
    close(npy(out[0]), ref.mean, tol=TOL32, what="fp32 cubature mean")
    close(gram(npy(out[1])), gram(ref.chol), tol=TOL32_COV, what="fp32 cubature cov")
