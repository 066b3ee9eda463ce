"""Sequential passes on the device (parallel=False: psm_seq_filter /
psm_seq_smoother, one lane per sequence) against the oracle's restatement of
parsmooth/sequential/_filtering.py and _smoothing.py."""
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
from parsmooth._engine import Models, Problem, seq_filter_pass, seq_smoother_pass  # noqa: E402
from parsmooth.linearization import cubature, extended, gauss_hermite, unscented  # noqa: E402
from parsmooth.linearization._methods import resolve  # noqa: E402

from test_oracle import _nonlinear_problem  # noqa: E402

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


def _device_fn(fn):
    if isinstance(fn, O.CoordinatedTurn):
        return M.CoordinatedTurnTransition(fn.dt)
    if isinstance(fn, O.Bearings):
        return M.BearingsObservation(fn.s1, fn.s2, nin=5)
    if isinstance(fn, O.Linear):
        return M.LinearFunction(fn.A)
    raise TypeError(fn)


def _dev_mvn(x, dtype=torch.float64):
    cls = MVNSqrt if isinstance(x, O.MVNSqrt) else MVNStandard
    return cls(t64(x[0]).to(dtype), t64(x[1]).to(dtype))


def _dev_model(model, dtype=torch.float64):
    return FunctionalModel(_device_fn(model.function), _dev_mvn(model.mvn, dtype))


def _nominal(rng, traj, sqrt):
    m = traj[0] + 0.05 * rng.randn(*traj[0].shape)
    L = np.tile(0.3 * np.eye(m.shape[1]), (m.shape[0], 1, 1))
    return O.MVNSqrt(m, L) if sqrt else O.MVNStandard(m, L @ L)


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
@pytest.mark.parametrize("with_nominal", [False, True])
def test_sequential_filter_smoother_matches_oracle(method, sqrt, with_nominal):
    """Coordinated turn + bearings, T = 40: filter (with log-likelihood) and
    smoother, linearized at the running estimate or at a nominal trajectory."""
    ys, x0, tr, ob = _nonlinear_problem("ct_bearings", sqrt, T=40, seed=8)
    om = OMETH[method]
    nom = _nominal(np.random.RandomState(2), O.filtering(ys, x0, tr, ob, om, None, False), sqrt) \
        if with_nominal else None
    rf, rell = O.filtering(ys, x0, tr, ob, om, nom, False, return_loglikelihood=True)
    rs = O.smoothing(tr, rf, om, nom, False)
    nomd = _dev_mvn(nom) if with_nominal else None
    trd, obd = _dev_model(tr), _dev_model(ob)
    filt, ell = parsmooth.filtering(t64(ys), _dev_mvn(x0), trd, obd, METHODS[method], nomd, parallel=False,
                                    return_loglikelihood=True)
    close(npy(filt[0]), rf[0], 1e-9, "filtered mean")
    close(npy(filt[1]), rf[1], 1e-9, "filtered chol/cov")
    close(float(ell), rell, 1e-9, "log-likelihood")
    sm = parsmooth.smoothing(trd, filt, METHODS[method], nomd, parallel=False)
    close(npy(sm[0]), rs[0], 1e-9, "smoothed mean")
    close(npy(sm[1]), rs[1], 1e-9, "smoothed chol/cov")
    fs = parsmooth.filter_smoother(t64(ys), _dev_mvn(x0), trd, obd, METHODS[method], nomd, parallel=False)
    close(npy(fs[0]), rs[0], 1e-9, "filter_smoother mean")


@pytest.mark.parametrize("method", ["extended", "cubature"])
@pytest.mark.parametrize("sqrt", [True, False])
def test_sequential_iterated_smoothing_matches_oracle(method, sqrt):
    """iterated_smoothing(parallel=False): the first pass at the running
    estimate, then three passes at the previous smoothed trajectory, and the
    log-likelihood at the last one."""
    ys, x0, tr, ob = _nonlinear_problem("ct_bearings", sqrt, T=40, seed=9)
    crit = lambda i, *_: i < 4
    ref, rell = O.iterated_smoothing(ys, x0, tr, ob, OMETH[method], None, False, criterion=crit,
                                     return_loglikelihood=True)
    got, ell = parsmooth.iterated_smoothing(t64(ys), _dev_mvn(x0), _dev_model(tr), _dev_model(ob),
                                            METHODS[method], None, parallel=False, criterion=crit,
                                            return_loglikelihood=True)
    close(npy(got[0]), ref[0], 1e-9, "mean")
    close(npy(got[1]), ref[1], 1e-9, "chol/cov")
    close(float(ell), rell, 1e-9, "log-likelihood")


@pytest.mark.parametrize("sqrt", [True, False])
def test_sequential_ragged_batch_through_the_abi(sqrt):
    """Five sequences of different lengths in one launch (one lane each):
    every sequence equals the oracle run on it alone."""
    lengths = [1, 7, 30, 2, 19]
    rng = np.random.RandomState(4)
    probs = [_nonlinear_problem("ct_bearings", sqrt, T=T, seed=20 + s) for s, T in enumerate(lengths)]
    _, x0, tr, ob = probs[0]
    x0s = [O.MVNSqrt(p[1][0] + 0.1 * rng.randn(5), p[1][1]) if sqrt else
           O.MVNStandard(p[1][0] + 0.1 * rng.randn(5), p[1][1]) for p in probs]
    prob = Problem(5, 2, lengths, torch.float64, sqrt, DEV)
    mods = Models(prob, _dev_model(tr), _dev_model(ob), resolve(cubature))
    ys = t64(np.concatenate([p[0] for p in probs]))
    m0 = t64(np.stack([x[0] for x in x0s]))
    L0 = t64(np.stack([x[1] for x in x0s]))
    fm, fL, ell = seq_filter_pass(prob, mods, ys, m0, L0, None, None, True)
    sm, sL = seq_smoother_pass(prob, mods, fm, fL, None, None)
    offs = prob.state_offsets_host
    for s, (p, x) in enumerate(zip(probs, x0s)):
        rf, rell = O.filtering(p[0], x, tr, ob, O.cubature, None, False, return_loglikelihood=True)
        rs = O.smoothing(tr, rf, O.cubature, None, False)
        a, b = offs[s], offs[s + 1]
        close(npy(fm[a:b]), rf[0], 1e-9, f"seq {s} filtered mean")
        close(npy(fL[a:b]), rf[1], 1e-9, f"seq {s} filtered chol/cov")
        close(float(ell[s]), rell, 1e-9, f"seq {s} log-likelihood")
        close(npy(sm[a:b]), rs[0], 1e-9, f"seq {s} smoothed mean")
        close(npy(sL[a:b]), rs[1], 1e-9, f"seq {s} smoothed chol/cov")


def test_sequential_fp32_square_root_matches_oracle():
    """FP32 square-root form (linearized in FP64, FP64 re-triangularisation
    fallback): within 2e-4 of the FP64 oracle, relative to max(1, |x|)."""
    ys, x0, tr, ob = _nonlinear_problem("ct_bearings", True, T=40, seed=8)
    rf = O.filtering(ys, x0, tr, ob, O.extended, None, False)
    rs = O.smoothing(tr, rf, O.extended, None, False)
    f32 = torch.float32
    x0d = MVNSqrt(t64(x0[0]).to(f32), t64(x0[1]).to(f32))
    trd, obd = _dev_model(tr, f32), _dev_model(ob, f32)
    filt = parsmooth.filtering(t64(ys).to(f32), x0d, trd, obd, extended, None, parallel=False)
    assert filt[0].dtype == f32
    sm = parsmooth.smoothing(trd, filt, extended, None, parallel=False)
    close(npy(filt[0]), rf[0], 2e-4, "FP32 filtered mean")
    close(npy(sm[0]), rs[0], 2e-4, "FP32 smoothed mean")
    close(npy(sm[1]), rs[1], 2e-4, "FP32 smoothed chol")


def test_sequential_torch_callable_needs_a_nominal():
    """A torch-callable transition is evaluated before the launch, so it runs
    at a nominal trajectory (and matches the oracle there) and is refused at
    the running estimate."""
    ys, x0, tr, ob = _nonlinear_problem("cv_bearings", True, T=30, seed=5)
    F = tr.function.A
    Ft = t64(F)
    trd = FunctionalModel(lambda x: x @ Ft.T, _dev_mvn(tr.mvn))   # rows of points
    obd = FunctionalModel(M.BearingsObservation(ob.function.s1, ob.function.s2, nin=4), _dev_mvn(ob.mvn))
    nom = _nominal(np.random.RandomState(3), O.filtering(ys, x0, tr, ob, O.cubature, None, False), True)
    rf = O.filtering(ys, x0, tr, ob, O.cubature, nom, False)
    got = parsmooth.filtering(t64(ys), _dev_mvn(x0), trd, obd, cubature, _dev_mvn(nom), parallel=False)
    close(npy(got[0]), rf[0], 1e-9, "filtered mean")
    close(npy(got[1]), rf[1], 1e-9, "filtered chol")
    with pytest.raises(NotImplementedError):
        parsmooth.filtering(t64(ys), _dev_mvn(x0), trd, obd, cubature, None, parallel=False)


def test_sequential_sampling_default_nominal():
    """sampling(parallel=False) takes the sequential smoother's trajectory as
    the default nominal; with the same normals the samples match the oracle."""
    ys, x0, tr, ob = _nonlinear_problem("ct_bearings", True, T=40, seed=8)
    filt = O.filtering(ys, x0, tr, ob, O.extended, None, False)
    nom = O.smoothing(tr, filt, O.extended, None, False)
    S = 16
    eps = np.random.RandomState(6).randn(41, S, 5)
    ref = O.sampling(eps, tr, filt, O.extended, nom)
    got = parsmooth.sampling(t64(eps), S, _dev_model(tr), _dev_mvn(filt), extended, None, parallel=False)
    close(npy(got), ref, 1e-9, "samples")



def _batch(lengths, sqrt, seed0):
    from parsmooth.batch import IteratedSmoother
    probs = [_nonlinear_problem("ct_bearings", sqrt, T=T, seed=seed0 + s) for s, T in enumerate(lengths)]
    _, x0, tr, ob = probs[0]
    ys = np.concatenate([p[0] for p in probs])
    m0 = np.stack([p[1][0] for p in probs])
    L0 = np.stack([p[1][1] for p in probs])
    return probs, (tr, ob), ys, m0, L0, IteratedSmoother


@pytest.mark.parametrize("sqrt", [True, False])
def test_batch_runner_sequential_mode_matches_oracle(sqrt):
    """IteratedSmoother(sequential=True): every pass at the previous smoothed
    trajectory runs as the sequential recursion; the iterates and the final
    log-likelihood equal the oracle's per sequence."""
    lengths = [1, 7, 30, 2, 19, 64]
    probs, (tr, ob), ys, m0, L0, IteratedSmoother = _batch(lengths, sqrt, 40)
    r = IteratedSmoother(lengths, _dev_model(tr), _dev_model(ob), cubature, sqrt=sqrt, sequential=True)
    r.set_data(t64(ys), t64(m0), t64(L0))
    nom_m = np.concatenate([np.tile(p[1][0], (T + 1, 1)) for p, T in zip(probs, lengths)])
    nom_L = np.concatenate([np.tile(p[1][1], (T + 1, 1, 1)) for p, T in zip(probs, lengths)])
    gm, gL, ell = r.run(3, (t64(nom_m), t64(nom_L)), return_loglikelihood=True)
    offs = r.prob.state_offsets_host
    crit = lambda i, *_: i < 3
    for s, p in enumerate(probs):
        a, b = offs[s], offs[s + 1]
        nom = type(p[1])(nom_m[a:b], nom_L[a:b])
        ref, rell = O.iterated_smoothing(p[0], p[1], tr, ob, O.cubature, nom, False, criterion=crit,
                                         return_loglikelihood=True)
        close(npy(gm[a:b]), ref[0], 1e-9, f"seq {s} mean")
        close(npy(gL[a:b]), ref[1], 1e-9, f"seq {s} chol/cov")
        close(float(ell[s]), rell, 1e-9, f"seq {s} log-likelihood")


def test_autotune_chooses_sequential_for_many_short_sequences():
    """8192 sequences of 4 steps: one lane per sequence beats the chunked scan,
    and the autotuned runner gives the same iterates as the parallel one."""
    n, T = 8192, 4
    ys1, x0, tr, ob = _nonlinear_problem("ct_bearings", True, T=T, seed=3)
    from parsmooth.batch import IteratedSmoother
    mk = lambda: IteratedSmoother([T] * n, _dev_model(tr), _dev_model(ob), extended)
    r, ref_r = mk(), mk()
    for rr in (r, ref_r):
        rr.set_data(t64(np.tile(ys1, (n, 1))), t64(np.tile(x0[0], (n, 1))), t64(np.tile(x0[1], (n, 1, 1))))
    r.autotune()
    assert r.tuned["sequential"], r.tuned
    assert r.tuned["sequential_ms"] < r.tuned["parallel_ms"]
    got = r.run(3)
    want = ref_r.run(3, use_graph=False)
    close(npy(got[0]), npy(want[0]), 1e-9, "mean")
    close(npy(got[1]), npy(want[1]), 1e-9, "chol")
