"""The reverse-mode adjoint (psm_adjoint_prepare / psm_filter_smoother_vjp /
psm_filter_vjp) against the forward-mode tangent passes of the same kernels,
which the gradient tests pin to central differences of the oracle:

* transpose identity of one filter-smoother pass: <J v, u> = <v, J^T u> for
  random nominal / parameter tangents v and output cotangents u (extended and
  cubature, a ragged batch of three sequences, FP64);
* the log-likelihood's gradient (Fisher's identity) against its tangent;
* the implicit gradient of the iterated smoother (parsmooth/_utils.py:118-133)
  by the reverse sum against the forward-mode evaluation of the same rule.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

from parsmooth import _native as N  # noqa: E402
from parsmooth.data import bearings_data_batch  # noqa: E402
from parsmooth.gradient import IteratedLoglikelihood, bearings_parameterisation, _dual  # noqa: E402
from parsmooth.linearization import cubature, extended  # noqa: E402

DEV = torch.device("cuda", 0)
S1, S2 = (-1.5, 0.5), (1.0, 1.0)


def _objective(method, lengths=(60, 17, 41), n_iter=3, dtype=torch.float64, seed=3):
    rng = np.random.default_rng(seed)
    n = len(lengths)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0]) + 0.3 * rng.standard_normal((n, 5))
    truth, obs = bearings_data_batch(np.asarray(lengths), m0[:, :4], 0.01, 0.5, S1, S2, q=np.sqrt(10.0),
                                     seed=seed + 1)
    L0 = np.tile(np.linalg.cholesky(np.diag([0.5, 0.4, 0.3, 0.6, 0.2])), (n, 1, 1))
    build = bearings_parameterisation(0.01, m0, L0)
    M = int(np.sum(lengths)) + n
    init = (np.concatenate(truth).astype(np.float64) + 0.01 * rng.standard_normal((M, 5)),
            np.tile(0.2 * np.eye(5) + 0.02 * np.tril(np.ones((5, 5)), -1), (M, 1, 1)))
    obj = IteratedLoglikelihood(build, N.MODEL_COORDINATED_TURN, N.MODEL_BEARINGS, 5, 2, list(lengths),
                                np.concatenate(obs), method, n_iter, init, dtype=dtype)
    theta = torch.tensor([np.log(0.01), np.log(0.1), np.log(0.5), -1.5, 0.5, 1.0, 1.0], dtype=torch.float64)
    return obj, theta, init


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("method", [extended, cubature], ids=["extended", "cubature"])
def test_pass_vjp_is_the_transpose_of_the_tangent_pass(method):
    obj, theta, init = _objective(method)
    rng = np.random.default_rng(7)
    M = obj.prob.n_states
    nom = (torch.as_tensor(init[0], device=DEV), torch.as_tensor(init[1], device=DEV))
    v_m = torch.as_tensor(rng.standard_normal((M, 5)), device=DEV)
    v_L = torch.as_tensor(np.tril(rng.standard_normal((M, 5, 5))), device=DEV)
    dtheta = torch.as_tensor(rng.standard_normal(7))
    u_m = torch.as_tensor(rng.standard_normal((M, 5)), device=DEV)
    u_L = torch.as_tensor(np.tril(rng.standard_normal((M, 5, 5))), device=DEV)
    _, (dm, dL) = obj.jvp_pass(theta, dtheta, nom, (v_m, v_L))
    (xb_m, xb_L), gth = obj.vjp_pass(theta, nom, u_m, u_L)
    lhs = float((dm * u_m).sum() + (dL * u_L).sum())
    rhs = float((v_m * xb_m).sum() + (v_L * xb_L).sum()) + float((gth.sum(0) * dtheta.to(gth.device)).sum())
    print(f"{method.__name__}: <J v, u> = {lhs:.12e}, <v, J^T u> = {rhs:.12e}, rel {_rel(lhs, rhs):.2e}",
          flush=True)
    assert _rel(lhs, rhs) < 1e-9


@pytest.mark.parametrize("method", [extended, cubature], ids=["extended", "cubature"])
def test_loglikelihood_vjp_matches_its_tangent(method):
    """d ell = <d ell / d x, v> + <d ell / d theta, dtheta> for the filter's
    log-likelihood at a nominal (one tangent filter pass on the right)."""
    obj, theta, init = _objective(method)
    rng = np.random.default_rng(8)
    M = obj.prob.n_states
    v_m = rng.standard_normal((M, 5))
    v_L = np.tril(rng.standard_normal((M, 5, 5)))
    dtheta = rng.standard_normal(7)
    # tangent of the log-likelihood at the nominal init along (v, dtheta)
    th = torch.as_tensor(theta)
    vals, tans = torch.func.jvp(obj.build, (th,), (torch.as_tensor(dtheta),))
    models = obj._dual_models(vals, tans)
    sigma, _keep = obj._sigma(True)
    nom = (_dual(torch.as_tensor(init[0], device=DEV), torch.as_tensor(v_m, device=DEV)),
           _dual(torch.as_tensor(init[1], device=DEV), torch.as_tensor(v_L, device=DEV)))
    ell_d = torch.empty(obj.prob.n_seq, 2, dtype=torch.float64, device=DEV)
    filt = obj._new(True)
    obj._filter_ell(obj.tprob, models, sigma, obj.ys_dual, models.keep["m0"], models.keep["L0"], nom, filt, ell_d)
    lhs = float(ell_d[:, 1].sum())
    # reverse: prepare at the nominal, then psm_filter_vjp
    import ctypes
    from parsmooth.gradient import _ModelSet
    mv = _ModelSet(obj.kinds, (5, 2), vals, DEV, obj.dtype)
    sg, _k = obj._sigma(False)
    aws = obj._adjoint_workspace(obj.prob)
    nm = (torch.as_tensor(init[0], device=DEV).contiguous(), torch.as_tensor(init[1], device=DEV).contiguous())
    base = (ctypes.byref(obj.prob.struct), ctypes.byref(mv.trans), ctypes.byref(mv.obs),
            ctypes.byref(sg) if sg is not None else None, N.ptr(obj.ys), N.ptr(mv.keep["m0"]),
            N.ptr(mv.keep["L0"]), N.ptr(nm[0]), N.ptr(nm[1]))
    f, s = obj._new(False), obj._new(False)
    N.check(N.lib().psm_adjoint_prepare(*base, N.ptr(f[0]), N.ptr(f[1]), N.ptr(s[0]), N.ptr(s[1]),
                                        N.ptr(obj.prob.workspace), obj.prob.ws_bytes, N.ptr(aws), aws.numel(),
                                        N.stream_ptr()), "psm_adjoint_prepare")
    xb_m = torch.empty((M, 5), dtype=torch.float64, device=DEV)
    xb_L = torch.empty((M, 5, 5), dtype=torch.float64, device=DEV)
    nth = 3 * 25 + 10 + 10 + 2 + 4
    thb = torch.empty((obj.prob.n_seq, nth), dtype=torch.float64, device=DEV)
    N.check(N.lib().psm_filter_vjp(*base, N.ptr(s[0]), N.ptr(s[1]), N.ptr(xb_m), N.ptr(xb_L), N.ptr(thb),
                                   N.ptr(aws), aws.numel(), N.stream_ptr()), "psm_filter_vjp")
    jac = torch.func.jacfwd(obj.build)(th)
    g = obj._model_cotangents(thb, vals, jac).sum(0)
    rhs = float((xb_m.cpu().numpy() * v_m).sum() + (xb_L.cpu().numpy() * v_L).sum()) + float(g @ torch.as_tensor(dtheta))
    print(f"{method.__name__}: d ell = {lhs:.12e}, adjoint {rhs:.12e}, rel {_rel(lhs, rhs):.2e}", flush=True)
    assert _rel(lhs, rhs) < 1e-9


@pytest.mark.parametrize("method", [extended, cubature], ids=["extended", "cubature"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["fp64", "fp32"])
def test_reverse_gradient_matches_forward_mode(method, dtype):
    """The implicit gradient by the reverse sum (N + 3 adjoint passes) equals the
    forward-mode evaluation of the same rule (P (N + 3) tangent passes): FP64 to
    1e-8, FP32 primal (FP64 adjoint) against the FP32 forward mode to 1e-4."""
    obj, theta, _ = _objective(method, dtype=dtype)
    ell_r, g_r = obj.value_and_grad(theta)
    ell_f, g_f = obj.value_and_grad(theta, mode="forward")
    g_r, g_f = g_r.cpu().numpy(), g_f.cpu().numpy()
    err = np.abs(g_r - g_f).max() / np.abs(g_f).max()
    print(f"{method.__name__} {dtype}: reverse vs forward gradient rel err {err:.2e}", flush=True)
    assert torch.equal(ell_r, ell_f)
    assert err < (1e-8 if dtype == torch.float64 else 1e-4), (g_r, g_f)


def test_adjoint_rejects_fp32_problems():
    """The adjoint ABI takes FP64 problems only: an FP32 problem gets a
    workspace size of 0 and PSM_E_UNSUPPORTED, and FP32 objectives reach the
    adjoint through their FP64 twin (test_reverse_gradient_matches_forward_mode)."""
    import ctypes
    from parsmooth.gradient import _ModelSet
    obj, theta, init = _objective(extended, dtype=torch.float32)
    assert int(N.lib().psm_adjoint_workspace_bytes(ctypes.byref(obj.prob.struct))) == 0
    vals = obj.build(torch.as_tensor(theta))
    mv = _ModelSet(obj.kinds, (5, 2), vals, DEV, torch.float32)
    nm = (torch.as_tensor(init[0], dtype=torch.float32, device=DEV).contiguous(),
          torch.as_tensor(init[1], dtype=torch.float32, device=DEV).contiguous())
    f, s = obj._new(False), obj._new(False)
    aws = torch.empty(1 << 20, dtype=torch.uint8, device=DEV)
    rc = N.lib().psm_adjoint_prepare(ctypes.byref(obj.prob.struct), ctypes.byref(mv.trans), ctypes.byref(mv.obs),
                                     None, N.ptr(obj.ys), N.ptr(mv.keep["m0"]), N.ptr(mv.keep["L0"]),
                                     N.ptr(nm[0]), N.ptr(nm[1]), N.ptr(f[0]), N.ptr(f[1]), N.ptr(s[0]),
                                     N.ptr(s[1]), N.ptr(obj.prob.workspace), obj.prob.ws_bytes, N.ptr(aws),
                                     aws.numel(), N.stream_ptr())
    assert rc != 0
    assert b"PSM_F64" in N.lib().psm_last_error()
