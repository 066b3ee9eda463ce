"""psm_filter_smoother_noalloc (the entry point with the library-owned static
workspace) against psm_filter_smoother with a caller-supplied workspace,
through the C ABI: two calls in a row, the second on a smaller problem, must
both equal the existing entry point's result bit for bit (same kernels, same
schedule).  The refusal of problems beyond the static buffer is checked on the
CPU (tests/test_abi.py)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

from parsmooth import FunctionalModel, MVNSqrt  # noqa: E402
from parsmooth import _native as N  # noqa: E402
from parsmooth._engine import Models, Problem  # noqa: E402
from parsmooth.data import bearings_data  # noqa: E402
from parsmooth.linearization import cubature, extended  # noqa: E402
from parsmooth.linearization._methods import resolve  # noqa: E402
from parsmooth.models import make_parameters  # noqa: E402

DEV = torch.device("cuda", 0)
S1, S2 = (-1.5, 0.5), (1.0, 1.0)


def _setup(lengths, dtype, method, seed):
    Q, R, obs_f, trans_f = make_parameters(0.01, 0.1, 0.5, 0.01, S1, S2)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float64), device=DEV).to(dtype)   # noqa: E731
    trans = FunctionalModel(trans_f, MVNSqrt(t(np.zeros(5)), t(np.linalg.cholesky(Q))))
    obs = FunctionalModel(obs_f, MVNSqrt(t(np.zeros(2)), t(np.linalg.cholesky(R))))
    prob = Problem(5, 2, lengths, dtype, True, DEV)
    mods = Models(prob, trans, obs, resolve(method))
    ys = torch.cat([t(bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, S1, S2,
                                    random_state=seed + k)[2]) for k, T in enumerate(lengths)])
    m0 = t(np.tile([0.1, 0.2, 1.0, 0.0, 0.0], (len(lengths), 1)))
    L0 = t(np.tile(np.eye(5), (len(lengths), 1, 1)))
    nom = (t(np.tile([-1.0, -1.0, 6.0, 4.0, 2.0], (prob.n_states, 1))), t(np.tile(np.eye(5), (prob.n_states, 1, 1))))
    return prob, mods, ys, m0, L0, nom


def _args(prob, mods, ys, m0, L0, nom, filt, out):
    return (ctypes.byref(prob.struct), ctypes.byref(mods.trans), ctypes.byref(mods.obs), mods.sigma_ref(),
            N.ptr(ys), N.ptr(m0), N.ptr(L0), N.ptr(nom[0]), N.ptr(nom[1]), N.ptr(filt[0]), N.ptr(filt[1]),
            N.ptr(out[0]), N.ptr(out[1]))


def _both(lengths, dtype, method, seed):
    prob, mods, ys, m0, L0, nom = _setup(lengths, dtype, method, seed)
    print(f"noalloc case: lengths {lengths} {dtype} {method.__name__}", flush=True)
    f_ref, o_ref = prob.new_traj(), prob.new_traj()
    rc = N.lib().psm_filter_smoother(*_args(prob, mods, ys, m0, L0, nom, f_ref, o_ref), N.ptr(prob.workspace),
                                     prob.ws_bytes, 0, N.stream_ptr())
    N.check(rc, "psm_filter_smoother")
    f_got, o_got = prob.new_traj(), prob.new_traj()
    rc = N.lib().psm_filter_smoother_noalloc(*_args(prob, mods, ys, m0, L0, nom, f_got, o_got), N.stream_ptr())
    N.check(rc, "psm_filter_smoother_noalloc")
    torch.cuda.synchronize()
    for a, b in zip(f_got + o_got, f_ref + o_ref):
        assert torch.equal(a, b)


@pytest.mark.parametrize("dtype,method", [(torch.float64, extended), (torch.float32, cubature)],
                         ids=["fp64-extended", "fp32-cubature"])
def test_noalloc_matches_workspace_entry_point(dtype, method):
    """Two calls in a row: a ragged batch, then a smaller single sequence (the
    buffer keeps the first call's contents, which the second must not see)."""
    _both([20000, 3001, 777], dtype, method, 1)
    _both([500], dtype, method, 9)
