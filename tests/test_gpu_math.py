"""The FP64 reciprocal and square root of the factorisations (hardware
estimates refined in software) against correctly rounded results, through
psm_math_check.

The reference divides and takes square roots through XLA (correctly rounded);
the library's FP64 factorisations use v_rcp_f64 + two Newton steps and
v_rsq_f64 + one Goldschmidt step + a residual correction.  Bar: at most 1 ulp
everywhere and correctly rounded for all but 1e-4 of the values.  A single
Newton step (~2^-48 relative) or a dropped correction fails it.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

from parsmooth import _native as N  # noqa: E402

DEV = torch.device("cuda", 0)


def _run(x):
    xt = torch.as_tensor(x, dtype=torch.float64, device=DEV)
    r, q = torch.empty_like(xt), torch.empty_like(xt)
    N.check(N.lib().psm_math_check(len(x), N.ptr(xt), N.ptr(r), N.ptr(q), N.stream_ptr()), "psm_math_check")
    torch.cuda.synchronize()
    return r.cpu().numpy(), q.cpu().numpy()


def _ulps(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def _inputs():
    rng = np.random.default_rng(0)
    n = 1 << 20
    specials = [1.0, 2.0, 0.5, 3.0, 4.0, 0.25, 1.0 + 2.0 ** -52, 1.0 - 2.0 ** -53, 2.0 - 2.0 ** -51,
                1e-300, 1e300, 2.0 ** -1000, 2.0 ** 1000, 7.0, 1.0 / 3.0]
    return np.concatenate([rng.uniform(1e-3, 1e3, n // 2), 10.0 ** rng.uniform(-300, 300, n // 2),
                           np.asarray(specials)])


def test_reciprocal_correctly_rounded():
    x = _inputs()
    r, _ = _run(x)
    u = _ulps(r, 1.0 / x)
    assert u.max() <= 1.0, f"reciprocal max error {u.max():.2f} ulp"
    assert np.mean(u > 0) <= 1e-4, f"reciprocal not correctly rounded for {np.mean(u > 0):.2e} of the values"
    # powers of two have exact reciprocals
    p2 = 2.0 ** np.arange(-1000, 1001, 37, dtype=np.float64)
    rp, _ = _run(p2)
    assert np.all(rp == 1.0 / p2), "reciprocal of a power of two"


def test_square_root_correctly_rounded():
    x = _inputs()
    _, q = _run(x)
    u = _ulps(q, np.sqrt(x))
    assert u.max() <= 1.0, f"square root max error {u.max():.2f} ulp"
    assert np.mean(u > 0) <= 1e-4, f"square root not correctly rounded for {np.mean(u > 0):.2e} of the values"
    # perfect squares are exact, and sqrt(0) = 0
    sq = np.arange(1, 4097, dtype=np.float64) ** 2
    _, qs = _run(np.concatenate([sq, [0.0]]))
    assert np.all(qs[:-1] == np.arange(1, 4097)), "sqrt of perfect squares"
    assert qs[-1] == 0.0


def test_math_check_rejects_bad_arguments():
    lib = N.lib()
    assert lib.psm_math_check(-1, None, None, None, None) != 0
    assert lib.psm_math_check(4, None, None, None, None) != 0
    assert lib.psm_math_check(0, None, None, None, None) == 0
    assert b"psm_math_check" in lib.psm_last_error()
