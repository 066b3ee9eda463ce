"""ctypes wrapper of oracle/psm_oracle_c.c (TEST INFRASTRUCTURE ONLY: CPU
baseline and second checker; see that file's header)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_c.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle_c.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.oracle_iterated_smoother.restype = ctypes.c_double
        L.oracle_iterated_smoother.argtypes = ([ctypes.c_int] * 3 + [P] * 4 + [ctypes.c_int, P, ctypes.c_int, P,
                                               ctypes.c_double] + [P] * 5 + [ctypes.c_int, ctypes.c_int, P, P])
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def iterated_smoother(ys, m0, L0, init_m, trans, obs, qm, cq, rm, cr, n_iter, want_ell=False):
    """trans / obs: ("linear", A) | ("ct", dt) | ("bearings", (s1, s2)).
    Runs n_iter filter + smoother passes from the nominal means init_m."""
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    T, ny = ys.shape
    nx = int(np.asarray(m0).shape[-1])
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    kinds = {"linear": 0, "ct": 1, "bearings": 2}
    trA = f(trans[1]) if trans[0] == "linear" else None
    obA = f(obs[1]) if obs[0] == "linear" else None
    dt = float(trans[1]) if trans[0] == "ct" else 0.0
    s = f(np.concatenate([np.ravel(obs[1][0]), np.ravel(obs[1][1])])) if obs[0] == "bearings" else None
    out_m = np.zeros((T + 1, nx))
    out_L = np.zeros((T + 1, nx, nx))
    args = [f(m0), f(L0), f(init_m)]
    ell = lib().oracle_iterated_smoother(T, nx, ny, _p(ys), _p(args[0]), _p(args[1]), _p(args[2]),
                                         kinds[trans[0]], _p(trA), kinds[obs[0]], _p(obA), dt, _p(s),
                                         _p(f(qm)), _p(f(cq)), _p(f(rm)), _p(f(cr)), int(n_iter),
                                         int(want_ell), _p(out_m), _p(out_L))
    return out_m, out_L, ell
