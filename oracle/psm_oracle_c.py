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
        L.oracle_iterated_smoother_batch.restype = ctypes.c_int
        L.oracle_iterated_smoother_batch.argtypes = (
            [ctypes.c_int, P, ctypes.c_int, ctypes.c_int] + [P] * 5 + [ctypes.c_int, P, ctypes.c_int, P, ctypes.c_double,
                                                                    P, ctypes.c_int, P, P, P] + [P] * 4
            + [ctypes.c_int, P, P, P, ctypes.c_int])
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


_KINDS = {"linear": 0, "ct": 1, "bearings": 2}


def _model_args(model, f):
    kind = _KINDS[model[0]]
    A = f(model[1]) if model[0] == "linear" else None
    dt = float(model[1]) if model[0] == "ct" else 0.0
    s = f(np.concatenate([np.ravel(model[1][0]), np.ravel(model[1][1])])) if model[0] == "bearings" else None
    return kind, A, dt, s


def iterated_smoother_batch(lengths, ys, m0, L0, init_m, init_L, trans, obs, qm, cq, rm, cr, n_iter,
                            sigma=None, want_ell=False, threads=0):
    """Ragged batch of independent sequences (sequence q: lengths[q]
    observations, lengths[q] + 1 states), n_iter filter + smoother passes per
    sequence from the nominal (init_m, init_L) (init_L None: identity), the
    extended linearization (sigma None) or the sigma-point regression with
    sigma = (xi (npts, nx), wm, wc).  One OpenMP thread per sequence at a time
    (threads 0: the runtime default).  Returns (mean, chol, ell or None,
    threads used)."""
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)   # noqa: E731
    lengths = np.ascontiguousarray(lengths, dtype=np.int32)
    ys = f(ys)
    ny = ys.shape[-1]
    m0 = f(m0)
    nx = m0.shape[-1]
    M = int(lengths.sum()) + lengths.size
    tk, tA, dt, _ = _model_args(trans, f)
    ok, oA, _, s = _model_args(obs, f)
    if sigma is None:
        npts, xi, wm, wc = 0, None, None, None
    else:
        xi, wm, wc = (f(a) for a in sigma)
        npts = int(wm.size)
    out_m = np.zeros((M, nx))
    out_L = np.zeros((M, nx, nx))
    ell = np.zeros(lengths.size) if want_ell else None
    keep = [f(L0), f(init_m), None if init_L is None else f(init_L), f(qm), f(cq), f(rm), f(cr)]
    used = lib().oracle_iterated_smoother_batch(
        int(lengths.size), _p(lengths), nx, ny, _p(ys), _p(m0), _p(keep[0]), _p(keep[1]), _p(keep[2]), tk, _p(tA),
        ok, _p(oA), dt, _p(s), npts, _p(xi), _p(wm), _p(wc), _p(keep[3]), _p(keep[4]), _p(keep[5]), _p(keep[6]),
        int(n_iter), _p(ell), _p(out_m), _p(out_L), int(threads))
    return out_m, out_L, ell, int(used)
