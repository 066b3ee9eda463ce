"""Whole-pass driver over a ragged batch of sequences: shapes and dtypes are
checked here, buffers are torch tensors on the HIP device, and every pass is
one call into libparsmooth_hip.so."""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native as N
from .models import TorchFunction
from ._base import MVNSqrt, MVNStandard
from .linearization._methods import Method, model_struct


# Default threshold of min|diag| / max|diag| below which an FP32 square-root
# step is recomputed in FP64 (see include/psm.h, psm_problem.fallback_ratio).
FP32_FALLBACK_RATIO = 1e-4


def set_fp32_fallback(ratio: float):
    """Set the FP64 re-triangularisation threshold for FP32 square-root runs
    created afterwards (0 disables the fallback)."""
    global FP32_FALLBACK_RATIO
    FP32_FALLBACK_RATIO = float(ratio)


def to_dev(x, dtype, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.float64)).to(device=dev, dtype=dtype).contiguous()


def compute_dtype(x0) -> torch.dtype:
    m = x0[0]
    if isinstance(m, torch.Tensor) and m.dtype in (torch.float32, torch.float64):
        return m.dtype
    if isinstance(m, np.ndarray) and m.dtype == np.float32:
        return torch.float32
    return torch.float64


class Problem:
    """Dims, dtype, form and the sequence layout of one batch.

    ``lengths[s]`` is the number of observations of sequence s; it owns
    ``lengths[s] + 1`` states (its prior state first)."""

    def __init__(self, nx: int, ny: int, lengths: Sequence[int], dtype: torch.dtype, sqrt: bool,
                 dev, chunk: int = 0, tangent_code: Optional[int] = None):
        lengths = np.asarray(lengths, dtype=np.int64).reshape(-1)
        if lengths.size == 0 or np.any(lengths < 1):
            raise ValueError("every sequence needs at least one observation")
        self.nx, self.ny = int(nx), int(ny)
        self.lengths = lengths
        self.n_seq = int(lengths.size)
        offs = np.zeros(self.n_seq + 1, dtype=np.int64)
        offs[1:] = np.cumsum(lengths + 1)
        self.state_offsets_host = offs
        self.n_states = int(offs[-1])
        self.n_obs = int(lengths.sum())
        self.dtype = dtype
        self.sqrt = sqrt
        self.dev = dev
        self.state_offsets = torch.as_tensor(offs).to(dev)
        code = N.dtype_code(dtype) if tangent_code is None else int(tangent_code)
        if not N.lib().psm_supported(code, N.FORM_SQRT if sqrt else N.FORM_STD, self.nx, self.ny):
            raise NotImplementedError(f"no compiled kernels for nx={nx}, ny={ny}, dtype code {code}")
        # FP64 re-triangularisation fallback of the FP32 square-root path
        self.fallback_counter = torch.zeros(1, dtype=torch.int64, device=dev)
        ratio = FP32_FALLBACK_RATIO if (code in (N.PSM_F32, N.PSM_F32_TANGENT) and sqrt) else 0.0
        self.struct = N.PsmProblem(code, N.FORM_SQRT if sqrt else N.FORM_STD,
                                   self.nx, self.ny, self.n_seq, int(chunk), self.n_states,
                                   self.state_offsets.data_ptr(), float(ratio), 0,
                                   self.fallback_counter.data_ptr())
        self.ws_bytes = int(N.lib().psm_workspace_bytes(ctypes.byref(self.struct)))
        if self.ws_bytes == 0:
            N.check(-1, "psm_workspace_bytes")
        self._ws = None

    # candidate schedules of the library's autotuner: chunk, and scan fan-in
    # with or without the team-cooperative upper scan levels (N.SCAN_TEAM), the
    # half-wave level 0 (N.SCAN_HALF) and the quad-cooperative level 0
    # (N.SCAN_QUAD, with team upper levels)
    TUNE_CHUNKS = (2, 4, 8, 16, 32)
    TUNE_SCAN_K = tuple(k | f for f in (0, N.SCAN_TEAM, N.SCAN_HALF, N.SCAN_TEAM | N.SCAN_HALF,
                                        N.SCAN_TEAM | N.SCAN_QUAD,
                                        N.SCAN_TEAM | N.SCAN_QUAD | N.SCAN_HALF)
                        for k in (2, 4, 8))

    def set_schedule(self, chunk: int, scan_k: int):
        """Fix the launch schedule (0 = library default) and resize the workspace."""
        self.struct.chunk = int(chunk)
        self.struct.scan_k = int(scan_k)
        nbytes = int(N.lib().psm_workspace_bytes(ctypes.byref(self.struct)))
        if nbytes == 0:
            N.check(-1, "psm_workspace_bytes")
        if self._ws is not None and self._ws.numel() < nbytes:
            self._ws = None
        self.ws_bytes = nbytes

    def autotune(self, mods, ys, m0, L0, nom_m, nom_L, filt, out):
        """Time every candidate schedule of one filter-smoother pass at the given
        nominal and keep the fastest; returns (chunk, scan_k, ms per pass)."""
        saved = (self.struct.chunk, self.struct.scan_k)
        need = 0
        for c in range(2, 65):  # the refinement stage may pick any chunk up to 64
            for k in self.TUNE_SCAN_K:
                self.struct.chunk, self.struct.scan_k = c, k
                need = max(need, int(N.lib().psm_workspace_bytes(ctypes.byref(self.struct))))
        self.struct.chunk, self.struct.scan_k = saved
        ws = torch.empty(need, dtype=torch.uint8, device=self.dev)
        bc, bk, ms = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_float(0.0)
        rc = N.lib().psm_autotune(ctypes.byref(self.struct), ctypes.byref(mods.trans), ctypes.byref(mods.obs),
                                  mods.sigma_ref(), N.ptr(ys), N.ptr(m0), N.ptr(L0), N.ptr(nom_m), N.ptr(nom_L),
                                  N.ptr(filt[0]), N.ptr(filt[1]), N.ptr(out[0]), N.ptr(out[1]), N.ptr(ws),
                                  need, N.stream_ptr(), ctypes.byref(bc), ctypes.byref(bk), ctypes.byref(ms))
        N.check(rc, "psm_autotune")
        del ws
        self.set_schedule(bc.value, bk.value)
        return bc.value, bk.value, ms.value

    @property
    def workspace(self) -> torch.Tensor:
        if self._ws is None:
            self._ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.dev)
        return self._ws

    def new_traj(self):
        m = torch.empty((self.n_states, self.nx), dtype=self.dtype, device=self.dev)
        L = torch.empty((self.n_states, self.nx, self.nx), dtype=self.dtype, device=self.dev)
        return m, L

    def default_nominal(self):
        """Zero mean and identity chol/cov at every state
        (parsmooth/parallel/_filtering.py:25-28, _smoothing.py:21-28)."""
        m = torch.zeros((self.n_states, self.nx), dtype=self.dtype, device=self.dev)
        L = torch.eye(self.nx, dtype=self.dtype, device=self.dev).expand(self.n_states, self.nx, self.nx).contiguous()
        return m, L


class Models:
    """ctypes views of the transition / observation models and the sigma table."""

    def __init__(self, prob: Problem, trans, obs, method: Method):
        self.prob = prob
        self.trans, keep_t = model_struct(trans, prob.dtype, prob.dev, prob.sqrt, nin=prob.nx)
        self.trans_fn = keep_t[3]
        self.obs, self.obs_fn = None, None
        keep_o = ()
        if obs is not None:
            self.obs, keep_o = model_struct(obs, prob.dtype, prob.dev, prob.sqrt, nin=prob.nx)
            self.obs_fn = keep_o[3]
            if self.obs_fn.nin != prob.nx or self.obs_fn.nout != prob.ny:
                raise ValueError("observation function dims do not match the problem")
        if self.trans_fn.nin != prob.nx or self.trans_fn.nout != prob.nx:
            raise ValueError("transition function dims do not match the state dimension")
        self.sigma, keep_s = method.sigma(prob.nx, prob.dtype, prob.dev)
        self._keep = (keep_t, keep_o, keep_s)
        self._tables = {}

    @property
    def tabulated(self) -> bool:
        return isinstance(self.trans_fn, TorchFunction) or isinstance(self.obs_fn, TorchFunction)

    def prepare(self, nom_m, nom_L, obs: bool = True):
        """Evaluate torch-callable models at the linearization points of this
        nominal trajectory (row i = state i) into persistent tables."""
        if not self.tabulated:
            return
        xi = self._keep[2][0] if self.sigma is not None else None
        L = nom_L
        if xi is not None and not self.prob.sqrt:
            L = torch.linalg.cholesky(nom_L)   # the kernels factor the covariance the same way
        for name, fn, struct in (("trans", self.trans_fn, self.trans),
                                 ("obs", self.obs_fn, self.obs if obs else None)):
            if struct is None or not isinstance(fn, TorchFunction):
                continue
            width = fn.row_width(xi.shape[0] if xi is not None else 0)
            buf = self._tables.get(name)
            if buf is None or buf.shape != (self.prob.n_states, width):
                buf = torch.empty((self.prob.n_states, width), dtype=self.prob.dtype, device=self.prob.dev)
                self._tables[name] = buf
            fn.tabulate(nom_m.reshape(-1, fn.nin), L.reshape(-1, fn.nin, fn.nin), xi, out=buf)
            struct.params = buf.data_ptr()

    def sigma_ref(self):
        return ctypes.byref(self.sigma) if self.sigma is not None else None


def filter_pass(prob: Problem, mods: Models, ys, m0, L0, nom_m, nom_L, want_ell: bool,
                out=None):
    out_m, out_L = out if out is not None else prob.new_traj()
    ell = torch.empty(prob.n_seq, dtype=torch.float64, device=prob.dev) if want_ell else None
    mods.prepare(nom_m, nom_L)
    rc = N.lib().psm_filter(ctypes.byref(prob.struct), ctypes.byref(mods.trans), ctypes.byref(mods.obs),
                            mods.sigma_ref(), N.ptr(ys), N.ptr(m0), N.ptr(L0), N.ptr(nom_m),
                            N.ptr(nom_L), N.ptr(out_m), N.ptr(out_L), N.ptr(ell),
                            N.ptr(prob.workspace), prob.ws_bytes, N.stream_ptr())
    N.check(rc, "psm_filter")
    return out_m, out_L, ell


def smoother_pass(prob: Problem, mods: Models, fm, fL, nom_m, nom_L, out=None):
    out_m, out_L = out if out is not None else prob.new_traj()
    mods.prepare(nom_m, nom_L, obs=False)
    rc = N.lib().psm_smoother(ctypes.byref(prob.struct), ctypes.byref(mods.trans), mods.sigma_ref(),
                              N.ptr(fm), N.ptr(fL), N.ptr(nom_m), N.ptr(nom_L), N.ptr(out_m),
                              N.ptr(out_L), N.ptr(prob.workspace), prob.ws_bytes, N.stream_ptr())
    N.check(rc, "psm_smoother")
    return out_m, out_L


def filter_smoother_pass(prob: Problem, mods: Models, ys, m0, L0, nom_m, nom_L, filt, out,
                         use_graph: bool = False):
    fm, fL = filt
    om, oL = out
    mods.prepare(nom_m, nom_L)
    rc = N.lib().psm_filter_smoother(ctypes.byref(prob.struct), ctypes.byref(mods.trans),
                                     ctypes.byref(mods.obs), mods.sigma_ref(), N.ptr(ys), N.ptr(m0),
                                     N.ptr(L0), N.ptr(nom_m), N.ptr(nom_L), N.ptr(fm), N.ptr(fL),
                                     N.ptr(om), N.ptr(oL), N.ptr(prob.workspace), prob.ws_bytes,
                                     int(bool(use_graph)), N.stream_ptr())
    N.check(rc, "psm_filter_smoother")
    return om, oL


def _seq_models_ready(mods: Models, nom_m, nom_L, obs: bool = True):
    if nom_m is None:
        if mods.tabulated:
            raise NotImplementedError(
                "parallel=False without a nominal trajectory linearizes at the running estimate, "
                "which a torch-callable or conditional-moments model cannot be evaluated at ahead "
                "of the recursion; pass a nominal trajectory or use a parsmooth.models function")
    else:
        mods.prepare(nom_m, nom_L, obs=obs)


def seq_filter_pass(prob: Problem, mods: Models, ys, m0, L0, nom_m, nom_L, want_ell: bool,
                    out=None):
    """Sequential filter, one lane per sequence (parallel=False).  nom_m None:
    the transition is linearized at the filtered and the observation at the
    predicted state."""
    out_m, out_L = out if out is not None else prob.new_traj()
    ell = torch.empty(prob.n_seq, dtype=torch.float64, device=prob.dev) if want_ell else None
    _seq_models_ready(mods, nom_m, nom_L)
    rc = N.lib().psm_seq_filter(ctypes.byref(prob.struct), ctypes.byref(mods.trans),
                                ctypes.byref(mods.obs), mods.sigma_ref(), N.ptr(ys), N.ptr(m0),
                                N.ptr(L0), N.ptr(nom_m), N.ptr(nom_L), N.ptr(out_m), N.ptr(out_L),
                                N.ptr(ell), N.stream_ptr())
    N.check(rc, "psm_seq_filter")
    return out_m, out_L, ell


def seq_smoother_pass(prob: Problem, mods: Models, fm, fL, nom_m, nom_L, out=None):
    """Sequential (Rauch-Tung-Striebel) smoother, one lane per sequence.
    nom_m None: the transition out of state k is linearized at the smoothed
    state k+1."""
    out_m, out_L = out if out is not None else prob.new_traj()
    _seq_models_ready(mods, nom_m, nom_L, obs=False)
    rc = N.lib().psm_seq_smoother(ctypes.byref(prob.struct), ctypes.byref(mods.trans), mods.sigma_ref(),
                                  N.ptr(fm), N.ptr(fL), N.ptr(nom_m), N.ptr(nom_L), N.ptr(out_m),
                                  N.ptr(out_L), N.stream_ptr())
    N.check(rc, "psm_seq_smoother")
    return out_m, out_L


def mean_sq_diff(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """mean((a - b)^2) on the device, accumulated in FP64 (one-element tensor)."""
    out = torch.empty(1, dtype=torch.float64, device=a.device)
    a = a.contiguous()
    b = b.to(a.dtype).contiguous()
    if a.shape != b.shape:
        raise ValueError("mean_sq_diff needs equal shapes")
    rc = N.lib().psm_mean_sq_diff(N.dtype_code(a.dtype), a.numel(), N.ptr(a), N.ptr(b), N.ptr(out),
                                  N.stream_ptr())
    N.check(rc, "psm_mean_sq_diff")
    return out
