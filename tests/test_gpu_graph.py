"""The whole fixed-count iteration as one launch sequence / hipGraph
(psm_iterated_smoother_graph) against the host-driven loop of single passes,
and the graph cache: replays read the buffers' current contents, graphs of
callers on their own streams run there, the least recently used graph is
dropped beyond the cache limit.

The iteration is parsmooth/methods.py:54-76 with the fixed_point loop of
parsmooth/_utils.py:136-146."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

import parsmooth  # noqa: E402
from parsmooth import FunctionalModel, MVNSqrt, MVNStandard  # noqa: E402
from parsmooth.batch import IteratedSmoother  # noqa: E402
from parsmooth.data import bearings_data  # noqa: E402
from parsmooth.linearization import cubature, extended  # noqa: E402
from parsmooth.models import make_parameters  # noqa: E402

DEV = torch.device("cuda", 0)
S1, S2 = (-1.5, 0.5), (1.0, 1.0)


def t(a):
    return torch.as_tensor(np.asarray(a, np.float64), device=DEV)


def _case(T, seed, sqrt=True):
    Q, R, obs_f, trans_f = make_parameters(0.01, 0.1, 0.5, 0.01, S1, S2)
    _, _, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, S1, S2, random_state=seed)
    mvn = MVNSqrt if sqrt else MVNStandard
    q, r = (np.linalg.cholesky(Q), np.linalg.cholesky(R)) if sqrt else (Q, R)
    trans = FunctionalModel(trans_f, mvn(t(np.zeros(5)), t(q)))
    obs = FunctionalModel(obs_f, mvn(t(np.zeros(2)), t(r)))
    x0 = mvn(t([0.1, 0.2, 1.0, 0.0, 0.0]), t(np.eye(5)))
    return t(ys), x0, trans, obs


def _host_loop_criterion(n):
    # reads the trajectories, so iterated_smoothing runs it on the host
    return lambda i, prev, cur: bool(prev[0].shape[0] > 0) and i < n


@pytest.mark.parametrize("sqrt", [True, False], ids=["sqrt", "std"])
@pytest.mark.parametrize("method", [extended, cubature], ids=["extended", "cubature"])
@pytest.mark.parametrize("with_init", [False, True], ids=["init-None", "init"])
def test_whole_loop_equals_host_loop(sqrt, method, with_init):
    """A count criterion runs the whole iteration (and the log-likelihood pass)
    as one captured graph: bit-identical to the host loop of single passes."""
    ys, x0, trans, obs = _case(300, 2, sqrt)
    init = None
    if with_init:
        m = t(np.tile([-1.0, -1.0, 6.0, 4.0, 2.0], (301, 1)))
        init = type(x0)(m, t(np.tile(np.eye(5), (301, 1, 1))))
    for _ in range(2):   # the second call replays the cached graph
        fast, ell_f = parsmooth.iterated_smoothing(ys, x0, trans, obs, method, init, criterion=lambda i, *_: i < 4,
                                                   return_loglikelihood=True)
        host, ell_h = parsmooth.iterated_smoothing(ys, x0, trans, obs, method, init, criterion=_host_loop_criterion(4),
                                                   return_loglikelihood=True)
        assert torch.equal(fast[0], host[0]) and torch.equal(fast[1], host[1])
        assert float(ell_f) == float(ell_h)


def test_graph_replay_reads_new_data():
    """The cached graph of one problem shape, replayed with other observations
    and prior, gives the host loop's result for that data."""
    parsmooth.clear_caches()
    for seed in (3, 4, 5):
        ys, x0, trans, obs = _case(257, seed)
        x0 = MVNSqrt(x0[0] + 0.01 * seed, x0[1])
        fast = parsmooth.iterated_smoothing(ys, x0, trans, obs, extended, criterion=parsmooth.FixedIterations(3))
        host = parsmooth.iterated_smoothing(ys, x0, trans, obs, extended, criterion=_host_loop_criterion(3))
        assert torch.equal(fast[0], host[0]) and torch.equal(fast[1], host[1])


def test_runner_graph_on_user_stream():
    """IteratedSmoother.run: the whole-loop graph captured and replayed on a
    caller's own stream equals the eager launch sequence on the default stream."""
    ys, x0, trans, obs = _case(1000, 6)
    r = IteratedSmoother([1000], trans, obs, cubature)
    r.set_data(ys, x0[0][None], x0[1][None])
    m_e, L_e, ell_e = (x.clone() for x in r.run(5, use_graph=False, return_loglikelihood=True))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(3):
            m_g, L_g, ell_g = r.run(5, use_graph=True, return_loglikelihood=True)
    s.synchronize()
    assert torch.equal(m_g, m_e) and torch.equal(L_g, L_e) and torch.equal(ell_g, ell_e)


def test_graph_cache_drops_least_recently_used():
    """More distinct graphs than the cache keeps (64): every call still returns
    its eager result, including calls whose graph was dropped and captured
    again."""
    parsmooth.clear_caches()
    runners = []
    for k in range(70):
        ys, x0, trans, obs = _case(16 + k, 7)
        r = IteratedSmoother([16 + k], trans, obs, extended)
        r.set_data(ys, x0[0][None], x0[1][None])
        runners.append(r)
    for sweep in range(2):
        for r in runners:
            want = [x.clone() for x in r.run(2, use_graph=False)]
            got = r.run(2, use_graph=True)
            assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1])
    parsmooth.clear_caches()
