"""Input generators (CPU)."""
import numpy as np
import pytest
import scipy.linalg as la

from parsmooth.data import bearings_data, bearings_data_batch, lgssm_data


def _turn(a, dt):
    G = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [0, 0, 0, a], [0, 0, -a, 0]], dtype=np.float64)
    return la.expm(G * dt)


def test_batch_simulator_follows_the_matrix_exponential():
    lengths = [5, 17, 3]
    states, obs = bearings_data_batch(lengths, np.tile([0.1, 0.2, 1.0, 0.0], (3, 1)), 0.01, 0.5,
                                      (-1.5, 0.5), (1.0, 1.0), seed=3)
    for st, ob, T in zip(states, obs, lengths):
        assert st.shape == (T + 1, 5) and ob.shape == (T, 2)
        for k in range(T):
            pred = _turn(float(st[k + 1, 4]), 0.01) @ st[k, :4].astype(np.float64)
            np.testing.assert_allclose(st[k + 1, :4], pred, atol=1e-5)


def test_single_simulator_matches_reference_recipe():
    """notebooks/bearing_data.py:154-198: same RandomState draws and recipe."""
    T = 20
    rs = np.random.RandomState(7)
    a_s = (1 + 10.0 * 0.01 * np.cumsum(rs.randn(T))).astype(np.float32)
    normals = rs.randn(T, 2).astype(np.float32)
    _, states, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0),
                                  random_state=7)
    x = np.array([0.1, 0.2, 1.0, 0.0], dtype=np.float32)
    for i, a in enumerate(a_s):
        x = _turn(float(a), 0.01).astype(np.float32) @ x
        np.testing.assert_allclose(states[i + 1, :4], x, atol=1e-6)
        np.testing.assert_allclose(ys[i, 0], np.arctan2(x[1] - 0.5, x[0] + 1.5) + 0.5 * normals[i, 0], atol=1e-6)


def test_lgssm_shapes_and_determinism():
    A = np.eye(2)
    s1, y1 = lgssm_data(np.zeros(2), A, np.eye(1, 2), np.eye(1), np.eye(2), np.zeros(2), np.zeros(1), 10, random_state=0)
    s2, y2 = lgssm_data(np.zeros(2), A, np.eye(1, 2), np.eye(1), np.eye(2), np.zeros(2), np.zeros(1), 10, random_state=0)
    assert y1.shape == (10, 1) and s1.shape == (11, 2)
    np.testing.assert_array_equal(y1, y2)


def test_population_data_shapes_and_dynamics():
    """notebooks/population_model.py:172-208: T+1 states with x0 first, T
    non-negative integer counts; without noise the states follow the Ricker map."""
    from parsmooth.data import population_data
    states, obs = population_data(np.array([1.0]), 50, np.array([[0.09]]), 10.0, random_state=3)
    assert states.shape == (51, 1) and obs.shape == (50, 1)
    assert states[0, 0] == 1.0
    assert np.all(obs >= 0) and np.all(obs == np.round(obs))
    quiet, _ = population_data(np.array([1.0]), 5, np.array([[1e-300]]), 10.0, random_state=0)
    x = 1.0
    for k in range(5):
        x = np.log(44.7) + x - np.exp(x)
        assert quiet[k + 1, 0] == pytest.approx(x, rel=1e-12)


def test_population_parameters_values():
    """The six callables of notebooks/population_model.py:132-169 at a point."""
    import torch
    from parsmooth.models import population_parameters
    mt, ct, mo, co, cht, cho = population_parameters(10.0, np.array([[0.09]]))
    x = torch.tensor([0.7], dtype=torch.float64)
    assert float(mt(x)) == pytest.approx(np.log(44.7) + 0.7 - np.exp(0.7), rel=1e-14)
    assert float(ct(x)) == pytest.approx(0.09) and float(cht(x)) == pytest.approx(0.3)
    assert float(mo(x)) == pytest.approx(10 * np.exp(0.7), rel=1e-14)
    assert float(co(x)) == pytest.approx(10 * np.exp(0.7), rel=1e-14)
    assert float(cho(x)) == pytest.approx(np.sqrt(10 * np.exp(0.7)), rel=1e-14)
    assert tuple(co(x).shape) == (1, 1) and tuple(cho(x).shape) == (1, 1)


def test_robustness_fixture_reproduces_the_published_nan_rates():
    """tests/golden/robustness_runs.npz holds the reference's 100 FP32
    robustness runs and its log-likelihood tables; the % of NaN runs per T of
    those tables are the published rates (R1_rts_fl32_*_ell.csv of the
    reference's notebooks/robustness_data, restated here)."""
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "robustness_runs.npz"), allow_pickle=False)
    published = {
        "extended_sqrt": [0] * 24,
        "extended_std": [0] * 15 + [1, 0, 1, 1, 1, 1, 1, 2, 2],
        "cubature_sqrt": [0] * 17 + [1, 0, 0, 0, 1, 2, 2],
        "cubature_std": [0, 0, 2, 4, 19] + [100] * 19,
    }
    assert d["ys"].shape == (100, 7500, 2) and d["ys"].dtype == np.float32
    assert list(d["Ts"]) == [20, 30, 40, 50, 80, 100, 200, 300, 400, 500, 1000, 1500, 2000, 2500, 3000, 3500,
                             4000, 4500, 5000, 5500, 6000, 6500, 7000, 7500]
    for key, pct in published.items():
        got = 100.0 * np.mean(np.isnan(d[f"ell_{key}"]), axis=1)
        assert np.array_equal(got, np.asarray(pct, np.float64)), (key, got)
