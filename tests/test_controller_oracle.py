"""CPU checks of the controller: the FP64 network restatement in oracle/mlp.py
is pinned by central finite differences, and the sample extraction from an
optimiser log follows its documented definition."""
import numpy as np

from eplace_amd import _lib as L
from eplace_amd.controller import NP, NW, normalization, samples_from_log
from oracle import mlp


def random_params(seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, NP)
    p[mlp.oXsd:mlp.oYmu] = rng.uniform(0.5, 2.0, mlp.F)
    p[mlp.oYsd:mlp.NP] = rng.uniform(0.5, 2.0, mlp.NO)
    return p


def test_layout_matches_header():
    assert mlp.NW == NW == L.EP_CTRL_WEIGHTS
    assert mlp.NP == NP == L.EP_CTRL_PARAMS


def test_oracle_gradient_matches_finite_differences():
    p = random_params(1)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (37, mlp.F))
    Y = rng.normal(0, 1, (37, mlp.NO))
    loss, g = mlp.loss_grad(p, X, Y)
    idx = np.concatenate([rng.choice(NW, 60, replace=False), [mlp.ob1, mlp.ob2 + 5, mlp.ob3, mlp.ob3 + 1]])
    for i in idx:
        e = np.zeros(NP)
        e[i] = 1e-6
        fd = (mlp.loss_grad(p + e, X, Y)[0] - mlp.loss_grad(p - e, X, Y)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_oracle_adam_decreases_loss():
    p = random_params(3)
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (64, mlp.F))
    Y = np.stack([np.sin(X[:, 0]), X[:, 1] * 0.5], 1)
    _, losses = mlp.adam_steps(p, X, Y, 200, 3e-3)
    assert losses[-1] < 0.5 * losses[0]


def test_samples_from_log():
    rows = np.zeros((5, L.EP_LOG_COLS))
    rows[:, L.EP_LOG_FEATURES:L.EP_LOG_FEATURES + 12] = np.arange(60).reshape(5, 12)
    rows[:, L.EP_LOG_STEP] = [1.0, 2.0, 3.0, 4.0, 1.0]
    rows[:, L.EP_LOG_STEP_PROP] = [1.0, 1.0, 6.0, 2.0, 1.0]
    rows[:, L.EP_LOG_LAMBDA_BEFORE] = [1.0, 1.0, 2.0, 4.0, 3.8]
    rows[:, L.EP_LOG_LAMBDA] = [1.0, 2.0, 2.1, 3.8, 4.0]
    rows[:, L.EP_LOG_HPWL] = 350000.0 * 1000  # a large design: the floor applies in full
    X, Y = samples_from_log(rows)
    assert X.shape == (4, 12) and Y.shape == (4, 2)
    np.testing.assert_allclose(X[0], np.arange(12, 24))
    # step target: only shrinking is learned
    np.testing.assert_allclose(Y[:, 0], [0.0, np.log(0.5), 0.0, 0.0], atol=1e-7)
    # penalty floor over a window of one iteration: plain's own multiplier
    # 2.0, 1.05, 0.95, 1.0526 -> no floor where it grows by 2% or more, the
    # full floor 1.01 where it does not grow
    full = np.log(1.01 / 0.95)
    _, Y1 = samples_from_log(rows, window=1)
    np.testing.assert_allclose(Y1[:, 1], [0.0, 0.0, full, 0.0], atol=1e-7)
    # a window of two: mean growth sqrt(2.1), sqrt(1.9), 1.0 (4.0 / 4.0), and
    # the last row's own multiplier
    _, Y2 = samples_from_log(rows, window=2)
    np.testing.assert_allclose(Y2[:, 1], [0.0, 0.0, full, 0.0], atol=1e-7)
    # linear between no growth (full floor) and 2% growth (no floor)
    rows[3, L.EP_LOG_LAMBDA] = 4.0 * 1.01
    _, Y3 = samples_from_log(rows, window=1)
    np.testing.assert_allclose(Y3[2, 1], np.log((0.95 + 0.5 * 0.06) / 0.95), rtol=1e-6)
    # faded out by design scale: none at HPWL = 300 x 350000, half at the
    # geometric middle of 300 and 800 (row 3: half the growth ramp as well)
    rows[:, L.EP_LOG_HPWL] = 350000.0 * 300
    _, Y4 = samples_from_log(rows, window=1)
    np.testing.assert_allclose(Y4[:, 1], 0.0, atol=1e-7)
    rows[:, L.EP_LOG_HPWL] = 350000.0 * np.sqrt(300 * 800)
    _, Y5 = samples_from_log(rows, window=1)
    np.testing.assert_allclose(Y5[2, 1], np.log((0.95 + 0.25 * 0.06) / 0.95), rtol=1e-6)
    xmu, xsd, ymu, ysd = normalization(X, Y)
    assert np.all(xsd > 0) and np.all(ysd > 0)
