"""Host-side logic of the surrogate's statistics (no GPU): the spread
calibration that decides whether the ensemble spread may route samples away
from the solver, and the fallback and control-variate estimators, with a
stand-in model whose predictions and spreads are given."""
import numpy as np

from alsvinn_amd.surrogate import optimal_beta, spread_calibration, surrogate_statistics


class FixedModel:
    """predict(X) -> (mean, spread) looked up by the first input column."""

    def __init__(self, mean, spread):
        self.mean = np.asarray(mean, dtype=np.float64)
        self.spread = np.asarray(spread, dtype=np.float64)

    def predict(self, X):
        i = np.asarray(X)[:, 0].astype(int)
        return self.mean[i], self.spread[i]


def _ids(n):
    return np.arange(n, dtype=np.float64)[:, None]


def test_calibration_rejects_a_confident_ensemble_that_is_wrong():
    n = 200
    y = 1.0 + np.random.default_rng(0).random((n, 2))
    # members agree (tiny spread) on the mean of y, not on y
    model = FixedModel(np.full((n, 2), y.mean(0)), np.full((n, 2), 1e-6))
    cov, acc = spread_calibration(model, _ids(n), y, threshold=0.05)
    assert acc == n and np.all(cov < 0.5)
    solved = []

    def solver(X):
        solved.append(len(X))
        return y[np.asarray(X)[:, 0].astype(int)]

    st = surrogate_statistics(model, _ids(n), solver=solver, threshold=0.05, X_solved=_ids(n), Y_solved=y,
                              min_coverage=0.9)
    assert st["spread_fallback_fraction"] == 0.0  # the spread test alone would accept every sample
    assert st["spread_trusted"] is False and st["fallback_fraction"] == 1.0 and solved == [n]
    np.testing.assert_allclose(st["fallback_mean"], y.mean(0))
    np.testing.assert_allclose(st["fallback_var"], y.var(0))


def test_calibration_accepts_an_accurate_ensemble():
    n = 200
    y = 1.0 + np.random.default_rng(1).random((n, 3))
    pred = y * (1 + 0.01 * np.random.default_rng(2).standard_normal((n, 3)))
    spread = np.abs(pred) * 0.02
    spread[::10] = np.abs(pred[::10]) * 0.2  # every tenth sample too uncertain
    model = FixedModel(pred, spread)
    cov, acc = spread_calibration(model, _ids(n), y, threshold=0.05)
    assert acc == n - n // 10 and np.all(cov >= 0.99)
    st = surrogate_statistics(model, _ids(n), solver=lambda X: y[np.asarray(X)[:, 0].astype(int)], threshold=0.05,
                              X_solved=_ids(n), Y_solved=y, min_coverage=0.9)
    assert st["spread_trusted"] and st["fallback_fraction"] == 0.1


def test_control_variate_with_optimal_beta_is_exact_for_a_linear_surrogate():
    rng = np.random.default_rng(3)
    f = rng.standard_normal((500, 2))
    g = 3.0 * f + 1.0
    np.testing.assert_allclose(optimal_beta(f, g), [1 / 3, 1 / 3])
    # with g an affine function of f, f - beta (g - E g) is constant: the
    # estimate of E f is exact whatever the solved subset
    n = 500
    model = FixedModel(g, np.zeros_like(g))
    st = surrogate_statistics(model, _ids(n), X_solved=_ids(n)[:40], Y_solved=f[:40])
    np.testing.assert_allclose(st["cv_mean"], f.mean(0), rtol=1e-12, atol=1e-12)
