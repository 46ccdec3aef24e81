"""Surrogate pipeline on the GPU: the observables of Kelvin-Helmholtz samples
read from the batched solver (kinetic energy of u_y at equally spaced times,
S_2(h) of rho at the end) against oracle runs of the same samples."""
import sys

import numpy as np
import pytest

from alsvinn_amd.initial_data import kelvin_helmholtz_2d
from alsvinn_amd.surrogate import kh_observables, kh_parameters

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_kh_observables_match_oracle():
    from oracle import pyoracle
    n, T, times, nh = 32, 0.04, (0.02, 0.04), 9
    P = kh_parameters(100, 3)
    _say("case kh observables", n, T)
    Y = kh_observables(P, n=n, end_time=T, times=times, number_of_h=nh, chunk=2)
    assert Y.shape == (3, len(times) + nh - 1)
    for s in range(3):
        o = pyoracle.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", (n, n, 1), cfl=0.4)
        o.set_interior(kelvin_helmholtz_2d(n, P[s]))
        o.add_fixed_interval(times[0])
        for k, t in enumerate(times):
            o.run(t)
            assert o.time == t
            ke = pyoracle.kinetic_energy(o.get_interior(), 1, 1.0 / (n * n))
            assert Y[s, k] == pytest.approx(ke, rel=1e-12)
        ref = pyoracle.structure_cube(o.get_interior()[0, 0], "periodic", nh, 2.0)
        np.testing.assert_allclose(Y[s, len(times):], ref[1:], rtol=1e-10)


def test_surrogate_gains_on_a_smooth_problem():
    """Burgers with random Fourier initial data u0 = sum_i a_i sin(2 pi (i x +
    b_i)) before the shock: the structure function S_2(h) of u at t is a
    smooth function of the 6 parameters, the network learns it, and the
    control-variate estimator needs >= 10x fewer solver samples than plain
    Monte Carlo for the same error; its estimate agrees with a large plain
    reference within its own statistical error."""
    from alsvinn_amd import BatchRunner
    from alsvinn_amd.surrogate import EnsembleMLP, control_variate_efficiency, surrogate_statistics
    from alsvinn_amd.uq import generate_parameters
    N, T, nh = 128, 0.02, 17
    x = (np.arange(N) + 0.5) / N

    def solve(P):
        br = BatchRunner("burgers", "weno2", "godunov", "rungekutta3", "periodic", [((N, 1, 1), P.shape[0])], T,
                         cfl=0.45, stats_saves=0)
        br.add_fixed_interval_writer(T)
        for s in range(P.shape[0]):
            a, b = 0.1 * P[s, :3], P[s, 3:]
            u = sum(a[i] * np.sin(2 * np.pi * ((i + 1) * x + b[i])) for i in range(3))
            br.set_state(0, s, u.reshape(1, 1, 1, N))
        br.run()
        return br.structure_cube(0, 0, nh, 2.0)[:, 1:]

    P = generate_parameters([("p", 6, 0.0, 1.0)], 0, 768 + 4096)["p"]
    Xtr, Xho, Xref = P[:512], P[512:768], P[768:]
    Y = solve(P)
    Ytr, Yho, Yref = Y[:512], Y[512:768], Y[768:]
    model = EnsembleMLP(6, nh - 1, width=64, hidden=3, members=5, batch=64, seed=3)
    model.fit_normalization(Xtr, Ytr)
    model.train(Xtr, Ytr, epochs=200, lr=2e-3)
    eff = control_variate_efficiency(model, Xho, Yho)
    _say("control-variate efficiency per h", np.round(eff, 1))
    assert np.median(eff) >= 10
    Xq = generate_parameters([("p", 6, 0.0, 1.0)], 20000, 8192)["p"]
    st = surrogate_statistics(model, Xq, solver=solve, threshold=0.05, X_solved=Xho, Y_solved=Yho)
    g, _ = model.predict(Xho)
    se = np.sqrt((Yho - g).var(0) / len(Xho) + Yref.var(0) / len(Xref) + Yref.var(0) / len(Xq))
    z = np.abs(st["cv_mean"] - Yref.mean(0)) / se
    _say("cv mean z-scores", np.round(z, 2), "fallback fraction", st["fallback_fraction"])
    assert np.all(z < 5)


def test_early_stopping_keeps_each_members_best_round():
    """On a target that is pure noise the validation error can only rise as
    the network fits the training noise: every member must end at the
    parameters of its best round, whose validation error is the minimum of
    the rounds, and the optimal-beta control variate is never worse than
    plain Monte Carlo (efficiency >= 1)."""
    from alsvinn_amd.surrogate import EnsembleMLP, control_variate_efficiency
    rng = np.random.default_rng(8)
    X = rng.random((1024, 6)).astype(np.float32)
    Y = (np.sin(3 * X[:, :1]) + rng.normal(size=(1024, 3))).astype(np.float32)
    Xv = rng.random((256, 6)).astype(np.float32)
    Yv = (np.sin(3 * Xv[:, :1]) + rng.normal(size=(256, 3))).astype(np.float32)
    net = EnsembleMLP(6, 3, width=64, hidden=3, members=3, batch=64, seed=5)
    net.fit_normalization(X, Y)
    net.set_weight_decay(0.01)
    hist, best = net.train_early_stopping(X, Y, Xv, Yv, max_epochs=120, check_every=10, lr=3e-3, patience=3)
    _say("validation error per round", np.round(hist.mean(1), 4), "best rounds", best)
    pm = net.predict_members(Xv).astype(np.float64)
    sd = Yv.astype(np.float64).std(0)
    final = np.mean(((pm - Yv[None]) / sd) ** 2, axis=(1, 2))
    np.testing.assert_allclose(final, hist.min(0), rtol=1e-5)
    eff = control_variate_efficiency(net, Xv, Yv)
    assert np.all(eff >= 1.0 - 1e-12)
