"""Batched Monte Carlo on the GPU against per-sample oracle runs.

Each sample of a batch must equal a single oracle run of that sample bit for
bit (state, step count, time).  Statistics are Welford on the device; they
are compared with the reference's naive sums (sum u, sum u^2 in sample
order, then /M and M2/M - mean^2) within the cancellation error of the
naive formula: |var_w - var_naive| <= 64 * M * eps * (mean^2 + var)."""
import sys

import numpy as np
import pytest

import alsvinn_amd
from alsvinn_amd import BatchRunner, generate_parameters
from alsvinn_amd.initial_data import kelvin_helmholtz_2d
from tests.cases import smooth_state

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def naive_stats(finals):
    m = len(finals)
    s1 = np.zeros_like(finals[0])
    s2 = np.zeros_like(finals[0])
    for u in finals:
        s1 = s1 + u
        s2 = s2 + u * u
    mean = s1 / m
    return mean, s2 / m - mean * mean


def check_stats(mean, var, finals):
    m = len(finals)
    mn, vn = naive_stats(finals)
    np.testing.assert_allclose(mean, mn, rtol=1e-13, atol=1e-14)
    bound = 64 * m * EPS * (mn * mn + np.abs(vn)) + 1e-300
    assert np.all(np.abs(var - vn) <= bound), float(np.max(np.abs(var - vn) / bound))


@pytest.mark.parametrize("integ", ["rungekutta3", "rungekutta2", "forwardeuler"])
def test_kh_batch_two_resolutions_matches_single_runs(integ):
    from oracle import pyoracle
    groups = [((32, 32, 1), 5), ((16, 16, 1), 3)]
    T = 0.04
    params = generate_parameters([("a", 40, 0.0, 1.0)], 0, 8)["a"]
    recon = "weno2" if integ != "forwardeuler" else "none"
    cfl = 0.4 if integ != "forwardeuler" else 0.8
    _say("case KH batch", integ, groups)
    br = BatchRunner("euler2", recon, "hll3", integ, "periodic", groups, T, cfl=cfl, stats_saves=1)
    br.set_kelvin_helmholtz(0, params[:5])
    br.set_kelvin_helmholtz(1, params[5:])
    # device KH fill equals the host initial data
    for g, (n, m) in enumerate(groups):
        for s in range(m):
            np.testing.assert_array_equal(br.get_state(g, s), kelvin_helmholtz_2d(n[0], params[(0 if g == 0 else 5) + s]))
    steps, left = br.run()
    assert left == 0
    k = 0
    for g, (n, m) in enumerate(groups):
        finals = []
        for s in range(m):
            o = pyoracle.Simulator("euler2", recon, "hll3", integ, "periodic", n, cfl=cfl)
            o.set_interior(kelvin_helmholtz_2d(n[0], params[k]))
            o.add_fixed_interval(T)  # FixedIntervalStatistics, numberOfSaves = 1
            o.run(T)
            p = br.progress(g, s)
            assert p.steps == o.steps and p.time == o.time == T
            got = br.get_state(g, s)
            np.testing.assert_array_equal(got, o.get_interior())
            finals.append(o.get_interior())
            k += 1
        mean, var, cnt = br.statistics(g, 0)
        assert cnt == m
        check_stats(mean, var, finals)
        # rho, u, p statistics
        prim = []
        for u in finals:
            rho = u[0]
            ux, uy = u[1] / rho, u[2] / rho
            pr = 0.4 * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / rho)
            prim.append(np.stack([rho, ux, uy, pr]))
        pm, pv, _ = br.statistics(g, 0, primitive=True)
        check_stats(pm, pv, prim)


def test_burgers_1d_batch_with_writer_adjuster_and_initial_statistics():
    """1D burgers batch; an extra fvm writer adjuster (10 saves) and the
    statistics recorded at t = 0 and at T."""
    from oracle import pyoracle
    N, M, T = 96, 6, 0.2
    br = BatchRunner("burgers", "weno2", "godunov", "rungekutta3", "periodic", [((N, 1, 1), M)], T, cfl=0.45,
                     stats_saves=1, stats_write_initial=True)
    br.add_fixed_interval_writer(T / 10)
    inits = []
    for s in range(M):
        q = smooth_state("burgers", (N,), seed=100 + s, jump=(s % 2 == 0))
        inits.append(q)
        br.set_state(0, s, q)
    br.run()
    finals = []
    for s in range(M):
        o = pyoracle.Simulator("burgers", "weno2", "godunov", "rungekutta3", "periodic", (N, 1, 1), cfl=0.45)
        o.set_interior(inits[s])
        o.add_fixed_interval(T, 0.0, True)
        o.add_fixed_interval(T / 10)
        o.run(T)
        np.testing.assert_array_equal(br.get_state(0, s), o.get_interior())
        assert br.progress(0, s).steps == o.steps
        finals.append(o.get_interior())
    m0, v0, c0 = br.statistics(0, 0)
    assert c0 == M
    check_stats(m0, v0, inits)
    m1, v1, c1 = br.statistics(0, 1)
    assert c1 == M
    check_stats(m1, v1, finals)


def test_constant_fields_statistics_are_parameter_moments():
    """A constant state is steady, so each sample keeps its parameter value;
    the statistics equal the parameters' mean and population variance."""
    M = 64
    x = generate_parameters([("u", 1, 0.0, 1.0)], 0, M)["u"].ravel()
    br = BatchRunner("burgers", "weno2", "godunov", "rungekutta3", "periodic", [((40, 1, 1), M)], 0.01,
                     cfl=0.45, stats_saves=1)
    for s in range(M):
        br.set_state(0, s, np.full((1, 1, 1, 40), x[s]))
    br.run()
    mean, var, cnt = br.statistics(0, 0)
    assert cnt == M
    np.testing.assert_allclose(mean.ravel(), np.full(40, x.mean()), rtol=1e-13)
    np.testing.assert_allclose(var.ravel(), np.full(40, x.var()), rtol=1e-12)


def test_violating_sample_stops_and_others_run_on():
    """A sample whose pressure goes negative stops at the step where the
    oracle raises; the batch reports it and check_constraints raises like
    the reference; a healthy sample of the same batch runs on unaffected."""
    from oracle import pyoracle
    n = (32, 32)
    bad = smooth_state("euler2", n, seed=5)
    bad[3] = bad[3] - 0.999 * (bad[3] - 0.5 * (bad[1] ** 2 + bad[2] ** 2) / bad[0])
    bad[1] += 3.0 * np.sign(np.sin(np.arange(n[0]) * 0.7))[None, :]
    good = smooth_state("euler2", n, seed=6)
    T = 0.05
    br = BatchRunner("euler2", "weno2", "hll3", "rungekutta3", "periodic", [((32, 32, 1), 3)], T, cfl=0.9)
    for s, q in enumerate([good, bad, good]):
        br.set_state(0, s, q)
    br.run()
    assert br.violations(0) == (1, 1)
    with pytest.raises(alsvinn_amd.ConstraintError):
        br.check_constraints()
    o = pyoracle.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", (32, 32, 1), cfl=0.9)
    o.set_interior(bad)
    steps = 0
    with pytest.raises(pyoracle.OracleError):
        for steps in range(1, 100):
            o.step()
    p = br.progress(0, 1)
    assert p.violated and p.steps == steps
    og = pyoracle.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", (32, 32, 1), cfl=0.9)
    og.set_interior(good)
    og.add_fixed_interval(T)  # the batch's statistics save at T
    og.run(T)
    for s in (0, 2):
        np.testing.assert_array_equal(br.get_state(0, s), og.get_interior())
        assert br.progress(0, s).time == og.time


def test_device_meanvar_of_single_solvers_matches_batch():
    """Per-sample solvers folded into the device statistics one at a time
    (the Runner's one-simulator-per-sample path) give the naive statistics of
    their states and the batch engine's statistics of the same samples."""
    n, m, T = 24, 5, 0.03
    params = generate_parameters([("a", 40, 0.0, 1.0)], 50, m)["a"]
    st = alsvinn_amd.MeanVariance("euler2", (n, n), nsaves=2)
    br = BatchRunner("euler2", "weno2", "hll3", "rungekutta3", "periodic", [((n, n, 1), m)], T, cfl=0.4,
                     stats_saves=1, stats_write_initial=True)
    br.set_kelvin_helmholtz(0, params)
    br.run()
    finals = []
    for s in range(m):
        sim = alsvinn_amd.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", (n, n, 1), cfl=0.4)
        sim.set_interior(kelvin_helmholtz_2d(n, params[s]))
        sim.add_fixed_interval_writer(T, 0.0, True)  # the statistics' save times
        sim.ack_saves()
        st.accumulate(sim, 0)
        sim.run(T)
        st.accumulate(sim, 1)
        finals.append(sim.get_interior())
        np.testing.assert_array_equal(finals[-1], br.get_state(0, s))
    for k in (0, 1):
        mean, var, cnt = st.read(k)
        bm, bv, bc = br.statistics(0, k)
        assert cnt == bc == m
        pm, pv, _ = st.read(k, primitive=True)
        bpm, bpv, _ = br.statistics(0, k, primitive=True)
        if k == 0:
            # every sample arrives at t = 0 together: the batch folds them in
            # sample order, as the sequential accumulation does
            for a, b in ((mean, bm), (var, bv), (pm, bpm), (pv, bpv)):
                np.testing.assert_array_equal(a, b)
        else:
            # samples reach T after different step counts, so the batch folds
            # them in arrival order
            for a, b in ((mean, bm), (pm, bpm)):
                np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)
            for a, b, mm in ((var, bv, bm), (pv, bpv, bpm)):
                assert np.all(np.abs(a - b) <= 64 * m * EPS * (mm * mm + np.abs(b)) + 1e-300)
    check_stats(st.read(1)[0], st.read(1)[1], finals)
    st.reset()
    assert st.read(0)[2] == 0
