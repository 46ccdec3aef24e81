"""Parity at the scale of the benchmark configurations.

Kelvin-Helmholtz amplifies a one-ulp difference exponentially, so short runs
on small grids do not show that the kernels stay on the oracle's rounding
over a whole run.  These tests take the benchmark problems to their end time
on grids the oracle finishes in well under a minute on the host's cores, and
run the full-size grids for a few steps:

  * 2D KH (HLL3 + WENO2 + RK3, CFL 0.4) at 256^2 to T = 2 (about 5.7k
    steps), and at 1024^2 for 20 steps;
  * 3D KH (CFL 0.3) at 48^3 to T = 1 (about 1k steps), and at 256^3 for 3
    steps;
  * a batch of three resolution groups (32^2, 64^2, 128^2) to T = 2, every
    sample against its own oracle run;
  * the surrogate network at its full shape (40 -> 256^4 -> 66, 5 members,
    minibatch 256): gradients of every member against the network oracle.

Bar: bit-identical states, step counts and times (the kernels keep the
oracle's operation order); relative L1 <= 1e-10 / 1e-8 is implied.
"""
import sys

import numpy as np
import pytest

import alsvinn_amd
from alsvinn_amd import BatchRunner, generate_parameters
from alsvinn_amd.initial_data import kelvin_helmholtz_2d, kelvin_helmholtz_3d, kh_params

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(equation, n, cfl):
    from oracle import pyoracle
    o = pyoracle.Simulator(equation, "weno2", "hll3", "rungekutta3", "periodic", n, cfl=cfl)
    h = alsvinn_amd.Simulator(equation, "weno2", "hll3", "rungekutta3", "periodic", n, cfl=cfl)
    return o, h


def _to_end(equation, n, cfl, q0, t_end):
    o, h = _pair(equation, n, cfl)
    o.set_interior(q0)
    h.set_interior(q0)
    # one writer save at T: the last step lands on T exactly
    o.add_fixed_interval(t_end)
    h.add_fixed_interval_writer(t_end)
    o.run(t_end)
    p = h.run(t_end)
    return o, h, p


def _steps(equation, n, cfl, q0, steps):
    o, h = _pair(equation, n, cfl)
    o.set_interior(q0)
    h.set_interior(q0)
    for _ in range(steps):
        o.step()
    for _ in range(steps):
        h.perform_step()
    return o, h


def test_kh2d_256_to_T2_bitwise():
    n = 256
    _say("case kh2d 256^2 to T=2")
    o, h, p = _to_end("euler2", (n, n, 1), 0.4, kelvin_helmholtz_2d(n, kh_params(2024)), 2.0)
    _say("steps", o.steps, "time", o.time)
    assert o.steps > 5000
    assert p.steps == o.steps and h.get_current_time() == o.time == 2.0
    np.testing.assert_array_equal(h.get_interior(), o.get_interior())


def test_kh2d_1024_20_steps_bitwise():
    n = 1024
    _say("case kh2d 1024^2 x 20 steps")
    o, h = _steps("euler2", (n, n, 1), 0.4, kelvin_helmholtz_2d(n, kh_params(2024)), 20)
    assert h.get_current_time() == o.time
    np.testing.assert_array_equal(h.get_interior(), o.get_interior())


def test_kh3d_48_to_T1_bitwise():
    n = 48
    _say("case kh3d 48^3 to T=1")
    o, h, p = _to_end("euler3", (n, n, n), 0.3, kelvin_helmholtz_3d(n, kh_params(2024, count=80)), 1.0)
    _say("steps", o.steps, "time", o.time)
    assert o.steps > 900
    assert p.steps == o.steps and h.get_current_time() == o.time == 1.0
    np.testing.assert_array_equal(h.get_interior(), o.get_interior())


def test_kh3d_256_3_steps_bitwise():
    n = 256
    _say("case kh3d 256^3 x 3 steps")
    o, h = _steps("euler3", (n, n, n), 0.3, kelvin_helmholtz_3d(n, kh_params(2024, count=80)), 3)
    assert h.get_current_time() == o.time
    np.testing.assert_array_equal(h.get_interior(), o.get_interior())


def test_kh_batch_three_resolutions_to_T2():
    from oracle import pyoracle
    groups = [((32, 32, 1), 4), ((64, 64, 1), 3), ((128, 128, 1), 2)]
    T = 2.0
    total = sum(m for _, m in groups)
    params = generate_parameters([("a", 40, 0.0, 1.0)], 0, total)["a"]
    _say("case KH batch 32^2/64^2/128^2 to T=2")
    br = BatchRunner("euler2", "weno2", "hll3", "rungekutta3", "periodic", groups, T, cfl=0.4, stats_saves=1)
    k = 0
    for g, (n, m) in enumerate(groups):
        br.set_kelvin_helmholtz(g, params[k:k + m])
        k += m
    steps, left = br.run()
    assert left == 0
    k = 0
    for g, (n, m) in enumerate(groups):
        finals = []
        for s in range(m):
            o = pyoracle.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", n, cfl=0.4)
            o.set_interior(kelvin_helmholtz_2d(n[0], params[k]))
            o.add_fixed_interval(T)
            o.run(T)
            pr = br.progress(g, s)
            assert pr.steps == o.steps and pr.time == o.time == T
            np.testing.assert_array_equal(br.get_state(g, s), o.get_interior())
            finals.append(o.get_interior())
            k += 1
        mean, var, cnt = br.statistics(g, 0)
        assert cnt == m
        np.testing.assert_allclose(mean, np.mean(finals, axis=0), rtol=1e-13, atol=1e-14)
        _say("group", n[:2], "steps", [br.progress(g, s).steps for s in range(m)])


def test_mlp_full_shape_gradients():
    """40 -> 256^4 -> 66, 5 members, minibatch 256 (the surrogate's shape)."""
    from oracle import mlp_oracle as mo
    from alsvinn_amd.surrogate import EnsembleMLP
    n_in, n_out, width, hidden, members, batch = 40, 66, 256, 4, 5, 256
    rng = np.random.default_rng(21)
    net = EnsembleMLP(n_in, n_out, width=width, hidden=hidden, members=members, batch=batch, seed=21)
    oracles = []
    for m in range(members):
        o = mo.MlpOracle(n_in, width, hidden, n_out)
        o.random_init(rng)
        o.set_flat(o.flat().astype(np.float32).astype(np.float64))
        net.set_params(m, o.flat())
        oracles.append(o)
    assert net.num_params == 224834
    X = rng.random((batch, n_in)).astype(np.float32)
    Y = (rng.normal(size=(batch, n_out)) * 3 + 2).astype(np.float32)
    net.fit_normalization(X, Y)
    norm = tuple(np.asarray(v, np.float32).astype(np.float64) for v in net.norm)
    _say("case mlp 40-256^4-66 x5 gradients")
    loss, g = net.gradients(X, Y)
    for m in range(members):
        lo, go, _ = oracles[m].loss_and_grads(X, Y, norm, exact=False)
        _say("member", m, "loss", loss[m], lo, "grad rel err", np.linalg.norm(g[m] - go) / np.linalg.norm(go))
        assert loss[m] == pytest.approx(lo, rel=1e-3)
        off = 0
        for (r, c) in oracles[m].shapes:
            for size in (r * c, r):
                sl = slice(off, off + size)
                assert np.linalg.norm(g[m][sl] - go[sl]) <= 2e-2 * np.linalg.norm(go[sl]) + 1e-7
                off += size


def _conserved_totals(q):
    """Sum over the cells of each conserved variable, compensated (math.fsum)."""
    import math
    return np.array([math.fsum(q[v].ravel()) for v in range(q.shape[0])])


def _check_conservation(q0, q1, steps):
    """Periodic boundaries: every conserved total stays what it was, up to
    rounding, relative to the total of |values| of that variable (before or
    after, whichever is larger).  The rounding drift grows linearly with the
    step count, not with its square root: the RK3 weights 1/3 and 3/4 and the
    dt/dx factor are rounded the same way in every cell, so each step scales
    the totals by the same 1 + O(1e-16).  The oracle drifts 7.8e-17 per step
    on KH 64^2 and 128^2; the bound is 4e-16 per step."""
    t0, t1 = _conserved_totals(q0), _conserved_totals(q1)
    scale = np.array([max(np.abs(q0[v]).sum(), np.abs(q1[v]).sum()) for v in range(q0.shape[0])])
    drift = np.abs(t1 - t0) / scale
    rel = 4e-16 * steps
    _say("conserved totals", t0, "->", t1, "relative drift", drift)
    assert np.all(drift <= rel), drift


def test_kh2d_1024_to_T2_conserves_and_lands_on_T():
    """The config-2 problem at its full size and end time (about 2.4e4
    steps; the oracle would need hours, so the check is by property): the
    last step lands on T = 2 exactly, the state stays physical, and mass,
    momentum and energy totals are conserved to rounding."""
    n = 1024
    _say("case kh2d 1024^2 to T=2 (conservation)")
    q0 = kelvin_helmholtz_2d(n, kh_params(2024))
    h = alsvinn_amd.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", (n, n, 1), cfl=0.4)
    h.set_interior(q0)
    h.add_fixed_interval_writer(2.0)
    p = h.run(2.0, chunk=512)
    _say("steps", p.steps, "time", h.get_current_time())
    assert h.get_current_time() == 2.0 and p.steps > 20000 and not p.violated
    _check_conservation(q0, h.get_interior(), p.steps)


def test_kh3d_256_to_T1_conserves_and_lands_on_T():
    """The config-4 problem at its full size and end time (about 5.8e3
    steps), by the same properties."""
    n = 256
    _say("case kh3d 256^3 to T=1 (conservation)")
    q0 = kelvin_helmholtz_3d(n, kh_params(2024, count=80))
    h = alsvinn_amd.Simulator("euler3", "weno2", "hll3", "rungekutta3", "periodic", (n, n, n), cfl=0.3)
    h.set_interior(q0)
    h.add_fixed_interval_writer(1.0)
    p = h.run(1.0, chunk=256)
    _say("steps", p.steps, "time", h.get_current_time())
    assert h.get_current_time() == 1.0 and p.steps > 4000 and not p.violated
    _check_conservation(q0, h.get_interior(), p.steps)
