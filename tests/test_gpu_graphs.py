"""Steps replayed from captured hipGraphs equal ordinary launches bit for
bit: single solvers of every integrator (with a writer adjuster whose save
times fall inside a captured chunk), graph replays interleaved with eager
steps, and the batched engine."""
import sys

import numpy as np
import pytest

import alsvinn_amd
from alsvinn_amd import BatchRunner, generate_parameters
from alsvinn_amd.initial_data import kelvin_helmholtz_2d
from tests.cases import smooth_state

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _solver(integ, graph_steps, n=(48, 40, 1)):
    s = alsvinn_amd.Simulator("euler2", "weno2", "hll3", integ, "periodic", n, cfl=0.4 if integ != "forwardeuler" else 0.2)
    s.set_graph_steps(graph_steps)
    s.set_interior(smooth_state("euler2", n[:2], seed=7))
    s.add_fixed_interval_writer(0.0037)
    return s


@pytest.mark.parametrize("integ", ["rungekutta3", "rungekutta2", "forwardeuler"])
def test_graph_chunks_equal_ordinary_launches(integ):
    _say("case graphs", integ)
    ref = _solver(integ, 0)
    ref.enqueue_steps(37)
    ref.synchronize()
    for gs in (4, 7):
        g = _solver(integ, gs)
        g.enqueue_steps(37)
        g.synchronize()
        np.testing.assert_array_equal(g.get_interior(), ref.get_interior())
        assert g.progress().steps == ref.progress().steps == 37
        assert g.progress().time == ref.progress().time
        assert g.progress().saves_pending == ref.progress().saves_pending


def test_graph_replays_interleaved_with_eager_steps():
    ref = _solver("rungekutta3", 0)
    g = _solver("rungekutta3", 5)
    for rep in range(4):
        ref.enqueue_steps(5)
        g.enqueue_steps(5)          # one replay of the captured chunk
        ref.perform_step()
        g.perform_step()            # eager step between replays
        ref.synchronize()
        g.synchronize()
        np.testing.assert_array_equal(g.get_interior(), ref.get_interior())
        assert g.progress().steps == ref.progress().steps == 6 * (rep + 1)


def test_graph_run_to_end_time_stops_on_device():
    ref = _solver("rungekutta3", 0)
    g = _solver("rungekutta3", 16)
    ref.run(0.01, chunk=16)
    g.run(0.01, chunk=16)
    # no save time at 0.01, so the last step overshoots it as in the reference
    assert g.progress().time == ref.progress().time >= 0.01
    assert ref.progress().time - ref.progress().last_dt < 0.01
    assert g.progress().steps == ref.progress().steps
    np.testing.assert_array_equal(g.get_interior(), ref.get_interior())


def test_batch_graphs_equal_ordinary_launches():
    groups = [((32, 32, 1), 3), ((16, 16, 1), 2)]
    params = generate_parameters([("a", 40, 0.0, 1.0)], 0, 5)["a"]
    out = []
    for gs in (0, 3):
        br = BatchRunner("euler2", "weno2", "hll3", "rungekutta3", "periodic", groups, 0.03, cfl=0.4, stats_saves=2)
        br.set_graph_steps(gs)
        br.set_kelvin_helmholtz(0, params[:3])
        br.set_kelvin_helmholtz(1, params[3:])
        br.run(chunk=10)
        out.append(br)
    a, b = out
    for g, (n, m) in enumerate(groups):
        for s in range(m):
            np.testing.assert_array_equal(a.get_state(g, s), b.get_state(g, s))
            assert a.progress(g, s).steps == b.progress(g, s).steps
        for k in range(2):
            ma, va, ca = a.statistics(g, k)
            mb, vb, cb = b.statistics(g, k)
            assert ca == cb == m
            np.testing.assert_array_equal(ma, mb)
            np.testing.assert_array_equal(va, vb)
