"""Batch observables on the GPU against the oracle: every sample paused
exactly at an intermediate time by the end-time control, then resumed; the
kinetic energy and the structure function of each sample's state at both
times (different summation order from the oracle: relative 1e-12 / 1e-10),
and the states themselves bit for bit."""
import sys

import numpy as np
import pytest

from alsvinn_amd import BatchRunner, generate_parameters
from alsvinn_amd.initial_data import kelvin_helmholtz_2d
from tests.cases import smooth_state

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_kh_observables_at_pause_and_end():
    from oracle import pyoracle
    n, m, T = 32, 4, 0.04
    params = generate_parameters([("a", 40, 0.0, 1.0)], 0, m)["a"]
    br = BatchRunner("euler2", "weno2", "hll3", "rungekutta3", "periodic", [((n, n, 1), m)], T, cfl=0.4)
    br.add_fixed_interval_writer(T / 2)
    br.set_kelvin_helmholtz(0, params)
    oracles = []
    for s in range(m):
        o = pyoracle.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", (n, n, 1), cfl=0.4)
        o.set_interior(kelvin_helmholtz_2d(n, params[s]))
        o.add_fixed_interval(T)
        o.add_fixed_interval(T / 2)
        oracles.append(o)
    for t in (T / 2, T):
        _say("case KH observables t =", t)
        br.set_end_time(t)
        done, left = br.run()
        assert left == 0
        ke = br.kinetic_energy(0, 1)
        s2 = br.structure_cube(0, 0, 9, 2.0)
        for s, o in enumerate(oracles):
            o.run(t)
            assert br.progress(0, s).time == o.time == t
            q = o.get_interior()
            np.testing.assert_array_equal(br.get_state(0, s), q)
            assert ke[s] == pytest.approx(pyoracle.kinetic_energy(q, 1, 1.0 / (n * n)), rel=1e-12)
            ref = pyoracle.structure_cube(q[0, 0], "periodic", 9, 2.0)
            assert s2[s, 0] == 0
            np.testing.assert_allclose(s2[s, 1:], ref[1:], rtol=1e-10)


@pytest.mark.parametrize("p", [1.0, 2.0, 2.5])
def test_structure_function_1d_neumann(p):
    from oracle import pyoracle
    N, m = 97, 3
    br = BatchRunner("euler1", "eno2", "hll", "rungekutta2", "neumann", [((N, 1, 1), m)], 1.0, cfl=0.4)
    qs = [smooth_state("euler1", (N,), seed=40 + s, jump=True) for s in range(m)]
    for s in range(m):
        br.set_state(0, s, qs[s])
    s2 = br.structure_cube(0, 2, 12, p)
    ke = br.kinetic_energy(0, 0)
    for s in range(m):
        ref = pyoracle.structure_cube(qs[s][2, 0, 0], "neumann", 12, p)
        np.testing.assert_allclose(s2[s], ref, rtol=1e-10)
        assert ke[s] == pytest.approx(pyoracle.kinetic_energy(qs[s], 0, 1.0 / N), rel=1e-12)
