"""Every tile shape of the stage kernel, forced, gives the oracle's result bit
for bit (grids with partial tiles in every direction), the autotuner picks a
valid shape without changing the state, and the batched kernel agrees across
shapes."""
import sys

import numpy as np
import pytest

import alsvinn_amd
from alsvinn_amd import BatchRunner
from tests.cases import smooth_state

pytestmark = pytest.mark.gpu

CASES = [
    # equation, recon, flux, integrator, bc, n, cfl, nshapes
    ("burgers", "weno2", "godunov", "rungekutta3", "periodic", (203, 1, 1), 0.45, 2),
    ("euler1", "eno3", "hll3", "rungekutta2", "neumann", (131, 1, 1), 0.4, 2),
    ("euler2", "weno2", "hll3", "rungekutta3", "periodic", (70, 37, 1), 0.4, 5),
    ("euler2", "eno2", "hll", "forwardeuler", "neumann", (45, 29, 1), 0.2, 5),
    ("burgers", "none", "godunov", "rungekutta3", "periodic", (33, 66, 1), 0.4, 5),
    ("euler3", "weno2", "hll3", "rungekutta3", "periodic", (35, 19, 40), 0.3, 7),
    ("euler3", "eno3", "hll", "rungekutta2", "neumann", (17, 33, 36), 0.3, 7),
    # RK3 on every 3D shape for the other reconstructions (the last shape's
    # stages run kernels compiled per stage mode, stage 2 with its U0 ring)
    ("euler3", "eno2", "hll", "rungekutta3", "neumann", (23, 17, 37), 0.3, 7),
    ("euler3", "none", "hll3", "rungekutta3", "periodic", (18, 31, 33), 0.3, 7),
]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", CASES, ids=["-".join(map(str, c[:2])) + "-%dx%dx%d" % c[5] for c in CASES])
def test_every_tile_shape_matches_oracle(case):
    from oracle import pyoracle
    eq, recon, flux, integ, bc, n, cfl, ns = case
    dim = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
    q0 = smooth_state(eq, n[:dim], seed=11, jump=True)
    o = pyoracle.Simulator(eq, recon, flux, integ, bc, n, cfl=cfl)
    o.set_interior(q0)
    for _ in range(12):
        o.step()
    for shape in range(ns):
        _say("case", case, "shape", shape)
        h = alsvinn_amd.Simulator(eq, recon, flux, integ, bc, n, cfl=cfl)
        h.set_tile_shape(shape)
        h.set_interior(q0)
        for _ in range(12):
            h.perform_step()
        np.testing.assert_array_equal(h.get_interior(), o.get_interior())
        assert h.progress().time == o.time
    with pytest.raises(ValueError):
        alsvinn_amd.Simulator(eq, recon, flux, integ, bc, n, cfl=cfl).set_tile_shape(ns)


def test_autotune_keeps_state_and_picks_a_shape():
    from oracle import pyoracle
    n = (96, 80, 1)
    q0 = smooth_state("euler2", n[:2], seed=5)
    h = alsvinn_amd.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", n, cfl=0.4)
    h.set_interior(q0)
    h.perform_step()
    before = h.get_interior()
    p0 = h.progress()
    best, ms = h.autotune()
    _say("autotune", best, ms)
    assert 0 <= best < len(ms) == 5 and all(t > 0 for t in ms)
    np.testing.assert_array_equal(h.get_interior(), before)
    assert h.progress().time == p0.time and h.progress().steps == p0.steps
    for _ in range(5):
        h.perform_step()
    o = pyoracle.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", n, cfl=0.4)
    o.set_interior(q0)
    for _ in range(6):
        o.step()
    np.testing.assert_array_equal(h.get_interior(), o.get_interior())


def test_batch_tile_shapes_agree():
    groups = [((40, 24, 1), 3), ((20, 20, 1), 2)]
    res = []
    for shape in (0, 1, 2, 3, 4, -1):
        br = BatchRunner("euler2", "weno2", "hll3", "rungekutta3", "periodic", groups, 0.02, cfl=0.4, stats_saves=1)
        br.set_tile_shape(shape)
        for g, (nn, m) in enumerate(groups):
            for s in range(m):
                br.set_state(g, s, smooth_state("euler2", nn[:2], seed=20 + 7 * g + s))
        br.run(chunk=8)
        res.append([[br.get_state(g, s) for s in range(m)] for g, (nn, m) in enumerate(groups)] +
                   [[br.statistics(g, 0)[0] for g in range(len(groups))]])
    for r in res[1:]:
        for a, b in zip(res[0], r):
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)
