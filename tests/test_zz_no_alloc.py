"""avh_step_noalloc -- a full time step on a caller's device state whose
stage buffers live in the library's static device scratch -- against the
solver's avh_step on the same problem: state, dt and constraint flag equal
bit for bit, on two calls in a row with the second on a smaller grid (the
scratch keeps the first call's contents), for each integrator; and states
larger than the scratch are rejected."""
import ctypes
import math

import numpy as np
import pytest

import alsvinn_amd
from alsvinn_amd import _abi
from alsvinn_amd.simulator import default_context
from tests.cases import smooth_state
from tests.devmem import DeviceArray

pytestmark = pytest.mark.gpu


def _noalloc_step(sim, state_dev, dt_dev, flag_dev, cap=math.inf):
    L = alsvinn_amd.load()
    ctx = default_context()
    _abi.check(L.avh_step_noalloc(ctx._h, ctypes.byref(sim.cfg), state_dev.ptr, cap, dt_dev.ptr, flag_dev.ptr, None))


@pytest.mark.parametrize("equation,recon,flux,integ,sizes", [
    ("euler2", "weno2", "hll3", "rungekutta3", [(96, 80, 1), (64, 48, 1)]),
    ("euler3", "weno2", "hll3", "rungekutta3", [(40, 36, 32), (24, 20, 16)]),
    ("euler1", "eno3", "hll", "rungekutta2", [(300, 1, 1), (130, 1, 1)]),
    ("burgers", "none", "godunov", "forwardeuler", [(512, 1, 1), (100, 1, 1)]),
])
def test_noalloc_step_equals_solver_step(equation, recon, flux, integ, sizes):
    for k, n in enumerate(sizes):
        print("case noalloc", equation, recon, integ, n, flush=True)
        q0 = smooth_state(equation, n, seed=40 + k, jump=True)
        sim = alsvinn_amd.Simulator(equation, recon, flux, integ, "periodic", n, cfl=0.4)
        sim.set_interior(q0)
        state = DeviceArray(q0)
        dt_dev = DeviceArray(np.full(1, -1.0))
        flag_dev = DeviceArray(np.full(1, 7.0))  # 8 bytes; the flag is the first int
        cap = math.inf if k == 0 else 1e-4
        if k == 1:
            sim.add_fixed_interval_writer(cap)
        for _ in range(2):
            # with the writer at interval cap and dt = cap, the solver's clip
            # candidate is (k * cap + 0) - (k - 1) * cap = cap at every step,
            # the candidate the no-allocation step's fresh clock gives
            _noalloc_step(sim, state, dt_dev, flag_dev, cap)
            dt = sim.perform_step()
            got_dt = dt_dev.download()[0]
            assert got_dt == dt
            assert np.frombuffer(flag_dev.download().tobytes()[:4], dtype=np.int32)[0] == 0
            np.testing.assert_array_equal(state.download(), sim.get_interior())


def test_noalloc_step_flags_constraint_violation():
    n = (64, 1, 1)
    q = smooth_state("euler1", n, seed=3)
    q[2, 0, 0, 10] = -5.0  # negative energy: the step's result breaks p > 0
    sim = alsvinn_amd.Simulator("euler1", "none", "hll", "forwardeuler", "periodic", n, cfl=0.4)
    state = DeviceArray(q)
    dt_dev = DeviceArray(np.zeros(1))
    flag_dev = DeviceArray(np.zeros(1))
    _noalloc_step(sim, state, dt_dev, flag_dev)
    assert np.frombuffer(flag_dev.download().tobytes()[:4], dtype=np.int32)[0] == 1


def test_noalloc_step_rejects_states_beyond_the_scratch():
    L = alsvinn_amd.load()
    ctx = default_context()
    sim = alsvinn_amd.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", (8, 8, 1), cfl=0.4)
    cfg = _abi.AvhConfig()
    ctypes.memmove(ctypes.byref(cfg), ctypes.byref(sim.cfg), ctypes.sizeof(cfg))
    cfg.n[0], cfg.n[1] = 1024, 1025  # 4 * 1024 * 1025 values > 4194304
    dummy = DeviceArray(np.zeros(8))
    with pytest.raises(_abi.InvalidArgument, match="exceeds the no-allocation scratch"):
        _abi.check(L.avh_step_noalloc(ctx._h, ctypes.byref(cfg), dummy.ptr, math.inf, None, None, None))
