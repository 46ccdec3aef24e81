"""The reference's per-call plug-in interfaces on device Volumes
(avh_numerical_flux, avh_reconstruction, avh_apply_boundary,
avh_obeys_constraints, avh_substep) against the oracle's restatement of the
same reference calls, bit for bit, on the reference's ghosted per-variable
layout.  Ghost cells are filled with values that are not the boundary fill,
so a kernel that re-derived them from the interior would fail."""
import ctypes
import math

import numpy as np
import pytest

import alsvinn_amd
from alsvinn_amd import _abi
from alsvinn_amd.simulator import default_context
from tests.cases import smooth_state
from tests.devmem import DeviceVolume

pytestmark = pytest.mark.gpu

CASES = [
    ("burgers", "weno2", "godunov", (40, 1, 1)),
    ("euler1", "eno3", "hll", (37, 1, 1)),
    ("euler2", "weno2", "hll3", (37, 29, 1)),
    ("euler2", "eno2", "hll", (33, 18, 1)),
    ("euler3", "weno2", "hll3", (13, 11, 9)),
    ("euler3", "eno3", "hll3", (12, 10, 9)),
    ("euler3", "none", "hll", (10, 9, 8)),
]
GHOSTS = {"none": 1, "weno2": 2, "eno2": 2, "eno3": 3}


def _cfg(oracle, equation, recon, flux, n, integ="rungekutta3", bc="periodic"):
    h = alsvinn_amd.Simulator  # noqa: F841  (the library must be loaded)
    cfg = _abi.AvhConfig()
    o = oracle.make_config(equation, recon, flux, integ, bc, n, cfl=0.4)
    ctypes.memmove(ctypes.byref(cfg), ctypes.byref(o), ctypes.sizeof(cfg))
    return cfg, o


def _geometry(n, g):
    dim = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
    ghosts = tuple(g if d < dim else 0 for d in range(3))
    total = tuple(n[d] + 2 * ghosts[d] for d in range(3))
    return dim, ghosts, total


def _volume_state(equation, total, seed):
    """A physical state over every cell of the volume, ghosts included."""
    return np.ascontiguousarray(smooth_state(equation, total, seed=seed, jump=True))


def _ivec(v):
    return (ctypes.c_int * 3)(*v)


def _boxes(n, g, dim):
    """whole interior, ConservedSystem's interior call and its side strips"""
    tot = [n[d] + 2 * (g if d < dim else 0) for d in range(3)]
    out = [((0, 0, 0), (0, 0, 0))]
    out.append((tuple(g if d < dim else 0 for d in range(3)), tuple(-g if d < dim else 0 for d in range(3))))
    for d in range(dim):
        eb = [0, 0, 0]
        eb[d] = -tot[d] + 3 * g
        out.append(((0, 0, 0), tuple(eb)))
        st = [0, 0, 0]
        st[d] = tot[d] - 3 * g
        out.append((tuple(st), (0, 0, 0)))
    return out


@pytest.mark.parametrize("equation,recon,flux,n", CASES)
def test_numerical_flux_boxes_bitwise(oracle, equation, recon, flux, n):
    g = GHOSTS[recon]
    dim, ghosts, total = _geometry(n, g)
    cfg, ocfg = _cfg(oracle, equation, recon, flux, n)
    ctx = default_context()
    L = alsvinn_amd.load()
    U = _volume_state(equation, total, seed=len(equation) + g)
    out0 = np.full_like(U, 44.0)
    dU = DeviceVolume(U, n, ghosts)
    print("case volume flux", equation, recon, flux, n, flush=True)
    for start, end in _boxes(n, g, dim):
        dOut = DeviceVolume(out0, n, ghosts)
        ws = (ctypes.c_double * 3)()
        _abi.check(L.avh_numerical_flux(ctx._h, ctypes.byref(cfg), ctypes.byref(dU.desc), ctypes.byref(dOut.desc),
                                        _ivec(start), _ivec(end), ws, None))
        want, wsw = oracle.compute_flux_box(ocfg, U, out0, start, end)
        np.testing.assert_array_equal(dOut.download(), want, err_msg="box %s %s" % (start, end))
        np.testing.assert_array_equal(np.array(ws[:]), wsw, err_msg="wave speeds, box %s %s" % (start, end))


@pytest.mark.parametrize("equation,recon,flux,n", CASES)
def test_reconstruction_bitwise(oracle, equation, recon, flux, n):
    g = GHOSTS[recon]
    dim, ghosts, total = _geometry(n, g)
    cfg, ocfg = _cfg(oracle, equation, recon, flux, n)
    ctx = default_context()
    L = alsvinn_amd.load()
    U = _volume_state(equation, total, seed=3)
    nv = U.shape[0]
    dU = DeviceVolume(U, n, ghosts)
    indicators = [0] + ([nv - 1] if recon in ("eno2", "eno3") and nv > 1 else [])
    print("case volume reconstruction", equation, recon, n, flush=True)
    for d in range(dim):
        for ind in indicators:
            for start, end in _boxes(n, g, dim)[:2]:
                l0 = np.full_like(U, 7.0)
                r0 = np.full_like(U, -7.0)
                dl, dr = DeviceVolume(l0, n, ghosts), DeviceVolume(r0, n, ghosts)
                _abi.check(L.avh_reconstruction(ctx._h, ctypes.byref(cfg), ctypes.byref(dU.desc), d, ind,
                                                ctypes.byref(dl.desc), ctypes.byref(dr.desc), _ivec(start),
                                                _ivec(end), None))
                wl, wr = oracle.reconstruct_box(ocfg, U, d, ind, l0, r0, start, end)
                np.testing.assert_array_equal(dl.download(), wl)
                np.testing.assert_array_equal(dr.download(), wr)


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
@pytest.mark.parametrize("equation,recon,flux,n", CASES)
def test_apply_boundary_bitwise(oracle, equation, recon, flux, n, bc):
    g = GHOSTS[recon]
    dim, ghosts, total = _geometry(n, g)
    cfg, ocfg = _cfg(oracle, equation, recon, flux, n, bc=bc)
    ctx = default_context()
    L = alsvinn_amd.load()
    U = _volume_state(equation, total, seed=5)
    dU = DeviceVolume(U, n, ghosts)
    _abi.check(L.avh_apply_boundary(ctx._h, {"periodic": 0, "neumann": 1}[bc], ctypes.byref(dU.desc), None))
    np.testing.assert_array_equal(dU.download(), oracle.apply_boundary(ocfg, g, U))


def test_obeys_constraints_checks_every_cell(oracle):
    n, g = (20, 16, 1), 2
    dim, ghosts, total = _geometry(n, g)
    cfg, _ = _cfg(oracle, "euler2", "weno2", "hll3", n)
    ctx = default_context()
    L = alsvinn_amd.load()
    U = _volume_state("euler2", total, seed=9)
    ok = ctypes.c_int(-1)

    def check(state):
        dU = DeviceVolume(state, n, ghosts)
        _abi.check(L.avh_obeys_constraints(ctx._h, ctypes.byref(cfg), ctypes.byref(dU.desc), ctypes.byref(ok), None))
        return ok.value

    assert check(U) == 1
    bad = U.copy()
    bad[3, 0, 0, 1] = 0.0  # E = 0 in a ghost corner: negative pressure
    assert check(bad) == 0
    bad = U.copy()
    bad[0, 0, 7, 9] = float("nan")
    assert check(bad) == 0
    bad = U.copy()
    bad[0, 0, 5, 5] = -1.0
    assert check(bad) == 0


@pytest.mark.parametrize("integ,cap", [("rungekutta3", None), ("rungekutta2", None), ("forwardeuler", None),
                                       ("rungekutta3", 1e-4)])
@pytest.mark.parametrize("equation,recon,flux,n", [CASES[2], CASES[5], CASES[0]])
def test_substeps_reproduce_simulator_steps(oracle, equation, recon, flux, n, integ, cap):
    """Simulator::incrementSolution (alsfvm/src/simulator/Simulator.cpp:191-214)
    driven through the entry points: every substep, the boundary after it,
    the swap, the constraint check -- equal to the oracle's steps bit for
    bit, ghost cells and dt included."""
    if integ == "forwardeuler":
        recon = "none"
    g = GHOSTS[recon]
    dim, ghosts, total = _geometry(n, g)
    cfg, ocfg = _cfg(oracle, equation, recon, flux, n, integ=integ)
    ctx = default_context()
    L = alsvinn_amd.load()
    q0 = smooth_state(equation, n, seed=11, jump=True)
    o = oracle.Simulator(equation, recon, flux, integ, "periodic", n, cfl=0.4)
    o.set_interior(q0)
    if cap is not None:
        o.add_fixed_interval(cap)
    U0 = o.get_state()
    nsub = {"forwardeuler": 1, "rungekutta2": 2, "rungekutta3": 3}[integ]
    vols = [DeviceVolume(U0, n, ghosts)] + [DeviceVolume(np.zeros_like(U0), n, ghosts) for _ in range(nsub)]
    t = 0.0
    saved = 1  # FixedIntervalWriter::numberSaved without an initial save
    print("case volume substeps", equation, recon, integ, cap, flush=True)
    for step in range(3):
        dt = ctypes.c_double(0.0)
        # the writer's adjustTimestep candidate (FixedIntervalWriter.cpp:45-53)
        dt_cap = math.inf if cap is None else (saved * cap + 0.0) - t
        for s in range(nsub):
            inputs = (ctypes.POINTER(_abi.AvhVolume) * (s + 1))(*[ctypes.pointer(v.desc) for v in vols[:s + 1]])
            _abi.check(L.avh_substep(ctx._h, ctypes.byref(cfg), inputs, s, ctypes.byref(vols[s + 1].desc),
                                     dt.value, dt_cap, ctypes.byref(dt), None))
            _abi.check(L.avh_apply_boundary(ctx._h, 0, ctypes.byref(vols[s + 1].desc), None))
        vols[0], vols[nsub] = vols[nsub], vols[0]
        ok = ctypes.c_int(0)
        _abi.check(L.avh_obeys_constraints(ctx._h, ctypes.byref(cfg), ctypes.byref(vols[0].desc), ctypes.byref(ok),
                                           None))
        assert ok.value == 1
        o.step()
        assert dt.value == o.last_dt
        t = t + dt.value
        assert t == o.time
        if cap is not None and t >= saved * cap + 0.0:
            saved += 1
        np.testing.assert_array_equal(vols[0].download(), o.get_state())


def test_compiled_caller_kh_steps_match_oracle(oracle, tmp_path):
    """tests/cabi/ref_step.cpp, a C++ program built against
    include/alsvinn_hip.h, runs 5 Kelvin-Helmholtz RK3 steps through the
    Volume entry points; its dt sequence and final ghosted state equal the
    oracle's steps bit for bit."""
    import os
    import subprocess
    from alsvinn_amd.initial_data import kelvin_helmholtz_2d, kh_params
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cabi", "ref_step")
    assert os.path.exists(exe), "build tests/cabi first (__graft_entry__.build())"
    n, steps = 48, 5
    o = oracle.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", (n, n, 1), cfl=0.4)
    o.set_interior(kelvin_helmholtz_2d(n, kh_params(17)))
    U0 = o.get_state()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    U0.astype(np.float64).tofile(fin)
    print("case compiled caller kh 48^2 x 5 steps", flush=True)
    r = subprocess.run([exe, str(fin), str(fout), str(n), str(n), str(steps)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    data = np.fromfile(fout, dtype=np.float64)
    dts, state = data[:steps], data[steps:].reshape(U0.shape)
    want_dts = []
    for _ in range(steps):
        o.step()
        want_dts.append(o.last_dt)
    np.testing.assert_array_equal(dts, np.array(want_dts))
    np.testing.assert_array_equal(state, o.get_state())
