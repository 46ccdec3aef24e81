"""HIP path vs the CPU oracle on identical seeded inputs.

Bar (stated per test): relative L1 per conserved variable <= 1e-10 on smooth
runs and <= 1e-8 past shocks; the kernels are built to be bit-equal to the
oracle, which the *_bitwise tests assert where the run is short enough that
the oracle itself is the reference-structure CPU path.
"""
import itertools
import sys

import numpy as np
import pytest

import alsvinn_amd
from tests.cases import alternating_euler3, kh2d, rel_l1, smooth_state

pytestmark = pytest.mark.gpu

SMOOTH_TOL = 1e-10
SHOCK_TOL = 1e-8


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(oracle, equation, recon, flux, integ, bc, n, cfl, lower=(0, 0, 0), upper=(1, 1, 1)):
    n = tuple(n) + (1,) * (3 - len(n))
    o = oracle.Simulator(equation, recon, flux, integ, bc, n, lower, upper, 1.4, cfl)
    h = alsvinn_amd.Simulator(equation, recon, flux, integ, bc, n, lower, upper, 1.4, cfl)
    return o, h


def run_both(oracle, equation, recon, flux, integ, bc, n, cfl, q0, steps=None, t_end=None, saves=None):
    o, h = make_pair(oracle, equation, recon, flux, integ, bc, n, cfl)
    o.set_interior(q0)
    h.set_interior(q0)
    if saves:
        o.add_fixed_interval(t_end / saves)
        h.add_fixed_interval_writer(t_end / saves)
    if t_end is not None:
        o.run(t_end, max_steps=steps if steps else 1 << 62)
        p = h.run(t_end, max_steps=steps if steps else 1 << 62)
        assert p.steps == o.steps
    else:
        for _ in range(steps):
            o.step()
        for _ in range(steps):
            h.perform_step()
    return o, h


def test_burgers_sine_config1_full_run():
    """Burgers sine N=1024, WENO2 + Godunov + RK3, CFL 0.45, periodic, to T=0.5
    past the shock (t=1/(2 pi)), landing on T with a one-save writer."""
    from oracle import pyoracle
    N = 1024
    x = (np.arange(N) + 0.5) / N
    q0 = np.sin(2 * np.pi * x).reshape(1, 1, 1, N)
    _say("case burgers N=1024 T=0.5")
    o, h = run_both(pyoracle, "burgers", "weno2", "godunov", "rungekutta3", "periodic", (N,), 0.45, q0,
                    t_end=0.5, saves=1)
    uo = o.get_interior()
    uh = h.get_interior()
    assert h.get_current_time() == o.time == 0.5
    err = rel_l1(uh, uo)
    _say("steps", o.steps, "rel L1", err, "max abs", np.abs(uh - uo).max())
    assert np.all(err <= SHOCK_TOL)
    np.testing.assert_array_equal(uh, uo)
    # ghosted export equals the oracle's ghosted volume (periodic fill)
    np.testing.assert_array_equal(h.get_state(), o.get_state())


def test_kh2d_64_rk3_bitwise():
    from oracle import pyoracle
    n = 64
    q0 = kh2d(n)
    _say("case kh2d 64 x 40 steps")
    o, h = run_both(pyoracle, "euler2", "weno2", "hll3", "rungekutta3", "periodic", (n, n), 0.4, q0, steps=40)
    err = rel_l1(h.get_interior(), o.get_interior())
    _say("rel L1", err)
    assert np.all(err <= SHOCK_TOL)
    np.testing.assert_array_equal(h.get_interior(), o.get_interior())
    np.testing.assert_array_equal(h.get_state(), o.get_state())
    assert h.progress().time == o.time


def test_euler3_16_rk3_bitwise():
    from oracle import pyoracle
    n = (16, 16, 16)
    q0 = smooth_state("euler3", n, seed=3, jump=True)
    _say("case euler3 16^3 x 4 steps")
    o, h = run_both(pyoracle, "euler3", "weno2", "hll3", "rungekutta3", "periodic", n, 0.3, q0, steps=4)
    np.testing.assert_array_equal(h.get_interior(), o.get_interior())


def test_compute_flux_alternating_states_bitwise():
    """CUDANumericalFluxTest::CompareAgainstCPU shape: euler3, alternating
    states, neumann, no reconstruction; net flux and wave speeds bit-equal."""
    from oracle import pyoracle
    for flux in ("hll", "hll3"):
        n = (10, 10, 10)
        q0 = alternating_euler3(n)
        o, h = make_pair(pyoracle, "euler3", "none", flux, "forwardeuler", "neumann", n, 0.5)
        o.set_interior(q0)
        h.set_interior(q0)
        U = o.get_state()
        out0 = np.full_like(U, 44.0)
        out_o, ws_o = pyoracle.compute_flux(o.cfg, U, out0)
        out_h, ws_h = h.compute_flux()
        np.testing.assert_array_equal(ws_h, ws_o)
        np.testing.assert_array_equal(out_h, o.interior(out_o))


def _extreme_state(equation, n, seed=7):
    """A smooth euler state with cells whose operands leave the range of the
    shared-reciprocal quotients: densities 1e-40 and 1e35, momenta 1e-130
    (nonzero below 2^-390), exact +0 and -0 momenta, and density jumps that
    drive a WENO2 weight below 2^-400."""
    q = smooth_state(equation, n, seed=seed)
    nsd = q.shape[0] - 2
    flat = q.reshape(q.shape[0], -1)
    ncell = flat.shape[1]
    rng = np.random.default_rng(seed)
    picks = rng.permutation(ncell)
    k = max(1, ncell // 12)
    tiny_rho, huge_rho, tiny_m, zero_m = np.split(picks[:4 * k], 4)
    flat[0, tiny_rho] = 1e-40
    flat[1:1 + nsd, tiny_rho] = 0.0
    flat[-1, tiny_rho] = 1e-40
    flat[0, huge_rho] = 1e35
    flat[-1, huge_rho] = flat[-1, huge_rho] * 1e35
    flat[1, tiny_m] = 1e-130 * np.where(rng.random(k) < 0.5, -1.0, 1.0)
    flat[1:1 + nsd, zero_m] = np.where(rng.random((nsd, k)) < 0.5, -0.0, 0.0)
    return flat.reshape(q.shape)


@pytest.mark.parametrize("equation,n", [("euler2", (40, 20)), ("euler1", (70,))])
@pytest.mark.parametrize("recon", ["weno2", "none"])
@pytest.mark.parametrize("flux", ["hll", "hll3"])
def test_shared_reciprocal_guards_bitwise(equation, n, recon, flux):
    """Interfaces whose operands leave the range of the shared-reciprocal
    quotients are redone with plain divisions: net flux and wave speeds keep
    every bit of the oracle's, signed zeros included."""
    from oracle import pyoracle
    _say("shared-reciprocal guards", equation, n, recon, flux)
    q0 = _extreme_state(equation, n)
    o, h = make_pair(pyoracle, equation, recon, flux, "forwardeuler", "periodic", n, 0.4)
    o.set_interior(q0)
    h.set_interior(q0)
    U = o.get_state()
    out_o, ws_o = pyoracle.compute_flux(o.cfg, U, np.zeros_like(U))
    out_h, ws_h = h.compute_flux()
    assert np.array_equal(np.asarray(ws_h).view(np.uint64), np.asarray(ws_o).view(np.uint64))
    ref = np.ascontiguousarray(o.interior(out_o))
    assert np.array_equal(np.ascontiguousarray(out_h).view(np.uint64), ref.view(np.uint64))


GRIDS = {
    "burgers": [(70,), (24, 20), (10, 9, 12)],
    "euler1": [(70,)],
    "euler2": [(70,), (40, 20)],
    "euler3": [(40,), (20, 18), (10, 9, 12)],
}
FLUX_OF = {"burgers": ["godunov"], "euler1": ["hll", "hll3"], "euler2": ["hll", "hll3"], "euler3": ["hll", "hll3"]}


def _matrix():
    out = []
    for eq, grids in GRIDS.items():
        for n in grids:
            for recon in ("none", "weno2", "eno2", "eno3"):
                for flux in FLUX_OF[eq]:
                    for integ in ("forwardeuler", "rungekutta2", "rungekutta3"):
                        for bc in ("periodic", "neumann"):
                            out.append((eq, n, recon, flux, integ, bc))
    return out


def test_variant_matrix_bitwise():
    """Every equation x grid dimension x reconstruction x flux x integrator x
    boundary, 3 steps on small grids with partial tiles."""
    from oracle import pyoracle
    failures = []
    cases = _matrix()
    for i, (eq, n, recon, flux, integ, bc) in enumerate(cases):
        q0 = smooth_state(eq, n, seed=i, jump=(i % 2 == 0))
        _say("case", i, eq, n, recon, flux, integ, bc)
        cfl = 0.3 if recon != "none" else 0.45
        o, h = run_both(pyoracle, eq, recon, flux, integ, bc, n, cfl, q0, steps=3)
        a, b = h.get_state(), o.get_state()
        if not np.array_equal(a, b):
            failures.append((eq, n, recon, flux, integ, bc, float(np.abs(a - b).max()),
                             rel_l1(a, b).max()))
    for f in failures:
        _say("MISMATCH", f)
    assert not failures, "%d of %d variants differ" % (len(failures), len(cases))


def test_constraint_violation_raises_like_reference():
    """Negative pressure after a step: both paths raise at the same step."""
    from oracle import pyoracle
    n = (32, 32)
    q0 = smooth_state("euler2", n, seed=5)
    q0[3] = q0[3] - 0.999 * (q0[3] - 0.5 * (q0[1] ** 2 + q0[2] ** 2) / q0[0])  # p ~ 0
    q0[1] += 3.0 * np.sign(np.sin(np.arange(n[0]) * 0.7))[None, :]
    o, h = make_pair(pyoracle, "euler2", "weno2", "hll3", "rungekutta3", "periodic", n, 0.9)
    o.set_interior(q0)
    h.set_interior(q0)
    so = sh = None
    for k in range(50):
        try:
            o.step()
        except pyoracle.OracleError:
            so = k
            break
    for k in range(50):
        try:
            h.perform_step()
        except alsvinn_amd.ConstraintError:
            sh = k
            break
    assert so is not None and so == sh


@pytest.mark.parametrize("recon,flux,integ,bc", [("weno2", "hll3", "rungekutta3", "periodic"),
                                                 ("eno3", "hll", "rungekutta2", "neumann"),
                                                 ("none", "hll3", "forwardeuler", "periodic"),
                                                 ("eno2", "hll3", "rungekutta3", "neumann")])
def test_euler3_several_z_chunks_bitwise(recon, flux, integ, bc):
    """3D grids taller than one z-chunk of the marching kernel, with partial
    x/y tiles, match the oracle bit for bit."""
    from oracle import pyoracle
    n = (40, 10, 70)
    q0 = smooth_state("euler3", n, seed=23, jump=True)
    _say("case euler3", n, recon, flux, integ, bc)
    o, h = run_both(pyoracle, "euler3", recon, flux, integ, bc, n, 0.3, q0, steps=3)
    np.testing.assert_array_equal(h.get_state(), o.get_state())


def test_burgers3d_several_z_chunks_bitwise():
    from oracle import pyoracle
    n = (33, 17, 65)
    q0 = smooth_state("burgers", n, seed=29, jump=True)
    o, h = run_both(pyoracle, "burgers", "weno2", "godunov", "rungekutta3", "periodic", n, 0.3, q0, steps=3)
    np.testing.assert_array_equal(h.get_state(), o.get_state())


@pytest.mark.parametrize("recon", ["none", "weno2", "eno2", "eno3"])
@pytest.mark.parametrize("eq,n,bc", [("burgers", (37, 1, 1), "periodic"), ("euler2", (21, 18, 1), "neumann"),
                                     ("euler3", (9, 10, 11), "periodic")])
def test_unfused_reconstruction_bitwise(recon, eq, n, bc):
    """avh_reconstruct (Reconstruction::performReconstruction) equals the
    oracle's ReconstructionCPU / ENOCPU passes on the boundary-filled state,
    cell for cell, every active direction."""
    from oracle import pyoracle
    dim = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
    flux = "godunov" if eq == "burgers" else "hll3"
    q = smooth_state(eq, n[:dim], seed=17, jump=True)
    h = alsvinn_amd.Simulator(eq, recon, flux, "rungekutta3", bc, n)
    h.set_interior(q)
    o = pyoracle.Simulator(eq, recon, flux, "rungekutta3", bc, n)
    o.set_interior(q)
    U = o.get_state()
    cfg = pyoracle.make_config(eq, recon, flux, "rungekutta3", bc, n)
    for d in range(dim):
        left, right = h.reconstruct(d)
        lo, ro = pyoracle.reconstruct(cfg, U, d)
        np.testing.assert_array_equal(left, h.interior(lo))
        np.testing.assert_array_equal(right, h.interior(ro))


def _upper_extreme_state(equation, n, kind, seed=12):
    """Smooth states that cross an upper bound of the shortcuts along x, so
    that interfaces on both sides of it share wavefronts.  Densities are
    about 1e29 (inside the shortcuts' density range) so that wave speeds stay
    small enough for every flux product to remain finite.
      "momentum": the first velocity rises and falls (periodically) as
                  10^(57.5..59), the momentum through 2^290, the energy a
                  hundred times the kinetic energy;
      "pressure": velocities O(1), the pressure as 10^(179..182), so
                  gamma*p and the Roe pressure numerator pass 2^600 (HLL's
                  products overflow there; HLL3's do not)."""
    q = smooth_state(equation, n, seed=seed)
    nsd = q.shape[0] - 2
    nx = q.shape[-1]
    x = np.sin(np.pi * (np.arange(nx) + 0.5) / nx) ** 2
    u = [q[1 + c] / q[0] for c in range(nsd)]
    p = 0.4 * (q[-1] - 0.5 * sum(q[1 + c] ** 2 for c in range(nsd)) / q[0])
    q[0] = q[0] * 1e29
    if kind == "momentum":
        u[0] = (1.5 + np.tanh(u[0])) * 10.0 ** (57.5 + 1.5 * x)
    else:
        p = p * 10.0 ** (179.0 + 3.0 * x)
    for c in range(nsd):
        q[1 + c] = q[0] * u[c]
    kin = 0.5 * sum(q[1 + c] ** 2 for c in range(nsd)) / q[0]
    q[-1] = 100.0 * kin + p / 0.4 if kind == "momentum" else kin + p / 0.4
    return q


UPPER_CASES = [(eq, n, recon, flux) for eq, n in (("euler1", (70,)), ("euler2", (40, 20)))
               for recon in ("none", "weno2", "eno2", "eno3") for flux in ("hll", "hll3")] + \
              [("euler3", (20, 18, 24), recon, "hll3") for recon in ("weno2", "eno3")] + \
              [("euler3", (20, 18, 24), "weno2", "hll")]


@pytest.mark.parametrize("equation,n,recon,flux", UPPER_CASES)
def test_shared_reciprocal_upper_guards_bitwise(equation, n, recon, flux):
    """Operands above the shortcuts' upper bounds (momenta > 2^290; gamma*p
    and the Roe pressure numerator > 2^600) next to operands below them,
    with every reconstruction: the net flux and wave speeds keep every bit
    of the oracle's, and so does a whole RK3 step (state and dt).  In 3D
    every march-kernel tile shape runs, the shapes with shared reciprocals
    included."""
    from oracle import pyoracle
    shapes = range(7) if len(n) == 3 else [-1]  # the 3D march-kernel tile shapes
    for kind in ("momentum", "pressure") if flux == "hll3" else ("momentum",):
        _say("upper guards", equation, n, recon, flux, kind)
        q0 = _upper_extreme_state(equation, n, kind)
        for shape in shapes:
            o, h = make_pair(pyoracle, equation, recon, flux, "rungekutta3", "periodic", n, 0.4)
            o.set_interior(q0)
            h.set_interior(q0)
            if shape >= 0:
                h.set_tile_shape(shape)
            U = o.get_state()
            out_o, ws_o = pyoracle.compute_flux(o.cfg, U, np.zeros_like(U))
            out_h, ws_h = h.compute_flux()
            ref = np.ascontiguousarray(o.interior(out_o))
            assert np.isfinite(ref).all()
            assert np.array_equal(np.asarray(ws_h).view(np.uint64), np.asarray(ws_o).view(np.uint64)), shape
            assert np.array_equal(np.ascontiguousarray(out_h).view(np.uint64), ref.view(np.uint64)), shape
            dt_o = o.step()
            dt_h = h.perform_step()
            assert dt_h == dt_o
            assert np.array_equal(h.get_interior().view(np.uint64), o.get_interior().view(np.uint64)), shape


@pytest.mark.parametrize("equation,n", [("euler2", (40, 20)), ("euler3", (20, 18, 24))])
def test_negative_pressure_cells_flux_and_step(equation, n):
    """Cells of negative pressure (square roots of negative arguments, NaN
    wave speeds): the flux equals the oracle's wherever the oracle's is a
    number (NaN where it is NaN), and the step raises the constraint
    violation as the reference does."""
    from oracle import pyoracle
    _say("negative pressure", equation, n)
    q0 = smooth_state(equation, n, seed=13)
    flat = q0.reshape(q0.shape[0], -1)
    neg = np.random.default_rng(13).permutation(flat.shape[1])[:max(1, flat.shape[1] // 50)]
    kin = 0.5 * sum(flat[1 + c, neg] ** 2 for c in range(flat.shape[0] - 2)) / flat[0, neg]
    flat[-1, neg] = kin - 0.5  # p < 0
    o, h = make_pair(pyoracle, equation, "weno2", "hll3", "rungekutta3", "periodic", n, 0.4)
    o.set_interior(q0)
    h.set_interior(q0)
    U = o.get_state()
    out_o, ws_o = pyoracle.compute_flux(o.cfg, U, np.zeros_like(U))
    out_h, ws_h = h.compute_flux()
    ref = np.ascontiguousarray(o.interior(out_o))
    assert np.isnan(ref).any()
    np.testing.assert_array_equal(np.isnan(out_h), np.isnan(ref))
    fin = ~np.isnan(ref)
    assert np.array_equal(out_h[fin].view(np.uint64), ref[fin].view(np.uint64))
    np.testing.assert_array_equal(ws_h, ws_o)
    with pytest.raises(pyoracle.OracleError):
        o.step()
    with pytest.raises(alsvinn_amd.ConstraintError):
        h.perform_step()
