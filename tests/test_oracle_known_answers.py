"""Pins of the CPU oracle against the reference's own known-answer tests and
against closed forms.  Each test cites the reference test it re-expresses."""
import numpy as np
import pytest

from tests.cases import smooth_state


def ghosted(oracle, equation, recon, n, interior, bc="periodic"):
    s = oracle.Simulator(equation, recon, "godunov" if equation == "burgers" else "hll3", "forwardeuler", bc, n)
    s.set_interior(interior)
    return s, s.get_state()


# --------------------------------------------------------------------------
# reconstruction

def test_weno2_closed_form_weights(oracle):
    """WenoTest.ReconstructionSimple (test/library_tests/src/fvm/reconstruction/
    WenoTest.cpp:228-299): rho = {1, 2, 0, 1, 1, ...} in ghosted indices, the
    faces of cell 2 follow the WENO2 closed form with beta0 = 1, beta1 = 4."""
    nx = 10
    cfg = oracle.make_config("euler3", "weno2", "hll3", "forwardeuler", "periodic", (nx, 1, 1))
    U = np.ones((5, 1, 1, nx + 4))
    U[4] = 10
    U[0, 0, 0, 1] = 2
    U[0, 0, 0, 2] = 0
    U[0, 0, 0, 3] = 1
    left, right = oracle.reconstruct(cfg, U, 0)
    eps = 1e-8
    d0, d1 = 2.0 / 3.0, 1.0 / 3.0
    beta0, beta1 = 1.0, 4.0
    a0, a1 = d0 / (beta0 + eps) ** 2, d1 / (beta1 + eps) ** 2
    a0t, a1t = d1 / (beta0 + eps) ** 2, d0 / (beta1 + eps) ** 2
    right_expected = (a0 / (a0 + a1)) * 0.5 + (a1 / (a0 + a1)) * (-1.0)
    left_expected = (a0t / (a0t + a1t)) * (-0.5) + (a1t / (a0t + a1t)) * 1.0
    assert right[0, 0, 0, 2] == pytest.approx(right_expected, rel=1e-14, abs=1e-15)
    assert left[0, 0, 0, 2] == pytest.approx(left_expected, rel=1e-14, abs=1e-15)


@pytest.mark.parametrize("recon", ["none", "weno2", "eno2", "eno3"])
@pytest.mark.parametrize("n", [(12, 1, 1), (12, 10, 1), (8, 7, 6)])
def test_constant_states_reconstruct_exactly(oracle, recon, n):
    """WenoTest.cpp:32-224, CUDAWenoTest.cpp:115-231: a constant state
    reconstructs to itself on every face."""
    cfg = oracle.make_config("euler3", recon, "hll3", "forwardeuler", "periodic", n)
    g = oracle.ghosts_for(recon)
    shape, gax = oracle.ghosted_shape(n, g)
    U = np.empty((5,) + shape)
    for v, c in enumerate((1.0, 1.0, 1.0, 1.0, 10.0)):
        U[v] = c
    dim = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
    for d in range(dim):
        left, right = oracle.reconstruct(cfg, U, d)
        gx, gy, gz = gax
        sl = (slice(None), slice(gz, shape[0] - gz), slice(gy, shape[1] - gy), slice(gx, shape[2] - gx))
        np.testing.assert_allclose(left[sl], U[sl], rtol=1e-14, atol=0)
        np.testing.assert_allclose(right[sl], U[sl], rtol=1e-14, atol=0)


def _recon_rates(oracle, recon):
    """ReconstructionConvergenceTest (test/library_tests/src/fvm/reconstruction/
    ReconstructionConvergenceTest.cpp:164-283): cell averages of sin(2 pi x)+2
    as density (p = 2.5, u = 0), L1 and Linf errors of the face values
    against the point values, slopes of the log2 fits over n = 2^5 .. 2^14."""
    L1l, L1r, Lil, Lir, res = [], [], [], [], []
    for k in range(5, 15):
        n = 1 << k
        dx = 1.0 / n
        a = np.arange(n) * dx
        b = (np.arange(n) + 1) * dx
        avg = (-np.cos(2 * np.pi * b) + np.cos(2 * np.pi * a)) / (2 * np.pi * (b - a)) + 2
        prim = np.stack([avg, np.zeros(n), np.zeros(n), np.zeros(n), np.full(n, 2.5)])
        cons = oracle.from_primitive("euler3", prim)
        cfg = oracle.make_config("euler3", recon, "hll3", "forwardeuler", "periodic", (n, 1, 1))
        g = oracle.ghosts_for(recon)
        U = np.zeros((5, 1, 1, n + 2 * g))
        U[:, 0, 0, g:g + n] = cons
        U = oracle.apply_boundary(cfg, g, U)
        left, right = oracle.reconstruct(cfg, U, 0)
        f = lambda x: np.sin(2 * np.pi * x) + 2
        dl = np.abs(left[0, 0, 0, g:g + n] - f(a))
        dr = np.abs(right[0, 0, 0, g:g + n] - f(b))
        L1l.append(np.log2(dl.sum() / n))
        L1r.append(np.log2(dr.sum() / n))
        Lil.append(np.log2(dl.max()))
        Lir.append(np.log2(dr.max()))
        res.append(np.log2(n))
    slope = lambda y: -np.polyfit(res, y, 1)[0]
    return slope(L1l), slope(L1r), slope(Lil), slope(Lir)


@pytest.mark.parametrize("recon,l1,linf", [("weno2", 1.99, 1.99), ("eno2", 1.98, 1.99),
                                           ("eno3", 2.95, 2.99), ("none", 0.999, 0.999)])
def test_reconstruction_convergence_rates(oracle, recon, l1, linf):
    """Thresholds of ReconstructionConvergenceTest.cpp:306-320 (double)."""
    r1l, r1r, ril, rir = _recon_rates(oracle, recon)
    assert r1l >= l1 and r1r >= l1
    assert ril >= linf and rir >= linf


def test_eno_exact_on_polynomials(oracle):
    """ENO of order k reproduces face values of polynomials of degree < k
    exactly (to rounding) from their cell averages."""
    n = 32
    dx = 1.0 / n
    a = np.arange(n) * dx
    b = a + dx
    # linear: averages = value at the midpoint
    lin_avg = 3.0 + 2.0 * (a + b) / 2
    # quadratic q(x) = 1 + x + x^2: average = 1 + (a+b)/2 + (a^2+ab+b^2)/3
    quad_avg = 1 + (a + b) / 2 + (a * a + a * b + b * b) / 3
    for recon, avg, f in (("eno2", lin_avg, lambda x: 3.0 + 2.0 * x), ("eno3", quad_avg, lambda x: 1 + x + x * x)):
        g = oracle.ghosts_for(recon)
        cfg = oracle.make_config("burgers", recon, "godunov", "forwardeuler", "neumann", (n, 1, 1))
        U = np.zeros((1, 1, 1, n + 2 * g))
        U[0, 0, 0, g:g + n] = avg
        left, right = oracle.reconstruct(cfg, U, 0)
        # away from the (neumann-mirrored, non-polynomial) ends
        inner = slice(g + 3, g + n - 3)
        np.testing.assert_allclose(left[0, 0, 0, inner], f(a[3:n - 3]), rtol=0, atol=1e-13)
        np.testing.assert_allclose(right[0, 0, 0, inner], f(b[3:n - 3]), rtol=0, atol=1e-13)


def _eno2_ghosted(nx, ny):
    """euler3 volume of ones (E = 10), ENO2 ghost width 2, as EnoTest fills it."""
    g = 2
    shape = (1, ny + 2 * g if ny > 1 else 1, nx + 2 * g)
    U = np.ones((5,) + shape)
    U[4] = 10.0
    return U


@pytest.mark.parametrize("case", ["simple", "reverse", "reverse_y"])
def test_eno2_shift_known_answers(oracle, case):
    """EnoTest.ReconstructionSimple, ReconstructionSimpleReverse and
    ReconstructionSimpleReverseYDirection (test/library_tests/src/fvm/
    reconstruction/EnoTest.cpp:218-327): three rho values around ghosted
    index 2 fix the direction of the stencil shift (ENOCPU.cpp:139-150) and
    the coefficient rows used for each face, exactly."""
    if case == "reverse_y":
        nx, ny = 10, 10
        U = _eno2_ghosted(nx, ny)
        U[0, 0, 1, 3], U[0, 0, 2, 3], U[0, 0, 3, 3] = 1.0, 0.0, 2.0
        cfg = oracle.make_config("euler3", "eno2", "hll3", "forwardeuler", "periodic", (nx, ny, 1))
        left, right = oracle.reconstruct(cfg, U, 1)
        assert right[0, 0, 2, 3] == -1.0 / 2.0
        assert left[0, 0, 2, 3] == 1.0 / 2.0
        return
    nx = 10
    U = _eno2_ghosted(nx, 1)
    vals = (2.0, 0.0, 1.0) if case == "simple" else (1.0, 0.0, 2.0)
    U[0, 0, 0, 1:4] = vals
    cfg = oracle.make_config("euler3", "eno2", "hll3", "forwardeuler", "periodic", (nx, 1, 1))
    left, right = oracle.reconstruct(cfg, U, 0)
    if case == "simple":
        assert right[0, 0, 0, 2] == 1.0 / 2.0
        assert left[0, 0, 0, 2] == -1.0 / 2.0
    else:
        assert right[0, 0, 0, 2] == -1.0 / 2.0
        assert left[0, 0, 0, 2] == 1.0 / 2.0


@pytest.mark.parametrize("vals,shift,left_val,right_val", [
    # level 0: |1| < |2| -> shift 1; level 1: |1| < |-1| false -> shift 1
    ((0.0, 0.0, 1.0, 3.0, 4.0), 1, 1.0 / 3.0, 11.0 / 6.0),
    # level 0: |0.5| < |1.5| -> 1; level 1: |1| < |2.5| -> 2
    ((0.0, 1.0, 1.5, 3.0, 7.0), 2, 4.0 / 3.0, 19.0 / 12.0),
    # level 0: |2| < |1| false -> 0; level 1: |-1| < |-0.5| false -> 0
    ((5.0, 1.0, 3.0, 4.0, 4.5), 0, 7.0 / 3.0, 43.0 / 12.0),
])
def test_eno3_shift_hand_computed(oracle, vals, shift, left_val, right_val):
    """ENO3 by the reference's rule (ENOCPU.cpp:139-150): at every level the
    undivided differences at (cell-1, cell) are compared and the shift grows
    by one when the left one is smaller in magnitude; the faces use the
    coefficient rows shift and shift+1 (ENOCoefficients.cpp:35-40) on cells
    i-shift .. i-shift+2.  Expected values worked out by hand for the five
    indicator values u[i-2..i+2]."""
    nx = 10
    g = 3
    U = np.ones((5, 1, 1, nx + 2 * g))
    U[4] = 10.0
    i = 6
    U[0, 0, 0, i - 2:i + 3] = vals
    cfg = oracle.make_config("euler3", "eno3", "hll3", "forwardeuler", "periodic", (nx, 1, 1))
    left, right = oracle.reconstruct(cfg, U, 0)
    assert left[0, 0, 0, i] == pytest.approx(left_val, rel=1e-14, abs=1e-14)
    assert right[0, 0, 0, i] == pytest.approx(right_val, rel=1e-14, abs=1e-14)
    # momentum variables are constant 1: every stencil reproduces them
    assert left[1, 0, 0, i] == pytest.approx(1.0, rel=1e-14)


# --------------------------------------------------------------------------
# equations and fluxes

def test_euler_point_flux(oracle):
    """EulerEquationTest.FluxTest (test/library_tests/src/fvm/equation/
    EulerEquationTest.cpp:22-59): all ones -> F = (1, m + p e_d, 2)."""
    a = np.array([1, 1, 1, 1, 1, 1, 1, 1, 1], dtype=float)  # rho, m(3), E, p, u(3)
    for d in range(3):
        F = oracle.point_flux_all(3, d, a)
        m = [1.0, 1.0, 1.0]
        m[d] = 2.0
        np.testing.assert_array_equal(F, [1.0] + m + [2.0])


def test_burgers_point_flux(oracle):
    """BurgerEquationTest.FluxTest: u = 42.42 -> F = u*u/2."""
    u = 42.42
    F = oracle.point_flux("burgers", 0, np.array([u]))
    assert F[0] == u * u / 2


def test_hll_consistency_exact(oracle):
    """HLLTest.ConsistencyTest (test/library_tests/src/fvm/flux/HLLTest.cpp:24-71):
    F(u, u) equals the point flux exactly in every direction."""
    a = np.array([1, 2, 3, 4, 5, 6, 7, 8, 9], dtype=float)
    for d in range(3):
        F, _ = oracle.flux_all(3, "hll", d, a, a)
        np.testing.assert_array_equal(F, oracle.point_flux_all(3, d, a))


@pytest.mark.parametrize("flux", ["hll", "hll3"])
def test_flux_consistency_subsonic(oracle, flux):
    """Consistency F(u, u) = f(u) on physical subsonic states, where both
    fluxes take their middle-state branch."""
    rng = np.random.default_rng(0)
    for nsd in (1, 2, 3):
        for _ in range(50):
            rho = 0.5 + rng.random()
            u = 0.3 * (rng.random(nsd) - 0.5)
            p = 0.5 + rng.random()
            q = np.concatenate([[rho], rho * u, [p / 0.4 + 0.5 * rho * np.dot(u, u)]])
            for d in range(nsd):
                eq = {1: "euler1", 2: "euler2", 3: "euler3"}[nsd]
                F, ws = oracle.interface_flux(eq, flux, d, q, q)
                f = oracle.point_flux(eq, d, q)
                np.testing.assert_allclose(F, f, rtol=1e-13, atol=1e-13)
                c = np.sqrt(1.4 * p / rho)
                assert ws == pytest.approx(abs(u[d]) + c, rel=1e-13)


def test_hll3_resolves_stationary_contact_exactly(oracle):
    """HLL3 is a three-wave solver: a stationary contact (u = 0, equal p,
    density jump) has flux (0, p, 0, 0) -- the star state equals the
    upwind state.  HLL smears it."""
    for rl, rr in ((1.0, 0.125), (2.0, 1.0), (0.3, 5.0)):
        p = 1.7
        ql = np.array([rl, 0.0, 0.0, p / 0.4])
        qr = np.array([rr, 0.0, 0.0, p / 0.4])
        F, _ = oracle.interface_flux("euler2", "hll3", 0, ql, qr)
        np.testing.assert_allclose(F, [0.0, p, 0.0, 0.0], rtol=1e-15, atol=1e-15)
        Fh, _ = oracle.interface_flux("euler2", "hll", 0, ql, qr)
        assert abs(Fh[0]) > 1e-3


def test_hll3_moving_contact_upwind_flux(oracle):
    """A contact moving with speed u > 0 (equal u and p across): the HLL3
    flux is the left state's point flux."""
    u, p = 0.4, 1.2
    for rl, rr in ((1.0, 0.125), (0.5, 3.0)):
        ql = np.array([rl, rl * u, p / 0.4 + 0.5 * rl * u * u])
        qr = np.array([rr, rr * u, p / 0.4 + 0.5 * rr * u * u])
        F, _ = oracle.interface_flux("euler1", "hll3", 0, ql, qr)
        np.testing.assert_allclose(F, oracle.point_flux("euler1", 0, ql), rtol=1e-13, atol=1e-14)


def test_godunov_flux_cases(oracle):
    """Godunov for Burgers: max(f(max(uL,0)), f(min(uR,0)))."""
    for ul, ur, f in ((1.0, 2.0, 0.5), (-2.0, -1.0, 0.5), (1.0, -1.0, 0.5), (-1.0, 1.0, 0.0), (3.0, -2.0, 4.5)):
        F, ws = oracle.interface_flux("burgers", "godunov", 0, np.array([ul]), np.array([ur]))
        assert F[0] == f
        assert ws == max(abs(ul), abs(ur))


@pytest.mark.parametrize("flux", ["hll", "hll3"])
def test_net_flux_of_constant_state_is_zero(oracle, flux):
    """CUDANumericalFluxTest.ConsistencyTest (test/library_tests/src/fvm/flux/
    CUDANumericalFluxTest.cpp:94-198): constant (1,1,1,1,10), neumann, the
    net flux on the interior is 0 within 1e-30; untouched ghosts keep 44."""
    n = (10, 10, 10)
    cfg = oracle.make_config("euler3", "none", flux, "forwardeuler", "neumann", n)
    shape, _ = oracle.ghosted_shape(n, 1)
    U = np.empty((5,) + shape)
    for v, c in enumerate((1.0, 1.0, 1.0, 1.0, 10.0)):
        U[v] = c
    out, ws = oracle.compute_flux(cfg, U, np.full_like(U, 44.0))
    assert np.all(np.abs(out[:, 1:-1, 1:-1, 1:-1]) <= 1e-30)
    assert out[0, 0, 0, 0] == 44.0
    assert np.all(ws > 0)


def test_burgers_godunov_constant_net_flux_zero(oracle):
    """BurgersFluxTest (test/library_tests/src/fvm/flux/BurgersFluxTest.cpp:101-139)."""
    n = (32, 1, 1)
    cfg = oracle.make_config("burgers", "weno2", "godunov", "forwardeuler", "periodic", n)
    U = np.full((1, 1, 1, 36), 0.7)
    out, ws = oracle.compute_flux(cfg, U, np.zeros_like(U))
    assert np.all(out[0, 0, 0, 2:-2] == 0.0)
    assert ws[0] == pytest.approx(0.7, rel=1e-15)  # WENO2 faces of a constant are exact to rounding


# --------------------------------------------------------------------------
# boundary

def test_neumann_two_ghost_index_map(oracle):
    """BoundaryTest.NeumannTest2CellsVarying (test/library_tests/src/fvm/
    boundary/BoundaryTest.cpp:79-140): values = flat index, ghosts mirror."""
    nx = ny = nz = 8
    g = 2
    cfg = oracle.make_config("euler3", "weno2", "hll3", "forwardeuler", "neumann", (nx, ny, nz))
    T = (nz + 2 * g, ny + 2 * g, nx + 2 * g)
    idx = np.arange(np.prod(T), dtype=float).reshape(T)
    U = np.zeros((5,) + T)
    U[:, g:g + nz, g:g + ny, g:g + nx] = idx[g:g + nz, g:g + ny, g:g + nx]
    V = oracle.apply_boundary(cfg, g, U)[0]
    for y in range(g, ny + g):
        for z in range(g, nz + g):
            assert V[z, y, 0] == idx[z, y, 3]
            assert V[z, y, 1] == idx[z, y, 2]
            assert V[z, y, g + nx + 1] == idx[z, y, nx + g - 2]
            assert V[z, y, g + nx] == idx[z, y, nx + g - 1]


@pytest.mark.parametrize("bc,mode", [("periodic", "wrap"), ("neumann", "symmetric")])
@pytest.mark.parametrize("n,g", [((9, 1, 1), 2), ((7, 6, 1), 3), ((5, 6, 7), 2), ((6, 5, 4), 1)])
def test_boundary_equals_pad(oracle, bc, mode, n, g):
    """The x -> y -> z ghost fill (BoundaryCPU.cpp:30-121), corners included,
    equals numpy's wrap / symmetric padding of the interior."""
    cfg = oracle.make_config("euler3", "weno2", "hll3", "forwardeuler", bc, n)
    shape, gax = oracle.ghosted_shape(n, g)
    rng = np.random.default_rng(1)
    inner = rng.random((5, n[2], n[1], n[0]))
    U = np.zeros((5,) + shape)
    gx, gy, gz = gax
    U[:, gz:shape[0] - gz, gy:shape[1] - gy, gx:shape[2] - gx] = inner
    V = oracle.apply_boundary(cfg, g, U)
    expected = np.pad(inner, ((0, 0), (gz, gz), (gy, gy), (gx, gx)), mode=mode)
    np.testing.assert_array_equal(V, expected)


# --------------------------------------------------------------------------
# system, integrators, CFL

@pytest.mark.parametrize("equation,n", [("burgers", (40, 1, 1)), ("burgers", (12, 10, 1)), ("euler1", (40, 1, 1)),
                                        ("euler2", (14, 12, 1)), ("euler3", (9, 8, 10))])
@pytest.mark.parametrize("recon", ["none", "weno2", "eno2", "eno3"])
def test_interior_and_strips_equal_whole_domain(oracle, equation, n, recon):
    """ConservedSystem computes the interior, then 2*dim side strips
    (alsfvm/src/simulator/ConservedSystem.cpp:45-95); the result and the wave
    speeds equal one whole-domain computeFlux bit for bit -- the identity the
    fused single-pass kernel relies on."""
    flux = "godunov" if equation == "burgers" else "hll3"
    cfg = oracle.make_config(equation, recon, flux, "rungekutta3", "periodic", n)
    s = oracle.Simulator(equation, recon, flux, "rungekutta3", "periodic", n)
    s.set_interior(smooth_state(equation, n, seed=4, jump=True))
    U = s.get_state()
    out0 = np.zeros_like(U)
    a, wa = oracle.system(cfg, U, out0, whole_domain=False)
    b, wb = oracle.system(cfg, U, out0, whole_domain=True)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(wa, wb)


def test_rk_amplification_polynomials(oracle):
    """One step of du/dt = u with step h through the integrators' own vector
    arithmetic gives the stability polynomials 1+h, 1+h+h^2/2,
    1+h+h^2/2+h^3/6 (to rounding)."""
    for h in (0.1, 0.37, 1e-3):
        assert oracle.integrate_ode("forwardeuler", h, 1) == 1 + h
        assert oracle.integrate_ode("rungekutta2", h, 1) == pytest.approx(1 + h + h * h / 2, rel=4e-16)
        assert oracle.integrate_ode("rungekutta3", h, 1) == pytest.approx(1 + h + h * h / 2 + h ** 3 / 6, rel=4e-16)


@pytest.mark.parametrize("integ,rate", [("forwardeuler", 0.9), ("rungekutta2", 1.9), ("rungekutta3", 2.9)])
def test_integrator_convergence_rates(oracle, integ, rate):
    """IntegratorConvergenceTest (test/library_tests/src/fvm/integrator/
    IntegratorConvergenceTest.cpp:77-163): du/dt = u to t = 1, N = 2<<k."""
    errs, Ns = [], []
    for k in range(5, 9):
        N = 2 << k
        errs.append(np.log(abs(np.e - oracle.integrate_ode(integ, 1.0 / N, N))))
        Ns.append(np.log(N))
    assert -np.polyfit(Ns, errs, 1)[0] >= rate


def test_cfl_dt_on_constant_state(oracle):
    """dt = cfl / sum_d(ws_d/dx_d) (Integrator.cpp:20-46) with ws = |u_d| + c
    on a constant state; lengths (b-a)/N (Grid.cpp:24-41)."""
    nx, ny = 20, 10
    rho, ux, uy, p = 1.3, 0.4, -0.7, 2.1
    c = np.sqrt(1.4 * p / rho)
    q = np.stack([np.full((1, ny, nx), v) for v in (rho, rho * ux, rho * uy, p / 0.4 + 0.5 * rho * (ux * ux + uy * uy))])
    s = oracle.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", (nx, ny, 1),
                         upper=(2.0, 1.0, 1.0), cfl=0.4)
    s.set_interior(q)
    dt = s.step()
    dx, dy = 2.0 / nx, 1.0 / ny
    assert dt == pytest.approx(0.4 / ((abs(ux) + c) / dx + (abs(uy) + c) / dy), rel=1e-14)
    np.testing.assert_allclose(s.get_interior(), q, rtol=1e-14, atol=1e-14)


def test_fixed_interval_writer_lands_on_save_times(oracle):
    """FixedIntervalWriter::adjustTimestep clips dt to the next save time
    (alsfvm/src/io/FixedIntervalWriter.cpp:45-53): with numberOfSaves = 4 up
    to T = 0.2 the run hits 0.05, 0.1, 0.15 and 0.2 exactly."""
    N = 128
    s = oracle.Simulator("burgers", "weno2", "godunov", "rungekutta3", "periodic", (N, 1, 1), cfl=0.45)
    x = (np.arange(N) + 0.5) / N
    s.set_interior(np.sin(2 * np.pi * x).reshape(1, 1, 1, N))
    s.add_fixed_interval(0.2 / 4)
    times = []
    while s.time < 0.2:
        s.step()
        times.append(s.time)
    for k in (1, 2, 3):
        assert any(t == k * 0.05 for t in times), k
    assert times[-1] == 0.2
    assert s.saves == 4


def test_conservation_under_periodic_bc(oracle):
    """Periodic finite-volume updates conserve every conserved variable
    (the net fluxes telescope) to rounding."""
    n = (48, 40, 1)
    s = oracle.Simulator("euler2", "weno2", "hll3", "rungekutta3", "periodic", n, cfl=0.4)
    q0 = smooth_state("euler2", n, seed=2, jump=True)
    s.set_interior(q0)
    for _ in range(20):
        s.step()
    q = s.get_interior()
    tot0 = q0.reshape(4, -1).sum(axis=1)
    tot = q.reshape(4, -1).sum(axis=1)
    np.testing.assert_allclose(tot, tot0, rtol=0, atol=1e-10 * np.abs(q0).sum() / q0[0].size)


def test_constraint_violation_raises(oracle):
    """Simulator::checkConstraints throws when rho or p is not positive
    (alsfvm/src/simulator/Simulator.cpp:180-189)."""
    n = (16, 1, 1)
    s = oracle.Simulator("euler1", "none", "hll3", "forwardeuler", "periodic", n, cfl=0.5)
    q = np.stack([np.ones((1, 1, 16)), np.zeros((1, 1, 16)), np.full((1, 1, 16), 2.5)])
    q[0, 0, 0, 5] = -1.0
    s.set_interior(q)
    with pytest.raises(oracle.OracleError):
        s.step()


# --------------------------------------------------------------------------
# random numbers

def test_mt19937_64_standard_value(oracle):
    """C++ [rand.predef]: the 10000th output of a default-constructed
    std::mt19937_64 is 9981545732273789042 (the engine STLMersenne uses,
    alsuq/src/generator/STLMersenne.cpp:25-85)."""
    v = oracle.mt19937_64(5489, 9999, 1)
    assert int(v[0]) == 9981545732273789042
