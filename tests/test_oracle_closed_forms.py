"""Pins of the oracle's full time stepping against exact solutions:
Burgers characteristics before the shock, Sod's Riemann problem (Toro's
exact solver), and exact preservation of a stationary contact by HLL3."""
import numpy as np
import pytest


def burgers_exact(x, t):
    """u(x,t) = sin(2 pi (x - u t)) for t < 1/(2 pi), Newton per point."""
    u = np.sin(2 * np.pi * x)
    for _ in range(100):
        g = u - np.sin(2 * np.pi * (x - u * t))
        dg = 1 + 2 * np.pi * t * np.cos(2 * np.pi * (x - u * t))
        u = u - g / dg
    return u


def test_burgers_sine_before_shock_second_order(oracle):
    errs, ns = [], []
    for n in (128, 256, 512, 1024):
        s = oracle.Simulator("burgers", "weno2", "godunov", "rungekutta3", "periodic", (n, 1, 1), cfl=0.45)
        x = (np.arange(n) + 0.5) / n
        a = np.arange(n) / n
        b = a + 1.0 / n
        avg0 = (np.cos(2 * np.pi * a) - np.cos(2 * np.pi * b)) / (2 * np.pi / n)  # exact cell averages
        s.set_interior(avg0.reshape(1, 1, 1, n))
        s.add_fixed_interval(0.1)
        s.run(0.1)
        assert s.time == 0.1
        u = s.get_interior().ravel()
        errs.append(np.abs(u - burgers_exact(x, 0.1)).mean())
        ns.append(n)
    rate = -np.polyfit(np.log(ns), np.log(errs), 1)[0]
    assert errs[-1] < 2e-5
    assert rate > 1.6


def sod_exact(x, t, x0=0.5, gamma=1.4, L=(1.0, 0.0, 1.0), R=(0.125, 0.0, 0.1)):
    """Exact Riemann solution (Toro, Riemann Solvers and Numerical Methods
    for Fluid Dynamics, ch. 4): Newton for p*, then sampling."""
    rl, ul, pl = L
    rr, ur, pr = R
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)

    def f(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((gamma + 1) * rk), (gamma - 1) / (gamma + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B)), np.sqrt(A / (B + p)) * (1 - (p - pk) / (2 * (B + p)))
        pr_ = p / pk
        return 2 * ck / (gamma - 1) * (pr_ ** ((gamma - 1) / (2 * gamma)) - 1), 1 / (rk * ck) * pr_ ** (-(gamma + 1) / (2 * gamma))

    p = 0.5 * (pl + pr)
    for _ in range(100):
        fl, dl = f(p, rl, pl, cl)
        fr, dr = f(p, rr, pr, cr)
        p = max(1e-12, p - (fl + fr + ur - ul) / (dl + dr))
    fl, _ = f(p, rl, pl, cl)
    fr, _ = f(p, rr, pr, cr)
    um = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    g = gamma
    rho = np.empty_like(x)
    for i, xi in enumerate(x):
        S = (xi - x0) / t
        if S <= um:  # left of contact
            if p > pl:
                SL = ul - cl * np.sqrt((g + 1) / (2 * g) * p / pl + (g - 1) / (2 * g))
                rho[i] = rl if S <= SL else rl * ((p / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pl + 1))
            else:
                SHL = ul - cl
                cml = cl * (p / pl) ** ((g - 1) / (2 * g))
                STL = um - cml
                if S <= SHL:
                    rho[i] = rl
                elif S > STL:
                    rho[i] = rl * (p / pl) ** (1 / g)
                else:
                    rho[i] = rl * (2 / (g + 1) + (g - 1) / ((g + 1) * cl) * (ul - S)) ** (2 / (g - 1))
        else:
            if p > pr:
                SR = ur + cr * np.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g))
                rho[i] = rr if S >= SR else rr * ((p / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pr + 1))
            else:
                SHR = ur + cr
                cmr = cr * (p / pr) ** ((g - 1) / (2 * g))
                STR = um + cmr
                if S >= SHR:
                    rho[i] = rr
                elif S <= STR:
                    rho[i] = rr * (p / pr) ** (1 / g)
                else:
                    rho[i] = rr * (2 / (g + 1) - (g - 1) / ((g + 1) * cr) * (ur - S)) ** (2 / (g - 1))
    return rho, p, um


def test_sod_shock_tube_converges_to_exact(oracle):
    """Euler1 HLL3 + WENO2 + RK3, neumann, Sod data to t = 0.2: the L1
    density error shrinks with the grid like a shock-capturing scheme should."""
    rho_star_p, p_star, u_star = sod_exact(np.array([0.5]), 0.2)[0], *sod_exact(np.array([0.5]), 0.2)[1:]
    assert p_star == pytest.approx(0.30313, rel=1e-4)  # Toro's tabulated p* for test 1
    assert u_star == pytest.approx(0.92745, rel=1e-4)
    errs, ns = [], []
    for n in (100, 200, 400, 800):
        x = (np.arange(n) + 0.5) / n
        rho = np.where(x < 0.5, 1.0, 0.125)
        p = np.where(x < 0.5, 1.0, 0.1)
        q = np.stack([rho, np.zeros(n), p / 0.4]).reshape(3, 1, 1, n)
        s = oracle.Simulator("euler1", "weno2", "hll3", "rungekutta3", "neumann", (n, 1, 1), cfl=0.45)
        s.set_interior(q)
        s.add_fixed_interval(0.2)
        s.run(0.2)
        assert s.time == 0.2
        r = s.get_interior()[0].ravel()
        errs.append(np.abs(r - sod_exact(x, 0.2)[0]).mean())
        ns.append(n)
    rate = -np.polyfit(np.log(ns), np.log(errs), 1)[0]
    assert errs[-1] < 4e-3
    assert rate > 0.6


def test_hll3_keeps_stationary_contact(oracle):
    """A stationary contact (u = 0, p = 1, density jump) is an exact steady
    solution that HLL3 keeps to rounding; HLL diffuses it."""
    n = 64
    x = (np.arange(n) + 0.5) / n
    rho = np.where((x > 0.3) & (x < 0.7), 2.0, 0.5)
    q = np.stack([rho, np.zeros(n), np.full(n, 1.0 / 0.4)]).reshape(3, 1, 1, n)
    for flux, keeps in (("hll3", True), ("hll", False)):
        s = oracle.Simulator("euler1", "weno2", flux, "rungekutta3", "periodic", (n, 1, 1), cfl=0.4)
        s.set_interior(q)
        for _ in range(20):
            s.step()
        err = np.abs(s.get_interior() - q).max()
        assert (err < 1e-13) == keeps, (flux, err)


# --------------------------------------------------------------------------
# HLL3 star states (alsfvm/include/alsfvm/numflux/euler/HLL3.hpp:52-148)

GAMMA = 1.4
EQ_OF = {1: "euler1", 2: "euler2", 3: "euler3"}


def conserved(rho, u, p):
    u = np.asarray(u, dtype=float)
    return np.concatenate([[rho], rho * u, [p / (GAMMA - 1) + 0.5 * rho * np.dot(u, u)]])


def primitive(q, nsd):
    rho, m, E = q[0], q[1:1 + nsd], q[1 + nsd]
    return rho, m / rho, (GAMMA - 1) * (E - 0.5 * np.dot(m, m) / rho), E


def euler_flux(q, nsd, d):
    rho, u, p, E = primitive(q, nsd)
    F = np.empty(nsd + 2)
    F[0] = rho * u[d]
    F[1:1 + nsd] = rho * u * u[d]
    F[1 + d] += p
    F[1 + nsd] = (E + p) * u[d]
    return F


def random_pair(rng, nsd, mach=0.6):
    """Two states with densities, pressures and velocities of order one."""
    out = []
    for _ in range(2):
        rho = 0.3 + 2.0 * rng.random()
        p = 0.3 + 2.0 * rng.random()
        u = mach * (2 * rng.random(nsd) - 1)
        out.append(conserved(rho, u, p))
    return out


def hllc_toro(qL, qR, nsd, d):
    """The HLLC flux in the form of Toro, Riemann Solvers and Numerical Methods
    for Fluid Dynamics (3rd ed.), eqs. (10.37)-(10.39): contact speed S*, star
    states U*_K = rho_K (S_K - u_K)/(S_K - S*) (1, S*, v, E_K/rho_K +
    (S* - u_K)(S* + p_K/(rho_K (S_K - u_K)))) and F*_K = F_K + S_K (U*_K - U_K).
    The transverse velocity v of both star states is that of the HLL
    intermediate state (Harten, Lax and van Leer 1983), U_hll = (S_R U_R -
    S_L U_L - (F_R - F_L)) / (S_R - S_L).  The wave speeds are HLL3's: Roe
    (sqrt(rho)) averages of u and p, the mean density, the per-side sound
    speeds and the compression correction 0.5 max(0, u_L - u_R)."""
    rL, uL, pL, EL = primitive(qL, nsd)
    rR, uR, pR, ER = primitive(qR, nsd)
    wl, wr = np.sqrt(rL), np.sqrt(rR)
    ub = (wl * uL[d] + wr * uR[d]) / (wl + wr)
    pb = (wl * pL + wr * pR) / (wl + wr)
    rb = 0.5 * (rL + rR)
    cL, cR, cb = np.sqrt(GAMMA * pL / rL), np.sqrt(GAMMA * pR / rR), np.sqrt(GAMMA * pb / rb)
    corr = 0.5 * max(0.0, uL[d] - uR[d])
    SL = min(uL[d] + corr - cL, ub - cb)
    SR = max(uR[d] - corr + cR, ub + cb)
    FL, FR = euler_flux(qL, nsd, d), euler_flux(qR, nsd, d)
    if SL > 0:
        return FL, max(abs(SL), abs(SR)), "left"
    if SR < 0:
        return FR, max(abs(SL), abs(SR)), "right"
    Ss = (pR - pL + rL * uL[d] * (SL - uL[d]) - rR * uR[d] * (SR - uR[d])) / (rL * (SL - uL[d]) - rR * (SR - uR[d]))
    Uhll = (SR * qR - SL * qL - (FR - FL)) / (SR - SL)
    v = Uhll[1:1 + nsd] / Uhll[0]
    v[d] = Ss
    if Ss >= 0:
        rK, uK, pK, EK, SK, qK, FK, side = rL, uL, pL, EL, SL, qL, FL, "star-left"
    else:
        rK, uK, pK, EK, SK, qK, FK, side = rR, uR, pR, ER, SR, qR, FR, "star-right"
    fac = rK * (SK - uK[d]) / (SK - Ss)
    Us = np.empty(nsd + 2)
    Us[0] = fac
    Us[1:1 + nsd] = fac * v
    Us[1 + nsd] = fac * (EK / rK + (Ss - uK[d]) * (Ss + pK / (rK * (SK - uK[d]))))
    return FK + SK * (Us - qK), max(abs(SL), abs(SR)), side


@pytest.mark.parametrize("nsd", [1, 2, 3])
def test_hll3_equals_toro_hllc(oracle, nsd):
    """HLL3 is the HLLC solver: on random subsonic and supersonic pairs, every
    branch (star-left, star-right, left, right) agrees with Toro's form of the
    star states to rounding, which fixes the star pressure p* = p_K +
    rho_K (u_K - S*)(u_K - S_K) inside E* (HLL3.hpp:88-103) and S*."""
    rng = np.random.default_rng(11 + nsd)
    seen = set()
    for k in range(400):
        mach = 0.6 if k % 4 else 2.5
        qL, qR = random_pair(rng, nsd, mach)
        for d in range(nsd):
            F, ws = oracle.interface_flux(EQ_OF[nsd], "hll3", d, qL, qR)
            Fe, wse, side = hllc_toro(qL, qR, nsd, d)
            seen.add(side)
            scale = max(1.0, np.abs(Fe).max())
            np.testing.assert_allclose(F, Fe, rtol=0, atol=1e-12 * scale)
            assert ws == pytest.approx(wse, rel=1e-14)
    assert seen == {"left", "right", "star-left", "star-right"}


@pytest.mark.parametrize("flux", ["hll", "hll3"])
@pytest.mark.parametrize("nsd", [1, 2, 3])
def test_flux_reflection_symmetry(oracle, flux, nsd):
    """Mirror symmetry of a Riemann solver: reflecting x_d -> -x_d swaps the
    sides and negates the normal velocity, so F(L, R) = -M F(M R, M L) with M
    negating the normal momentum, and the wave speed is unchanged.  For HLL3
    an interface with S* > 0 takes the left star formula and its mirror the
    right one, so the two star branches check each other."""
    rng = np.random.default_rng(5 + nsd)
    branches = {True: 0, False: 0}
    for k in range(300):
        qL, qR = random_pair(rng, nsd, 0.6 if k % 3 else 2.0)
        for d in range(nsd):
            M = np.ones(nsd + 2)
            M[1 + d] = -1.0
            F, ws = oracle.interface_flux(EQ_OF[nsd], flux, d, qL, qR)
            Fm, wsm = oracle.interface_flux(EQ_OF[nsd], flux, d, M * qR, M * qL)
            scale = max(1.0, np.abs(F).max())
            np.testing.assert_allclose(F, -M * Fm, rtol=0, atol=1e-12 * scale)
            assert ws == pytest.approx(wsm, rel=1e-14)
            if flux == "hll3":
                _, _, side = hllc_toro(qL, qR, nsd, d)
                if side.startswith("star"):
                    branches[side == "star-left"] += 1
    if flux == "hll3":
        assert branches[True] > 50 and branches[False] > 50
