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
