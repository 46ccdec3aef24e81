"""Initial data of the benchmark problems, evaluated on the host at cell
midpoints x_i = a + dx*i + dx/2 (alsfvm/src/grid/Grid.cpp:148-171) and
converted to conserved variables with E = p/(gamma-1) + 0.5*rho*|u|^2
(alsfvm/include/alsfvm/equation/euler/Euler.hpp:236-245).  The reference
evaluates user Python snippets per midpoint
(alsfvm/src/init/PythonInitialData.cpp:84-281); here the same formulas are
numpy expressions, and both the hip path and its checks receive identical
arrays.
"""
import numpy as np


def midpoints(n, lower=0.0, upper=1.0):
    dx = (upper - lower) / n
    return lower + dx * np.arange(n) + dx / 2.0


def sine_burgers(n):
    """u0(x) = sin(2 pi x) on [0,1] (examples/sine)."""
    x = midpoints(n)
    return np.sin(2 * np.pi * x).reshape(1, 1, 1, n)


def conserved_from_primitive(rho, u, p, gamma=1.4):
    """Euler::computeConserved, same association: m = rho*u component-wise,
    E = p/(gamma-1) + (0.5*rho)*(u.u) with u.u summed left to right."""
    m = [uc * rho for uc in u]
    dot = u[0] * u[0]
    for uc in u[1:]:
        dot = dot + uc * uc
    E = p / (gamma - 1) + 0.5 * rho * dot
    return np.stack([rho] + m + [E])


def kh_perturbation(x, a, b):
    """delta(x) = 0.01 * sum_{i=1..10} a_i cos(2 pi (i x + b_i)) with a
    normalised to sum 1 (a sum below 1e-10 counts as 1, as in
    examples/kelvinhelmholtz/kelvinhelmholtz.py:8-14)."""
    a = np.asarray(a, dtype=np.float64)
    norm = a.sum()
    if abs(norm) < 1e-10:
        norm = 1.0
    out = np.zeros_like(x)
    for i in range(len(a)):
        out = out + a[i] * np.cos(2 * np.pi * ((i + 1) * x + b[i]))
    return 0.01 * out / norm


def kh_interfaces(n, params):
    """Band interfaces at the x midpoints: (lower, upper) = (0.25 + d0(x),
    0.75 + d1(x)).  The 40 parameters follow the reference example's layout
    (examples/kelvinhelmholtz/kelvinhelmholtz.py:2-6): a[0:10] amplitudes and
    a[20:30] phases of the upper interface, a[10:20] and a[30:40] of the
    lower one."""
    params = np.asarray(params, dtype=np.float64)
    x = midpoints(n)
    lo = 0.25 + kh_perturbation(x, params[10:20], params[30:40])
    hi = 0.75 + kh_perturbation(x, params[0:10], params[20:30])
    return lo, hi


def kelvin_helmholtz_2d(n, params, gamma=1.4):
    """2D KH on the unit square: rho=2, ux=-0.5 inside the band
    0.25+d0(x) < y < 0.75+d1(x), else rho=1, ux=0.5; uy=0, p=2.5."""
    lo, hi = kh_interfaces(n, params)
    y = midpoints(n)
    Y = y[:, None]
    band = (Y > lo[None, :]) & (Y < hi[None, :])
    rho = np.where(band, 2.0, 1.0)
    ux = np.where(band, -0.5, 0.5)
    uy = np.zeros_like(rho)
    p = np.full_like(rho, 2.5)
    q = conserved_from_primitive(rho, [ux, uy], p, gamma)
    return q.reshape(4, 1, n, n)


def kelvin_helmholtz_3d(n, params, gamma=1.4):
    """3D KH, the 2D data extruded in z with the band interfaces perturbed
    over (x, z): params = 2 x (10 amplitudes + 10 phases) per interface for
    each of x and z (80 numbers: a0x, a1x, b0x, b1x, a0z, a1z, b0z, b1z)."""
    params = np.asarray(params, dtype=np.float64)
    x = midpoints(n)
    p4 = params.reshape(8, 10)
    # the interfaces depend on (z, x) only: evaluate on the plane, broadcast over y
    Zp, Xp = np.meshgrid(x, x, indexing="ij")
    d0 = (kh_perturbation(Xp, p4[0], p4[2]) + kh_perturbation(Zp, p4[4], p4[6]))[:, None, :]
    d1 = (kh_perturbation(Xp, p4[1], p4[3]) + kh_perturbation(Zp, p4[5], p4[7]))[:, None, :]
    Y = x[None, :, None]
    band = (Y > 0.25 + d0) & (Y < 0.75 + d1)
    rho = np.where(band, 2.0, 1.0)
    ux = np.where(band, -0.5, 0.5)
    uy = np.zeros_like(rho)
    uz = np.zeros_like(rho)
    p = np.full_like(rho, 2.5)
    return conserved_from_primitive(rho, [ux, uy, uz], p, gamma)


def kh_params(seed, sample=0, count=40):
    """Seeded U[0,1) perturbation parameters for single runs and tests:
    component c of sample s is draw s*count+c of one numpy MT19937 stream."""
    rng = np.random.Generator(np.random.MT19937(seed))
    rng.random(sample * count)
    return rng.random(count)
