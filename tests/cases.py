"""Seeded initial states shared by the oracle and GPU parity tests."""
import numpy as np

from alsvinn_amd.initial_data import conserved_from_primitive, kelvin_helmholtz_2d, kh_params, midpoints

NV = {"burgers": 1, "euler1": 3, "euler2": 4, "euler3": 5}


def smooth_state(equation, n, seed=0, gamma=1.4, jump=False):
    """Positive density/pressure fields with smooth random modes (plus an
    optional contact/shear jump) on the (nz, ny, nx) interior."""
    rng = np.random.default_rng(seed)
    n = tuple(n) + (1,) * (3 - len(n))
    nx, ny, nz = n
    x = midpoints(nx)
    y = midpoints(ny) if ny > 1 else np.array([0.5])
    z = midpoints(nz) if nz > 1 else np.array([0.5])
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")

    def field(amp):
        f = np.zeros_like(X)
        for _ in range(3):
            kx, ky, kz = rng.integers(1, 4, size=3)
            ph = rng.random(3) * 2 * np.pi
            f = f + amp * rng.random() * np.sin(2 * np.pi * kx * X + ph[0]) * \
                np.cos(2 * np.pi * ky * Y * (ny > 1) + ph[1]) * np.cos(2 * np.pi * kz * Z * (nz > 1) + ph[2])
        return f

    if equation == "burgers":
        u = field(1.0)
        if jump:
            u = u + np.where(X > 0.5, -0.5, 0.5)
        return u[None].astype(np.float64)
    nsd = {"euler1": 1, "euler2": 2, "euler3": 3}[equation]
    rho = 1.5 + field(0.3)
    p = 2.0 + field(0.4)
    us = [field(0.3) for _ in range(nsd)]
    if jump:
        band = (Y > 0.3) & (Y < 0.7) if ny > 1 else (X > 0.3) & (X < 0.7)
        rho = np.where(band, rho + 0.8, rho)
        us[0] = np.where(band, us[0] - 0.5, us[0] + 0.5)
    return conserved_from_primitive(rho, us, p, gamma)


def kh2d(n, seed=1):
    return kelvin_helmholtz_2d(n, kh_params(seed))


def alternating_euler3(n):
    """CUDANumericalFluxTest::CompareAgainstCPU states: every other cell
    (1,1,1,1,10) and (2,2,2,2,40) in fill order
    (test/library_tests/src/fvm/flux/CUDANumericalFluxTest.cpp:233-251)."""
    nx, ny, nz = n
    q = np.empty((5, nz, ny, nx))
    k = 0
    for zz in range(nz):
        for yy in range(ny):
            for xx in range(nx):
                if k % 2:
                    q[:, zz, yy, xx] = (1, 1, 1, 1, 10)
                else:
                    q[:, zz, yy, xx] = (2, 2, 2, 2, 40)
                k += 1
    return q


def rel_l1(a, b):
    """Per-variable relative L1 distance sum|a-b| / sum|b|."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    num = np.abs(a - b).sum(axis=1)
    den = np.abs(b).sum(axis=1)
    return num / np.where(den > 0, den, 1.0)
