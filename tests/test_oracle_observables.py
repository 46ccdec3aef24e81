"""Observable oracles pinned by the reference's own structure-function tests
(test/library_tests/src/uq/statistics/StructureTestBoundary.cpp: TestConstant
:189-212, TestLinear :214-315 on N = 16, TestLinear2D :317-628 on 16x16, all
with numberOfH = 8 and p = 1..8, periodic and neumann) and by closed forms of
the kinetic-energy functional."""
import numpy as np
import pytest

from oracle import pyoracle

NH = 8


def _close(expected, got, p):
    # the reference asserts ASSERT_DOUBLE_EQ (4 ulp) on the CPU for p < 7 and
    # eps * 1e6 relative otherwise
    if p < 7:
        assert got == pytest.approx(expected, rel=4 * np.finfo(float).eps, abs=0)
    else:
        assert abs(got - expected) <= np.finfo(float).eps * 1e6 * expected


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
@pytest.mark.parametrize("p", range(1, 9))
def test_structure_constant_is_zero(bc, p):
    assert np.all(pyoracle.structure_cube(np.full(16, 3.25), bc, NH, p) == 0)
    assert np.all(pyoracle.structure_cube(np.full((16, 16), -1.5), bc, NH, p) == 0)


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
@pytest.mark.parametrize("p", range(1, 9))
def test_structure_linear_1d_reference_values(bc, p):
    N = 16
    got = pyoracle.structure_cube(np.arange(N, dtype=float), bc, NH, p)
    for h in range(NH):
        correct = 0.0
        for i in range(N):
            for j in (-h, h):
                if 0 <= i + j < N:
                    correct += abs(-j) ** p
                elif i + j < 0:
                    correct += abs(i - (N + (i + j))) ** p if bc == "periodic" else abs(i) ** p
                else:
                    correct += abs(i - (i + j - N)) ** p if bc == "periodic" else abs(i - (N - 1)) ** p
        correct /= N
        if h == 0:
            assert got[0] == 0  # the reference loops h from 1; its h = 0 entry stays zero
            continue
        _close(correct, got[h], p)


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
@pytest.mark.parametrize("p", range(1, 9))
def test_structure_linear_2d_reference_values(bc, p):
    nx = ny = 16
    j, i = np.mgrid[0:ny, 0:nx]
    u = (i + nx * j).astype(float)
    got = pyoracle.structure_cube(u, bc, NH, p)

    def boundary_index(n, i, offset, left, right, factor):
        v = -factor * offset
        if i + offset < 0:
            v = factor * (i - (n + (i + offset))) if bc == "periodic" else factor * i - left
        elif i + offset >= n:
            v = factor * (i - (i + offset - n)) if bc == "periodic" else factor * i - right
        return v

    for h in range(1, NH):
        correct = 0.0
        for jj in range(ny):
            for ii in range(nx):
                def upd(x, y):
                    xo = boundary_index(nx, ii, x, 0, nx - 1, 1)
                    yo = boundary_index(ny, jj, y, 0, (ny - 1) * nx, nx)
                    return abs(xo + yo) ** p
                for x in (-h, h):
                    for y in range(-h, h + 1):
                        correct += upd(x, y)
                        if -h < y < h:
                            correct += upd(y, x)
        correct /= nx * ny
        _close(correct, got[h], p)


def test_structure_2d_matches_vectorised_restatement():
    """Brute force by whole-array shifts (np.roll / clamped indexing) in a
    different summation order: equal to rounding."""
    rng = np.random.default_rng(3)
    u = rng.random((24, 20))
    for bc in ("periodic", "neumann"):
        got = pyoracle.structure_cube(u, bc, 9, 2.0)
        ny, nx = u.shape
        J, I = np.mgrid[0:ny, 0:nx]

        def at(dx, dy):
            if bc == "periodic":
                return u[(J + dy) % ny, (I + dx) % nx]
            return u[np.clip(J + dy, 0, ny - 1), np.clip(I + dx, 0, nx - 1)]
        for h in range(1, 9):
            ring = [(s * h, y) for s in (-1, 1) for y in range(-h + 1, h)] + \
                   [(x, s * h) for s in (-1, 1) for x in range(-h, h + 1)]
            assert len(ring) == 8 * h
            ref = sum(np.sum((u - at(dx, dy)) ** 2) for dx, dy in ring) / u.size
            assert got[h] == pytest.approx(ref, rel=1e-12)


def test_kinetic_energy_closed_forms():
    n = 32
    # uniform flow: rho = 2, uy = 0.3 on the unit square -> 0.5 * 2 * 0.09 = 0.09
    q = np.zeros((4, 1, n, n))
    q[0] = 2.0
    q[2] = 2.0 * 0.3
    q[3] = 5.0
    assert pyoracle.kinetic_energy(q, 1, 1.0 / (n * n)) == pytest.approx(0.09, rel=1e-14)
    assert pyoracle.kinetic_energy(q, 0, 1.0 / (n * n)) == 0.0
    # rho = 1, uy = sin(2 pi x): integral of 0.5 sin^2 = 1/4, exact for the
    # midpoint rule on a periodic grid
    x = (np.arange(n) + 0.5) / n
    q[0] = 1.0
    q[2] = np.sin(2 * np.pi * x)[None, :]
    assert pyoracle.kinetic_energy(q, 1, 1.0 / (n * n)) == pytest.approx(0.25, rel=1e-13)
