"""ctypes front end of the CPU oracle (test infrastructure).

The oracle (oracle/oracle.cpp) restates the alsvinn/alsvinn CPU path for one
finite-volume step.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module: it is the checker and the timed CPU
baseline, never the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the oracle (the mutation tests load
# deliberately broken copies through it).
_LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

EQ = {"burgers": 0, "euler1": 1, "euler2": 2, "euler3": 3}
RECON = {"none": 0, "weno2": 1, "eno2": 2, "eno3": 3}
FLUX = {"godunov": 0, "hll": 1, "hll3": 2}
INTEGRATOR = {"forwardeuler": 0, "rungekutta2": 1, "rungekutta3": 2}
BC = {"periodic": 0, "neumann": 1}
NVARS = {"burgers": 1, "euler1": 3, "euler2": 4, "euler3": 5}


class OracleConfig(ctypes.Structure):
    """Same field layout as avh_config in include/alsvinn_hip.h."""

    _fields_ = [
        ("equation", ctypes.c_int),
        ("reconstruction", ctypes.c_int),
        ("flux", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("boundary", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("lower", ctypes.c_double * 3),
        ("upper", ctypes.c_double * 3),
        ("gamma", ctypes.c_double),
        ("cfl", ctypes.c_double),
    ]


def make_config(equation, reconstruction, flux, integrator, boundary, n,
                lower=(0.0, 0.0, 0.0), upper=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.45):
    c = OracleConfig()
    c.equation = EQ[equation]
    c.reconstruction = RECON[reconstruction]
    c.flux = FLUX[flux]
    c.integrator = INTEGRATOR[integrator]
    c.boundary = BC[boundary]
    n = tuple(n) + (1,) * (3 - len(n))
    for d in range(3):
        c.n[d] = int(n[d])
        c.lower[d] = float(lower[d])
        c.upper[d] = float(upper[d])
    c.gamma = float(gamma)
    c.cfl = float(cfl)
    return c


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        d = ctypes.c_double
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_sim_create.restype = ctypes.c_void_p
        L.oracle_sim_create.argtypes = [P(OracleConfig)]
        L.oracle_sim_destroy.argtypes = [ctypes.c_void_p]
        for f in ("oracle_sim_ghosts", "oracle_sim_nvars"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_int
        L.oracle_sim_cells.argtypes = [ctypes.c_void_p]
        L.oracle_sim_cells.restype = ctypes.c_size_t
        L.oracle_sim_set_whole_domain.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_sim_set_state.argtypes = [ctypes.c_void_p, P(d)]
        L.oracle_sim_get_state.argtypes = [ctypes.c_void_p, P(d)]
        L.oracle_sim_add_fixed_interval.argtypes = [ctypes.c_void_p, d, d, ctypes.c_int]
        L.oracle_sim_set_time.argtypes = [ctypes.c_void_p, d, ctypes.c_long]
        for f in ("oracle_sim_time", "oracle_sim_last_dt"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = d
        for f in ("oracle_sim_steps", "oracle_sim_saves"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_long
        L.oracle_sim_step.argtypes = [ctypes.c_void_p]
        L.oracle_sim_run.argtypes = [ctypes.c_void_p, d, ctypes.c_long]
        L.oracle_compute_flux.argtypes = [P(OracleConfig), P(d), P(d), P(d)]
        L.oracle_system.argtypes = [P(OracleConfig), P(d), P(d), P(d), ctypes.c_int]
        L.oracle_reconstruct.argtypes = [P(OracleConfig), P(d), ctypes.c_int, P(d), P(d)]
        ivec = ctypes.c_int * 3
        L.oracle_compute_flux_box.argtypes = [P(OracleConfig), P(d), P(d), P(d), ivec, ivec]
        L.oracle_reconstruct_box.argtypes = [P(OracleConfig), P(d), ctypes.c_int, ctypes.c_int, P(d), P(d),
                                             ivec, ivec]
        L.oracle_interface_flux.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, d, P(d), P(d), P(d)]
        L.oracle_interface_flux.restype = d
        L.oracle_point_flux.argtypes = [ctypes.c_int, ctypes.c_int, d, P(d), P(d)]
        L.oracle_apply_boundary.argtypes = [P(OracleConfig), ctypes.c_int, P(d)]
        L.oracle_obeys.argtypes = [ctypes.c_void_p, P(d)]
        L.oracle_from_primitive.argtypes = [ctypes.c_int, d, P(d), P(d), ctypes.c_size_t]
        L.oracle_point_flux_all.argtypes = [ctypes.c_int, ctypes.c_int, P(d), P(d)]
        L.oracle_flux_all.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, d, P(d), P(d), P(d)]
        L.oracle_flux_all.restype = d
        L.oracle_integrate_ode.argtypes = [ctypes.c_int, d, ctypes.c_long, d]
        L.oracle_integrate_ode.restype = d
        L.oracle_mt19937_64.argtypes = [ctypes.c_uint64] * 3 + [P(ctypes.c_uint64)]
        L.oracle_uniform01.argtypes = [ctypes.c_uint64] * 3 + [P(d)]
        L.oracle_kinetic_energy.argtypes = [P(d), ctypes.c_long, ctypes.c_int, d]
        L.oracle_kinetic_energy.restype = d
        L.oracle_structure_cube.argtypes = [P(d), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, d, P(d)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ghosts_for(reconstruction):
    return {"none": 1, "weno2": 2, "eno2": 2, "eno3": 3}[reconstruction]


def ghosted_shape(cfg_n, g):
    """(nz, ny, nx) of a ghosted volume: ghosts only on axes with n > 1."""
    n = list(cfg_n) + [1] * (3 - len(cfg_n))
    dim = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
    gx, gy, gz = g, (g if dim > 1 else 0), (g if dim > 2 else 0)
    return (n[2] + 2 * gz, n[1] + 2 * gy, n[0] + 2 * gx), (gx, gy, gz)


class OracleError(RuntimeError):
    pass


class Simulator:
    """Reference-structure CPU simulator (Simulator + ConservedSystem +
    NumericalFluxCPU + Integrator + BoundaryCPU)."""

    def __init__(self, equation, reconstruction, flux, integrator, boundary, n,
                 lower=(0.0, 0.0, 0.0), upper=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.45):
        self.cfg = make_config(equation, reconstruction, flux, integrator, boundary, n,
                               lower, upper, gamma, cfl)
        self.equation = equation
        self.nv = NVARS[equation]
        h = lib().oracle_sim_create(ctypes.byref(self.cfg))
        if not h:
            raise OracleError(lib().oracle_last_error().decode())
        self._h = ctypes.c_void_p(h)
        self.ghosts = lib().oracle_sim_ghosts(self._h)
        self.shape, self.ghost_axes = ghosted_shape([self.cfg.n[d] for d in range(3)], self.ghosts)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.oracle_sim_destroy(h)
            self._h = None

    def set_whole_domain(self, on=True):
        lib().oracle_sim_set_whole_domain(self._h, int(bool(on)))

    def interior(self, ghosted):
        gx, gy, gz = self.ghost_axes
        nz, ny, nx = self.shape
        return ghosted[:, gz:nz - gz, gy:ny - gy, gx:nx - gx]

    def embed(self, interior):
        """Interior (nv, nz, ny, nx) -> ghosted array with zero ghosts."""
        out = np.zeros((self.nv,) + self.shape)
        gx, gy, gz = self.ghost_axes
        nz, ny, nx = self.shape
        out[:, gz:nz - gz, gy:ny - gy, gx:nx - gx] = interior
        return out

    def set_state(self, ghosted):
        a = np.ascontiguousarray(ghosted, dtype=np.float64).reshape((self.nv,) + self.shape)
        lib().oracle_sim_set_state(self._h, _dp(a))

    def set_interior(self, interior):
        self.set_state(self.embed(np.asarray(interior, dtype=np.float64).reshape(
            (self.nv,) + tuple(self.cfg.n[d] for d in (2, 1, 0)))))

    def get_state(self):
        out = np.empty((self.nv,) + self.shape)
        lib().oracle_sim_get_state(self._h, _dp(out))
        return out

    def get_interior(self):
        return np.ascontiguousarray(self.interior(self.get_state()))

    def add_fixed_interval(self, interval, start=0.0, write_initial=False):
        lib().oracle_sim_add_fixed_interval(self._h, float(interval), float(start), int(bool(write_initial)))

    def set_time(self, t, steps=0):
        lib().oracle_sim_set_time(self._h, float(t), int(steps))

    @property
    def time(self):
        return lib().oracle_sim_time(self._h)

    @property
    def steps(self):
        return lib().oracle_sim_steps(self._h)

    @property
    def last_dt(self):
        return lib().oracle_sim_last_dt(self._h)

    @property
    def saves(self):
        return lib().oracle_sim_saves(self._h)

    def step(self):
        rc = lib().oracle_sim_step(self._h)
        if rc == 3:
            raise OracleError("Simulation state does not obey constraints! At time %r, number of "
                              "timesteps performed %d" % (self.time, self.steps))
        if rc != 0:
            raise OracleError(lib().oracle_last_error().decode())
        return self.last_dt

    def run(self, t_end, max_steps=1 << 62):
        rc = lib().oracle_sim_run(self._h, float(t_end), int(max_steps))
        if rc == 3:
            raise OracleError("Simulation state does not obey constraints! At time %r, number of "
                              "timesteps performed %d" % (self.time, self.steps))
        if rc != 0:
            raise OracleError(lib().oracle_last_error().decode())

    def obeys(self, ghosted):
        a = np.ascontiguousarray(ghosted, dtype=np.float64)
        return bool(lib().oracle_obeys(self._h, _dp(a)))


def compute_flux(cfg, U_ghosted, out_ghosted):
    U = np.ascontiguousarray(U_ghosted, dtype=np.float64)
    out = np.ascontiguousarray(out_ghosted, dtype=np.float64).copy()
    ws = np.zeros(3)
    rc = lib().oracle_compute_flux(ctypes.byref(cfg), _dp(U), _dp(out), _dp(ws))
    if rc:
        raise OracleError(lib().oracle_last_error().decode())
    return out, ws


def compute_flux_box(cfg, U_ghosted, out_ghosted, start, end):
    """NumericalFlux::computeFlux with the reference's start/end offsets."""
    U = np.ascontiguousarray(U_ghosted, dtype=np.float64)
    out = np.ascontiguousarray(out_ghosted, dtype=np.float64).copy()
    ws = np.zeros(3)
    rc = lib().oracle_compute_flux_box(ctypes.byref(cfg), _dp(U), _dp(out), _dp(ws), (ctypes.c_int * 3)(*start),
                                       (ctypes.c_int * 3)(*end))
    if rc:
        raise OracleError(lib().oracle_last_error().decode())
    return out, ws


def reconstruct_box(cfg, U_ghosted, direction, indicator, left, right, start, end):
    """Reconstruction::performReconstruction with an indicator variable and
    start/end offsets; left/right are the volumes written into."""
    U = np.ascontiguousarray(U_ghosted, dtype=np.float64)
    left = np.ascontiguousarray(left, dtype=np.float64).copy()
    right = np.ascontiguousarray(right, dtype=np.float64).copy()
    rc = lib().oracle_reconstruct_box(ctypes.byref(cfg), _dp(U), int(direction), int(indicator), _dp(left),
                                      _dp(right), (ctypes.c_int * 3)(*start), (ctypes.c_int * 3)(*end))
    if rc:
        raise OracleError(lib().oracle_last_error().decode())
    return left, right


def system(cfg, U_ghosted, out_ghosted, whole_domain=False):
    U = np.ascontiguousarray(U_ghosted, dtype=np.float64)
    out = np.ascontiguousarray(out_ghosted, dtype=np.float64).copy()
    ws = np.zeros(3)
    rc = lib().oracle_system(ctypes.byref(cfg), _dp(U), _dp(out), _dp(ws), int(whole_domain))
    if rc:
        raise OracleError(lib().oracle_last_error().decode())
    return out, ws


def reconstruct(cfg, U_ghosted, direction, left=None, right=None):
    U = np.ascontiguousarray(U_ghosted, dtype=np.float64)
    left = np.zeros_like(U) if left is None else np.ascontiguousarray(left, dtype=np.float64).copy()
    right = np.zeros_like(U) if right is None else np.ascontiguousarray(right, dtype=np.float64).copy()
    rc = lib().oracle_reconstruct(ctypes.byref(cfg), _dp(U), int(direction), _dp(left), _dp(right))
    if rc:
        raise OracleError(lib().oracle_last_error().decode())
    return left, right


def interface_flux(equation, flux, direction, left, right, gamma=1.4):
    nv = NVARS[equation]
    ql = np.ascontiguousarray(left, dtype=np.float64)
    qr = np.ascontiguousarray(right, dtype=np.float64)
    F = np.zeros(nv)
    ws = lib().oracle_interface_flux(EQ[equation], FLUX[flux], int(direction), float(gamma),
                                     _dp(ql), _dp(qr), _dp(F))
    return F, ws


def point_flux(equation, direction, q, gamma=1.4):
    nv = NVARS[equation]
    q = np.ascontiguousarray(q, dtype=np.float64)
    F = np.zeros(nv)
    lib().oracle_point_flux(EQ[equation], int(direction), float(gamma), _dp(q), _dp(F))
    return F


def apply_boundary(cfg, ghosts, U_ghosted):
    U = np.ascontiguousarray(U_ghosted, dtype=np.float64).copy()
    rc = lib().oracle_apply_boundary(ctypes.byref(cfg), int(ghosts), _dp(U))
    if rc:
        raise OracleError(lib().oracle_last_error().decode())
    return U


def from_primitive(equation, prim, gamma=1.4):
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    nv = NVARS[equation]
    ncells = prim.size // nv
    cons = np.empty_like(prim)
    lib().oracle_from_primitive(EQ[equation], float(gamma), _dp(prim), _dp(cons), ncells)
    return cons


def mt19937_64(seed, skip, n):
    out = np.empty(n, dtype=np.uint64)
    lib().oracle_mt19937_64(seed, skip, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out


def uniform01(seed, skip, n):
    out = np.empty(n)
    lib().oracle_uniform01(seed, skip, n, _dp(out))
    return out


def point_flux_all(nsd, direction, all_vars):
    """all_vars = (rho, m..., E, p, u...)"""
    a = np.ascontiguousarray(all_vars, dtype=np.float64)
    F = np.zeros(nsd + 2)
    lib().oracle_point_flux_all(int(nsd), int(direction), _dp(a), _dp(F))
    return F


def flux_all(nsd, flux, direction, left_all, right_all, gamma=1.4):
    aL = np.ascontiguousarray(left_all, dtype=np.float64)
    aR = np.ascontiguousarray(right_all, dtype=np.float64)
    F = np.zeros(nsd + 2)
    ws = lib().oracle_flux_all(int(nsd), FLUX[flux], int(direction), float(gamma), _dp(aL), _dp(aR), _dp(F))
    return F, ws


def integrate_ode(integrator, h, steps, u0=1.0):
    return lib().oracle_integrate_ode(INTEGRATOR[integrator], float(h), int(steps), float(u0))


def kinetic_energy(interior, component, dv):
    """sum of 0.5*rho*u_c^2 over the interior SoA state, times dv."""
    q = np.ascontiguousarray(interior, dtype=np.float64)
    ncells = q[0].size
    return lib().oracle_kinetic_energy(_dp(q), ncells, int(component), float(dv))


def structure_cube(field, boundary, number_of_h, p=2.0):
    """Structure function of one 1D (nx,) or 2D (ny, nx) field, reference loop order."""
    u = np.ascontiguousarray(field, dtype=np.float64)
    dims = 1 if u.ndim == 1 or u.shape[0] == 1 else 2
    ny, nx = (1, u.shape[-1]) if dims == 1 else u.shape[-2:]
    out = np.zeros(int(number_of_h))
    lib().oracle_structure_cube(_dp(u), int(nx), int(ny), dims, BC[boundary], int(number_of_h), float(p), _dp(out))
    return out
