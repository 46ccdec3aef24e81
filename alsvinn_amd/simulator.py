"""Host-side mirror of the reference's simulator surface for the hip platform.

Names, argument meaning and error behaviour follow alsfvm:
  Simulator(...)                  Simulator::Simulator + factories
                                  (alsfvm/src/simulator/Simulator.cpp:26-64)
  set_simulation_state(vol)       Simulator::setSimulationState (:100-103)
  perform_step()                  Simulator::performStep (:81-85)
  at_end(), get_current_time()    Simulator::atEnd / getCurrentTime (:77-79, :95)
  add_fixed_interval_writer(...)  FixedIntervalWriter registered as writer and
                                  timestep adjuster (alsfvm/src/config/SimulatorSetup.cpp:205-228)
  compute_flux()                  NumericalFlux::computeFlux over the whole domain
  run()                           alsvinncli's `while (!atEnd()) performStep()`
A constraint violation raises ConstraintError (the reference throws
std::runtime_error from Simulator::checkConstraints, Simulator.cpp:180-189).

Every call goes through libalsvinn_hip.so; nothing here computes the step.
"""
import ctypes

import numpy as np

from . import _abi

EQUATIONS = {"burgers": 0, "euler1": 1, "euler2": 2, "euler3": 3}
RECONSTRUCTIONS = {"none": 0, "weno2": 1, "eno2": 2, "eno3": 3}
FLUXES = {"godunov": 0, "hll": 1, "hll3": 2}
INTEGRATORS = {"forwardeuler": 0, "rungekutta2": 1, "rungekutta3": 2}
BOUNDARIES = {"periodic": 0, "neumann": 1}
NVARS = {"burgers": 1, "euler1": 3, "euler2": 4, "euler3": 5}
GHOSTS = {"none": 1, "weno2": 2, "eno2": 2, "eno3": 3}
CONSERVED_NAMES = {
    "burgers": ["u"],
    "euler1": ["rho", "mx", "E"],
    "euler2": ["rho", "mx", "my", "E"],
    "euler3": ["rho", "mx", "my", "mz", "E"],
}


def resolve_cfl(cfl, reconstruction):
    """cfl "auto": 0.9 without reconstruction, else 0.475
    (alsfvm/src/config/SimulatorSetup.cpp:305-319)."""
    if isinstance(cfl, str):
        if cfl.strip() == "auto":
            return 0.9 if reconstruction == "none" else 0.475
        return float(cfl)
    return float(cfl)


def resolve_integrator(integrator, reconstruction):
    """integrator "auto": forwardeuler without reconstruction, else rungekutta2
    (alsfvm/src/config/SimulatorSetup.cpp:321-336)."""
    if integrator == "auto":
        return "forwardeuler" if reconstruction == "none" else "rungekutta2"
    return integrator


def _lookup(table, key, what):
    k = str(key).strip().lower()
    if k not in table:
        raise ValueError("Unknown %s \"%s\" (supported: %s)" % (what, key, ", ".join(sorted(table))))
    return table[k]


class Context:
    """A HIP device (avh_ctx)."""

    def __init__(self, device=0):
        L = _abi.load()
        h = ctypes.c_void_p()
        _abi.check(L.avh_ctx_create(int(device), ctypes.byref(h)))
        self._h = h
        self.device = device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _abi._lib is not None:
            _abi._lib.avh_ctx_destroy(h)
            self._h = None


_default_ctx = {}


def default_context(device=0):
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


class Simulator:
    def __init__(self, equation, reconstruction, flux, integrator, boundary, n,
                 lower=(0.0, 0.0, 0.0), upper=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.475,
                 end_time=float("inf"), platform="hip", device=0):
        if platform != "hip":
            raise ValueError("Unknown platform %s (this engine provides \"hip\")" % platform)
        integrator = resolve_integrator(integrator, reconstruction)
        self.equation = str(equation).strip().lower()
        self.reconstruction = str(reconstruction).strip().lower()
        self.flux_name = str(flux).strip().lower()
        self.integrator = integrator
        self.boundary = str(boundary).strip().lower()
        n = tuple(int(v) for v in n) + (1,) * (3 - len(n))
        self.n = n
        self.cfl = resolve_cfl(cfl, self.reconstruction)
        self.end_time = float(end_time)
        cfg = _abi.AvhConfig()
        cfg.equation = _lookup(EQUATIONS, equation, "equation")
        cfg.reconstruction = _lookup(RECONSTRUCTIONS, reconstruction, "reconstruction")
        cfg.flux = _lookup(FLUXES, flux, "flux")
        cfg.integrator = _lookup(INTEGRATORS, integrator, "integrator")
        cfg.boundary = _lookup(BOUNDARIES, boundary, "boundary")
        for d in range(3):
            cfg.n[d] = n[d]
            cfg.lower[d] = float(lower[d]) if d < len(lower) else 0.0
            cfg.upper[d] = float(upper[d]) if d < len(upper) else 1.0
        cfg.gamma = float(gamma)
        cfg.cfl = self.cfl
        self.cfg = cfg
        self.ctx = default_context(device)
        L = _abi.load()
        h = ctypes.c_void_p()
        _abi.check(L.avh_solver_create(self.ctx._h, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.nv = NVARS[self.equation]
        self.ghosts = GHOSTS[self.reconstruction]
        dim = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
        self.dim = dim
        g = self.ghosts
        self.ghost_axes = (g, g if dim > 1 else 0, g if dim > 2 else 0)
        gx, gy, gz = self.ghost_axes
        self.shape = (n[2] + 2 * gz, n[1] + 2 * gy, n[0] + 2 * gx)
        self.interior_shape = (n[2], n[1], n[0])
        _abi.check(L.avh_set_end_time(self._h, self.end_time))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _abi._lib is not None:
            _abi._lib.avh_solver_destroy(h)
            self._h = None

    # -- state -----------------------------------------------------------
    def embed(self, interior):
        out = np.zeros((self.nv,) + self.shape)
        gx, gy, gz = self.ghost_axes
        nz, ny, nx = self.shape
        out[:, gz:nz - gz, gy:ny - gy, gx:nx - gx] = np.asarray(interior).reshape(
            (self.nv,) + self.interior_shape)
        return out

    def interior(self, ghosted):
        gx, gy, gz = self.ghost_axes
        nz, ny, nx = self.shape
        return ghosted[:, gz:nz - gz, gy:ny - gy, gx:nx - gx]

    def set_simulation_state(self, ghosted):
        a = np.ascontiguousarray(ghosted, dtype=np.float64).reshape((self.nv,) + self.shape)
        _abi.check(_abi.load().avh_set_state(self._h, a.ctypes.data))

    def set_interior(self, interior):
        self.set_simulation_state(self.embed(interior))

    def get_state(self):
        out = np.empty((self.nv,) + self.shape)
        _abi.check(_abi.load().avh_get_state(self._h, out.ctypes.data))
        return out

    def get_interior(self):
        return np.ascontiguousarray(self.interior(self.get_state()))

    def set_state_device(self, dev_ptr):
        _abi.check(_abi.load().avh_set_state_device(self._h, ctypes.c_void_p(dev_ptr)))

    def get_state_device(self, dev_ptr):
        _abi.check(_abi.load().avh_get_state_device(self._h, ctypes.c_void_p(dev_ptr)))

    # -- time ------------------------------------------------------------
    def set_time(self, t, steps=0):
        _abi.check(_abi.load().avh_set_time(self._h, float(t), int(steps)))

    def add_fixed_interval_writer(self, interval, start_time=0.0, write_initial_timestep=False):
        _abi.check(_abi.load().avh_add_fixed_interval_adjuster(
            self._h, float(interval), float(start_time), int(bool(write_initial_timestep))))

    def set_end_time(self, t_end):
        self.end_time = float(t_end)
        _abi.check(_abi.load().avh_set_end_time(self._h, self.end_time))

    def progress(self):
        p = _abi.AvhProgress()
        _abi.check(_abi.load().avh_get_progress(self._h, ctypes.byref(p)))
        return p

    def get_current_time(self):
        return self.progress().time

    def get_end_time(self):
        return self.end_time

    def at_end(self):
        return self.get_current_time() >= self.end_time

    # -- stepping --------------------------------------------------------
    def perform_step(self):
        dt = ctypes.c_double()
        _abi.check(_abi.load().avh_step(self._h, ctypes.byref(dt)))
        return dt.value

    def run(self, end_time=None, max_steps=1 << 62, chunk=64, pause_on_save=False):
        if end_time is not None:
            self.end_time = float(end_time)
        p = _abi.AvhProgress()
        _abi.check(_abi.load().avh_advance(self._h, self.end_time, int(max_steps), int(chunk),
                                           int(bool(pause_on_save)), ctypes.byref(p)))
        return p

    def ack_saves(self):
        _abi.check(_abi.load().avh_ack_saves(self._h))

    def enqueue_steps(self, steps):
        _abi.check(_abi.load().avh_enqueue_steps(self._h, int(steps)))

    def synchronize(self):
        _abi.check(_abi.load().avh_synchronize(self._h))

    def set_tile_shape(self, shape):
        """Tile shape of the stage kernel (-1: autotune on first use)."""
        _abi.check(_abi.load().avh_set_tile_shape(self._h, int(shape)))

    def autotune(self):
        """Time every tile shape now; returns (best shape, [ms per shape])."""
        best = ctypes.c_int()
        n = ctypes.c_int()
        ms = (ctypes.c_double * 8)()
        _abi.check(_abi.load().avh_autotune(self._h, ctypes.byref(best), ms, ctypes.byref(n)))
        return best.value, [ms[i] for i in range(n.value)]

    def set_graph_steps(self, steps):
        """Steps per captured hipGraph (0: ordinary launches)."""
        _abi.check(_abi.load().avh_set_graph_steps(self._h, int(steps)))

    def prepare_graph(self):
        """Resolve the tile shape and capture the step-chunk graph now, so
        that later steps only replay it."""
        _abi.check(_abi.load().avh_prepare_graph(self._h))

    def stream(self):
        return _abi.load().avh_stream(self._h)

    def reconstruct(self, direction):
        """(left, right) face values of every interior cell along `direction`
        (Reconstruction::performReconstruction), each (nv, nz, ny, nx)."""
        shape = (self.nv, self.n[2], self.n[1], self.n[0])
        left = np.empty(shape)
        right = np.empty(shape)
        _abi.check(_abi.load().avh_reconstruct(self._h, int(direction), left.ctypes.data, right.ctypes.data))
        return left, right

    def compute_flux(self):
        """NumericalFlux::computeFlux on the current state: unscaled net flux
        (interior layout) and the per-direction max wave speeds."""
        out = np.empty((self.nv,) + self.interior_shape)
        ws = (ctypes.c_double * 3)()
        _abi.check(_abi.load().avh_compute_flux(self._h, out.ctypes.data, ws))
        return out, np.array(list(ws))
