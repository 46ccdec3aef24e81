"""alsvinn_amd -- MI355X-native FP64 finite-volume engine (platform "hip").

The compute path is libalsvinn_hip.so (hand-written HIP kernels for gfx950
behind the C ABI in include/alsvinn_hip.h); this package is the host-side
mirror of the reference's Simulator surface on top of it.
"""
from ._abi import ConstraintError, HipError, InvalidArgument, LibraryMissing, SYMBOLS, load  # noqa: F401
from .simulator import (  # noqa: F401
    BOUNDARIES,
    CONSERVED_NAMES,
    EQUATIONS,
    FLUXES,
    GHOSTS,
    INTEGRATORS,
    NVARS,
    RECONSTRUCTIONS,
    Context,
    Simulator,
    resolve_cfl,
    resolve_integrator,
)
from . import initial_data  # noqa: F401
from .uq import BatchRunner, MeanVariance, generate_parameters  # noqa: F401
