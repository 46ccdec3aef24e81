"""ctypes binding of the C ABI in include/alsvinn_hip.h.

The shared library is loaded from the package directory only (built in-tree
by alsvinn_amd/build.py).  There is no fallback: if the library is missing or
cannot be loaded, every entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libalsvinn_hip.so")

# Exported entry points, in the order include/alsvinn_hip.h declares them.
SYMBOLS = [
    "avh_last_error",
    "avh_ctx_create",
    "avh_ctx_destroy",
    "avh_solver_create",
    "avh_solver_destroy",
    "avh_ghost_cells",
    "avh_num_vars",
    "avh_set_state",
    "avh_get_state",
    "avh_set_state_device",
    "avh_get_state_device",
    "avh_set_time",
    "avh_add_fixed_interval_adjuster",
    "avh_step",
    "avh_advance",
    "avh_ack_saves",
    "avh_get_progress",
    "avh_set_end_time",
    "avh_enqueue_steps",
    "avh_synchronize",
    "avh_set_graph_steps",
    "avh_prepare_graph",
    "avh_step_noalloc",
    "avh_set_tile_shape",
    "avh_autotune",
    "avh_time_step_kernels",
    "avh_stream",
    "avh_compute_flux",
    "avh_reconstruct",
    "avh_numerical_flux",
    "avh_reconstruction",
    "avh_apply_boundary",
    "avh_obeys_constraints",
    "avh_substep",
    "avh_arith_selftest",
    "avh_batch_create",
    "avh_batch_destroy",
    "avh_batch_add_fixed_interval_adjuster",
    "avh_batch_set_state",
    "avh_batch_get_state",
    "avh_batch_set_kh2d",
    "avh_batch_advance",
    "avh_batch_get_progress",
    "avh_batch_violations",
    "avh_batch_set_end_time",
    "avh_batch_kinetic_energy",
    "avh_batch_structure_cube",
    "avh_batch_set_graph_steps",
    "avh_batch_set_tile_shape",
    "avh_batch_autotune",
    "avh_batch_time_step_kernels",
    "avh_batch_get_statistics",
    "avh_uq_generate",
    "avh_stats_create",
    "avh_stats_destroy",
    "avh_stats_reset",
    "avh_stats_accumulate",
    "avh_stats_read",
    "avh_mlp_create",
    "avh_mlp_destroy",
    "avh_mlp_num_params",
    "avh_mlp_set_normalization",
    "avh_mlp_get_params",
    "avh_mlp_set_params",
    "avh_mlp_gradients",
    "avh_mlp_set_weight_decay",
    "avh_mlp_adam_step",
    "avh_mlp_train",
    "avh_mlp_predict",
]

AVH_OK, AVH_EINVAL, AVH_EHIP, AVH_ECONSTRAINT, AVH_ENOMEM = range(5)


class AvhConfig(ctypes.Structure):
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


class AvhMlpConfig(ctypes.Structure):
    _fields_ = [
        ("n_in", ctypes.c_int),
        ("width", ctypes.c_int),
        ("hidden", ctypes.c_int),
        ("n_out", ctypes.c_int),
        ("members", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float),
        ("seed", ctypes.c_uint64),
    ]


class AvhProgress(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double),
        ("last_dt", ctypes.c_double),
        ("steps", ctypes.c_int64),
        ("violated", ctypes.c_int32),
        ("saves_pending", ctypes.c_int32),
    ]


class AvhVolume(ctypes.Structure):
    _fields_ = [
        ("var", ctypes.c_void_p * 5),
        ("nvars", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("ghosts", ctypes.c_int * 3),
    ]


class AvhGroup(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("nsamples", ctypes.c_int)]


class LibraryMissing(RuntimeError):
    pass


_lib = None


def load():
    """Load libalsvinn_hip.so and declare every prototype.  Raises
    LibraryMissing if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            "%s is not built; run alsvinn_amd/build.py (or __graft_entry__.build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    P = ctypes.POINTER
    d = ctypes.c_double
    i64 = ctypes.c_int64
    st = ctypes.c_int
    L.avh_last_error.restype = ctypes.c_char_p
    L.avh_last_error.argtypes = []
    protos = {
        "avh_ctx_create": [ctypes.c_int, P(vp)],
        "avh_ctx_destroy": [vp],
        "avh_solver_create": [vp, P(AvhConfig), P(vp)],
        "avh_solver_destroy": [vp],
        "avh_ghost_cells": [vp, P(ctypes.c_int)],
        "avh_num_vars": [vp, P(ctypes.c_int)],
        "avh_set_state": [vp, vp],
        "avh_get_state": [vp, vp],
        "avh_set_state_device": [vp, vp],
        "avh_get_state_device": [vp, vp],
        "avh_set_time": [vp, d, i64],
        "avh_add_fixed_interval_adjuster": [vp, d, d, ctypes.c_int],
        "avh_step": [vp, P(d)],
        "avh_advance": [vp, d, i64, ctypes.c_int, ctypes.c_int, P(AvhProgress)],
        "avh_ack_saves": [vp],
        "avh_get_progress": [vp, P(AvhProgress)],
        "avh_set_end_time": [vp, d],
        "avh_enqueue_steps": [vp, i64],
        "avh_set_graph_steps": [vp, ctypes.c_int],
        "avh_prepare_graph": [vp],
        "avh_step_noalloc": [vp, P(AvhConfig), vp, d, vp, vp, vp],
        "avh_set_tile_shape": [vp, ctypes.c_int],
        "avh_autotune": [vp, P(ctypes.c_int), P(d), P(ctypes.c_int)],
        "avh_batch_set_tile_shape": [vp, ctypes.c_int],
        "avh_batch_set_end_time": [vp, d],
        "avh_batch_violations": [vp, ctypes.c_int, P(i64), P(i64)],
        "avh_batch_kinetic_energy": [vp, ctypes.c_int, ctypes.c_int, vp],
        "avh_batch_structure_cube": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, d, vp],
        "avh_batch_autotune": [vp, P(ctypes.c_int), P(d), P(ctypes.c_int)],
        "avh_batch_set_graph_steps": [vp, ctypes.c_int],
        "avh_batch_time_step_kernels": [vp, i64, P(d), P(ctypes.c_int)],
        "avh_synchronize": [vp],
        "avh_compute_flux": [vp, vp, P(d)],
        "avh_reconstruct": [vp, ctypes.c_int, vp, vp],
        "avh_arith_selftest": [vp, ctypes.c_int, ctypes.c_uint64, i64, P(i64), P(d)],
        "avh_numerical_flux": [vp, P(AvhConfig), P(AvhVolume), P(AvhVolume), P(ctypes.c_int), P(ctypes.c_int),
                               P(d), vp],
        "avh_reconstruction": [vp, P(AvhConfig), P(AvhVolume), ctypes.c_int, ctypes.c_int, P(AvhVolume),
                               P(AvhVolume), P(ctypes.c_int), P(ctypes.c_int), vp],
        "avh_apply_boundary": [vp, ctypes.c_int, P(AvhVolume), vp],
        "avh_obeys_constraints": [vp, P(AvhConfig), P(AvhVolume), P(ctypes.c_int), vp],
        "avh_substep": [vp, P(AvhConfig), P(P(AvhVolume)), ctypes.c_int, P(AvhVolume), d, d, P(d), vp],
        "avh_time_step_kernels": [vp, i64, P(d), P(ctypes.c_int)],
        "avh_batch_create": [vp, P(AvhConfig), ctypes.c_int, P(AvhGroup), d, ctypes.c_int, ctypes.c_int, P(vp)],
        "avh_batch_destroy": [vp],
        "avh_batch_add_fixed_interval_adjuster": [vp, d, d, ctypes.c_int],
        "avh_batch_set_state": [vp, ctypes.c_int, ctypes.c_int, vp],
        "avh_batch_get_state": [vp, ctypes.c_int, ctypes.c_int, vp],
        "avh_batch_set_kh2d": [vp, ctypes.c_int, vp],
        "avh_batch_advance": [vp, i64, ctypes.c_int, P(i64), P(i64)],
        "avh_batch_get_progress": [vp, ctypes.c_int, ctypes.c_int, P(AvhProgress)],
        "avh_batch_get_statistics": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, P(i64)],
        "avh_uq_generate": [ctypes.c_int, P(ctypes.c_int), P(d), P(d), i64, i64, vp],
        "avh_stats_create": [vp, ctypes.c_int, P(ctypes.c_int), d, ctypes.c_int, P(vp)],
        "avh_stats_destroy": [vp],
        "avh_stats_reset": [vp],
        "avh_stats_accumulate": [vp, vp, ctypes.c_int],
        "avh_stats_read": [vp, ctypes.c_int, ctypes.c_int, vp, vp, P(i64)],
        "avh_mlp_create": [vp, P(AvhMlpConfig), P(vp)],
        "avh_mlp_destroy": [vp],
        "avh_mlp_num_params": [vp, P(i64)],
        "avh_mlp_set_normalization": [vp, vp, vp, vp, vp],
        "avh_mlp_get_params": [vp, ctypes.c_int, vp],
        "avh_mlp_set_params": [vp, ctypes.c_int, vp],
        "avh_mlp_gradients": [vp, vp, vp, i64, vp, vp],
        "avh_mlp_adam_step": [vp, ctypes.c_float],
        "avh_mlp_set_weight_decay": [vp, ctypes.c_float],
        "avh_mlp_train": [vp, vp, vp, i64, ctypes.c_int, ctypes.c_float, ctypes.c_float, vp],
        "avh_mlp_predict": [vp, vp, i64, vp],
    }
    for name, args in protos.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = st
    L.avh_stream.argtypes = [vp]
    L.avh_stream.restype = vp
    _lib = L
    return L


class HipError(RuntimeError):
    """Raised for AVH_EHIP / AVH_ENOMEM / AVH_EINVAL."""


class InvalidArgument(ValueError):
    """AVH_EINVAL: a rejected argument (no such shape, group, sample ...)."""


class ConstraintError(RuntimeError):
    """AVH_ECONSTRAINT: the reference throws std::runtime_error from
    Simulator::checkConstraints (alsfvm/src/simulator/Simulator.cpp:180-189)."""


def check(status):
    if status == AVH_OK:
        return
    msg = load().avh_last_error().decode(errors="replace")
    if status == AVH_ECONSTRAINT:
        raise ConstraintError(msg)
    if status == AVH_EINVAL:
        raise InvalidArgument(msg)
    raise HipError("avh status %d: %s" % (status, msg))
