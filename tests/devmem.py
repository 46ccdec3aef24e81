"""Device buffers for the tests of the Volume entry points, through the HIP
runtime's C API (hipMalloc / hipMemcpy / hipFree), and reference Volumes
(one device array per conserved variable, ghost cells included) built from
numpy arrays of shape (nvars, nz_total, ny_total, nx_total)."""
import ctypes

import numpy as np

import alsvinn_amd
from alsvinn_amd import _abi

_hip = None
H2D, D2H = 1, 2


def hip():
    """The HIP runtime the hip platform library runs on.  Loaded by soname
    after the library, so that it is the runtime object already in the
    process (torch, when imported first, brings its own build of it, and a
    second runtime from another path would not match torch's HSA runtime)."""
    global _hip
    if _hip is None:
        alsvinn_amd.load()
        L = ctypes.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _hip = L
    return _hip


class DeviceArray:
    def __init__(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.shape = host.shape
        self.nbytes = host.nbytes
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)) == 0
        self.ptr = p
        self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.nbytes == self.nbytes
        assert hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    def download(self):
        out = np.empty(self.shape)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and _hip is not None:
            _hip.hipFree(self.ptr)
            self.ptr = None


class DeviceVolume:
    """A reference Volume on the device: host array (nv, tz, ty, tx)."""

    def __init__(self, host, n, ghosts):
        host = np.asarray(host, dtype=np.float64)
        self.arrays = [DeviceArray(host[v]) for v in range(host.shape[0])]
        self.desc = _abi.AvhVolume()
        for v, a in enumerate(self.arrays):
            self.desc.var[v] = a.ptr.value
        self.desc.nvars = len(self.arrays)
        for d in range(3):
            self.desc.n[d] = int(n[d])
            self.desc.ghosts[d] = int(ghosts[d])

    def download(self):
        return np.stack([a.download() for a in self.arrays])

    def upload(self, host):
        for v, a in enumerate(self.arrays):
            a.upload(host[v])
