"""NetCDF output in the reference's layout (alsfvm/src/io/NetCDFWriter.cpp:27-141).

One file per snapshot, named `<basename>_<k>.nc` with k counting the
snapshots written so far (alsfvm/src/io/io_utils.cpp:33-38); statistics
files are `<basename>_<statistic>_<k>.nc` (alsuq/src/config/Setup.cpp:250-259).
Each file holds
  dimension  t (1), x, y, z (interior cells)
  variable   time(t)            the simulation time of the snapshot
  variables  <name>(x, y, z)    one per conserved variable (rho, mx, my, mz,
                                E or u), double, interior cells only
The reference passes its x-fastest interior buffer to a variable declared
with dimensions (x, y, z); the same bytes are written here, so the files
read back identically with the reference's readers
(python/alsvinn/alsvinn.py).  The reference creates NetCDF-4 (HDF5) files;
no HDF5 library exists on this platform, so the files are classic NetCDF
(64-bit offset), which every NetCDF reader, including netCDF4-python and
xarray, opens the same way.
"""
import getpass
import os
import platform
import socket
import sys
import time as _time

import numpy as np
from scipy.io import netcdf_file

CONSERVED = {
    "burgers": ["u"],
    "euler1": ["rho", "mx", "E"],
    "euler2": ["rho", "mx", "my", "E"],
    "euler3": ["rho", "mx", "my", "mz", "E"],
}


def _cpu_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def basic_report():
    """The run report every reference NetCDF file carries as global text
    attributes `alsvinn_report.<key>` (alsfvm/src/io/netcdf_write_report.cpp:25-29,
    the keys of alsutils/src/make_basic_report.cpp:46-90), with this
    platform's values."""
    try:
        user = getpass.getuser()
    except Exception:
        user = "unknown"
    libc = " ".join(platform.libc_ver()).strip() or "unknown"
    report = {
        "software": "alsvinn_amd (alsvinn hip platform)",
        "softwareVersion": "hip",
        "currentWorkingDirectory": os.getcwd(),
        "operatingSystem": "%s %s" % (platform.system(), platform.release()),
        "username": user,
        "host": socket.gethostname(),
        "standardCLibrary": libc,
        "generatedAt": _time.strftime("%Y%m%dT%H%M%S", _time.localtime()),
        "CPU": _cpu_name(),
        "platform": "hip",
        "pythonVersion": sys.version.split()[0],
    }
    return {"alsvinn_report." + k: str(v) for k, v in report.items()}


def write_snapshot(filename, variables, time, attributes=None, report=True):
    """variables: {name: interior array of shape (nz, ny, nx)} (x fastest).
    The run report's attributes come first, then `attributes`, as the
    reference writes them."""
    arrs = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in variables.items()}
    shapes = {a.shape for a in arrs.values()}
    if len(shapes) != 1:
        raise ValueError("all variables need the same shape")
    nz, ny, nx = next(iter(shapes))
    f = netcdf_file(filename, "w", version=2)
    try:
        for k, v in (basic_report() if report else {}).items():
            setattr(f, k, v)
        for k, v in (attributes or {}).items():
            setattr(f, k, v)
        f.createDimension("t", 1)
        tv = f.createVariable("time", "d", ("t",))
        tv[:] = np.array([time], dtype=np.float64)
        f.createDimension("x", nx)
        f.createDimension("y", ny)
        f.createDimension("z", nz)
        for name, a in arrs.items():
            var = f.createVariable(name, "d", ("x", "y", "z"))
            # the x-fastest buffer under dimensions (x, y, z), as the reference writes it
            var[:] = a.reshape(-1).reshape(nx, ny, nz)
    finally:
        f.close()


def read_snapshot(filename):
    """(time, {name: interior array (nz, ny, nx)}) from a snapshot file."""
    f = netcdf_file(filename, "r", mmap=False)
    try:
        time = float(f.variables["time"][0])
        nx, ny, nz = f.dimensions["x"], f.dimensions["y"], f.dimensions["z"]
        out = {}
        for name, var in f.variables.items():
            if name == "time":
                continue
            out[name] = np.array(var[:], dtype=np.float64).reshape(-1).reshape(nz, ny, nx)
        return time, out
    finally:
        f.close()


class NetCDFWriter:
    """The reference's NetCDFWriter: every write() makes the next numbered
    file."""

    def __init__(self, basename, attributes=None):
        self.basename = basename
        self.snapshot = 0
        self.attributes = dict(attributes or {})

    def filename(self, k=None):
        return "%s_%d.nc" % (self.basename, self.snapshot if k is None else k)

    def write(self, variables, time):
        name = self.filename()
        d = os.path.dirname(name)
        if d:
            os.makedirs(d, exist_ok=True)
        write_snapshot(name, variables, time, self.attributes)
        self.snapshot += 1
        return name
