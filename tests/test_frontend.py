"""XML front end and NetCDF output (CPU): reference node names and auto
rules, both forms of Python initial data, cell midpoints in the reference's
operation order, and the NetCDF file layout."""
import os

import numpy as np
import pytest

from alsvinn_amd import config as C
from alsvinn_amd.initial_data import kelvin_helmholtz_2d
from alsvinn_amd.netcdf import NetCDFWriter, read_snapshot, write_snapshot
from alsvinn_amd.uq import generate_parameters

HERE = os.path.join(os.path.dirname(__file__), "configs")


def test_read_sod_config_auto_rules():
    cfg = C.read_config(os.path.join(HERE, "sod.xml"))
    assert cfg["equation"] == "euler1" and cfg["flux"] == "hll3" and cfg["boundary"] == "neumann"
    assert cfg["dimension"] == [200, 1, 1] and cfg["lower"] == [-5, 0, 0] and cfg["upper"] == [5, 0, 0]
    assert cfg["cfl"] == 0.475 and cfg["integrator"] == "rungekutta2"  # auto with weno2
    assert cfg["writer"] == {"type": "netcdf", "basename": "out/sod", "number_of_saves": 4, "write_initial": True}
    assert cfg["init_parameters"] == {"X": [0.25]}
    assert cfg["uq"] is None and cfg["gamma"] == 1.4


def test_read_kh_uq_config():
    cfg = C.read_config(os.path.join(HERE, "kh.xml"))
    assert cfg["platform"] == "cuda"
    assert cfg["init_parameters"]["a"] == [0.0] * 40  # one listed value fills the length
    u = cfg["uq"]
    assert u["samples"] == 6 and u["sample_start"] == 0 and u["parameters"] == [("a", 40, 0.0, 1.0)]
    assert u["stats"][0]["number_of_saves"] == 1 and u["stats"][0]["write_initial"] is True
    assert u["stats"][0]["basename"] == "out/kh"


def test_cell_midpoints_reference_order():
    xs, ys, zs = C.cell_midpoints([-5.0, 0.0, 0.0], [5.0, 0.0, 0.0], (200, 1, 1))
    length = 10.0 / 200
    assert np.array_equal(xs, -5.0 + length * np.arange(200) + length / 2.0)
    assert ys.tolist() == [0.5] and zs.tolist() == [0.5]  # single-cell axes have length 1


def test_snippet_initial_data_sod():
    cfg = C.read_config(os.path.join(HERE, "sod.xml"))
    q = C.initial_state(cfg)
    assert q.shape == (3, 1, 1, 200)
    xs = C.cell_midpoints(cfg["lower"], cfg["upper"], (200, 1, 1))[0]
    left = xs < 0.25
    np.testing.assert_array_equal(q[0, 0, 0], np.where(left, 1.0, 0.125))
    np.testing.assert_array_equal(q[1, 0, 0], 0.0)
    np.testing.assert_array_equal(q[2, 0, 0], np.where(left, 1.0 / (1.4 - 1), 0.1 / (1.4 - 1)))


def test_init_global_kh_equals_builtin_initial_data():
    cfg = C.read_config(os.path.join(HERE, "kh.xml"))
    a = generate_parameters([("a", 40, 0.0, 1.0)], 0, 3)["a"]
    for s in range(3):
        q = C.initial_state(cfg, {"a": list(a[s])})
        np.testing.assert_array_equal(q, kelvin_helmholtz_2d(32, a[s]))


def test_equation_parameters_reach_initial_data(tmp_path):
    (tmp_path / "g.py").write_text("u = gamma\n")
    (tmp_path / "g.xml").write_text(open(os.path.join(HERE, "sod.xml")).read()
                                    .replace("sod.py", "g.py").replace("euler1", "burgers")
                                    .replace("<gamma>1.4</gamma>", "<gamma>1.7</gamma>"))
    cfg = C.read_config(str(tmp_path / "g.xml"))
    np.testing.assert_array_equal(C.initial_state(cfg), 1.7)


@pytest.mark.parametrize("edit,msg", [
    (("<name>none</name>", "<name>tecno</name>"), "diffusion"),
    (("<type>netcdf</type>", "<type>hdf5</type>"), None),
    (("<boundary>neumann</boundary>", "<boundary>wall</boundary>"), "boundary"),
    (("<numberOfSaves>4</numberOfSaves>", "<numberOfCoarseSaves>4</numberOfCoarseSaves>"), "coarse"),
])
def test_unsupported_nodes_are_rejected(tmp_path, edit, msg):
    (tmp_path / "sod.py").write_text(open(os.path.join(HERE, "sod.py")).read())
    (tmp_path / "c.xml").write_text(open(os.path.join(HERE, "sod.xml")).read().replace(*edit, 1))
    if msg is None:
        cfg = C.read_config(str(tmp_path / "c.xml"))  # the writer type is checked by the CLI
        assert cfg["writer"]["type"] == "hdf5"
        return
    with pytest.raises(C.ConfigError, match=msg):
        C.read_config(str(tmp_path / "c.xml"))


def test_netcdf_layout_and_round_trip(tmp_path):
    from scipy.io import netcdf_file
    nz, ny, nx = 1, 3, 4
    rho = np.arange(nx * ny, dtype=float).reshape(nz, ny, nx)
    f = str(tmp_path / "s_0.nc")
    write_snapshot(f, {"rho": rho, "E": -rho}, 0.75, {"fvm_name": "t"})
    nc = netcdf_file(f, "r", mmap=False)
    assert nc.dimensions["x"] == nx and nc.dimensions["y"] == ny and nc.dimensions["z"] == nz
    assert nc.variables["rho"].dimensions == ("x", "y", "z")
    # the x-fastest buffer under dimensions (x, y, z), as the reference writes it
    np.testing.assert_array_equal(np.array(nc.variables["rho"][:]).ravel(), rho.ravel())
    assert float(nc.variables["time"][0]) == 0.75 and nc.fvm_name == b"t"
    nc.close()
    t, v = read_snapshot(f)
    assert t == 0.75
    np.testing.assert_array_equal(v["rho"], rho)
    np.testing.assert_array_equal(v["E"], -rho)


def test_netcdf_writer_numbers_snapshots(tmp_path):
    w = NetCDFWriter(str(tmp_path / "sub" / "run"))
    names = [w.write({"u": np.zeros((1, 1, 5))}, t) for t in (0.0, 0.5)]
    assert [os.path.basename(n) for n in names] == ["run_0.nc", "run_1.nc"]


def test_cli_decomposition_flags_accepted_and_checked():
    """alsvinncli / alsuqcli accept the reference's MPI options
    (alsvinncli/src/main.cpp:55-67, alsuqcli/src/main.cpp:40-51); one GPU
    means every count must be 1."""
    import argparse
    from alsvinn_amd.alsvinncli import add_decomposition_flags, check_decomposition
    for sample in (False, True):
        ap = argparse.ArgumentParser()
        ap.add_argument("input")
        add_decomposition_flags(ap, sample=sample)
        args = ["cfg.xml", "--multi-x", "1", "--multi-y", "1", "--multi-z", "1"]
        if sample:
            args += ["--multi-sample", "1"]
        else:
            args += ["-x"]
        a = ap.parse_args(args)
        check_decomposition(ap, a, sample=sample)
        bad = ["cfg.xml", "--multi-sample" if sample else "--multi-y", "2"]
        with pytest.raises(SystemExit):
            check_decomposition(ap, ap.parse_args(bad), sample=sample)


def test_netcdf_files_carry_the_run_report(tmp_path):
    """alsvinn_report.* global attributes (alsfvm/src/io/netcdf_write_report.cpp:25-29),
    then the writer's own attributes."""
    from scipy.io import netcdf_file
    from alsvinn_amd.netcdf import NetCDFWriter
    w = NetCDFWriter(str(tmp_path / "kh"), {"fvm_name": "kh"})
    fn = w.write({"rho": np.ones((1, 3, 4))}, 0.25)
    f = netcdf_file(fn, "r", mmap=False)
    try:
        attrs = {k: (v.decode() if isinstance(v, bytes) else v) for k, v in f._attributes.items()}
    finally:
        f.close()
    for key in ("software", "host", "CPU", "generatedAt", "currentWorkingDirectory", "operatingSystem"):
        assert attrs["alsvinn_report." + key]
    assert attrs["fvm_name"] == "kh"
