"""The C-ABI library loads and exports every entry point include/alsvinn_hip.h
declares (no compute: runs without a GPU)."""
import os
import re
import subprocess

import pytest

import alsvinn_amd
from alsvinn_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "alsvinn_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(avh_[a-z0-9_]+)\s*\(", text)))


def test_library_loads():
    L = alsvinn_amd.load()
    assert L is not None


def test_every_declared_symbol_is_exported():
    L = alsvinn_amd.load()
    declared = header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_abi.SYMBOLS) == declared


def test_exports_via_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (avh_[a-z0-9_]+)", out))
    assert set(header_functions()) <= exported


def test_config_struct_layout_matches_header():
    # avh_config: 8 ints (equation..boundary, n[3]) then 8 doubles
    assert _abi.AvhConfig.n.offset == 20
    assert _abi.AvhConfig.lower.offset == 32
    assert _abi.AvhConfig.gamma.offset == 80
    assert _abi.ctypes.sizeof(_abi.AvhConfig) == 96


def test_unknown_names_raise():
    with pytest.raises(ValueError):
        alsvinn_amd.Simulator("euler2", "weno7", "hll3", "rungekutta3", "periodic", (8, 8, 1))


def test_auto_rules():
    assert alsvinn_amd.resolve_cfl("auto", "none") == 0.9
    assert alsvinn_amd.resolve_cfl("auto", "weno2") == 0.475
    assert alsvinn_amd.resolve_integrator("auto", "none") == "forwardeuler"
    assert alsvinn_amd.resolve_integrator("auto", "eno2") == "rungekutta2"
