"""The stage kernels replace some square roots and divisions by cheaper
sequences where the operands admit it (sqrt without range fix-ups; one
refined reciprocal shared by quotients with a common denominator).  These
must give the bits of the compiler's sqrt and '/' on the device.  Each kind
draws 10^8 operands over the range the kernels' guards admit, bounds
included, and counts results that differ in any bit.  The guards read
binary exponents with the frexp instruction: kind 4 checks it against the
exponent the bits give (0 for zeros, the true exponent for subnormals)."""
import ctypes

import pytest

import alsvinn_amd
from alsvinn_amd import _abi
from alsvinn_amd.simulator import default_context

pytestmark = pytest.mark.gpu

KINDS = {0: "sqrt on [2^-767, DBL_MAX]",
         1: "n / rho, rho in [2^-100, 2^100], n 0 or +-[2^-800, 2^600]",
         2: "WENO2 weight a0 / (a0 + a1), a0, a1 in [2^-400, 2^53]",
         3: "Roe quotient, denominator in [2^-50, 2^51]",
        4: "binary exponent (frexp) of signed zeros, subnormals and normals"}


@pytest.mark.parametrize("kind", sorted(KINDS))
def test_arith_shortcuts_bitwise(kind):
    L = alsvinn_amd.load()
    ctx = default_context()
    bad = ctypes.c_int64(-1)
    first = (ctypes.c_double * 2)()
    print("case arith kind", kind, KINDS[kind], flush=True)
    _abi.check(L.avh_arith_selftest(ctx._h, kind, 0x5EED + kind, 100_000_000, ctypes.byref(bad), first))
    assert bad.value == 0, "%d mismatches, e.g. operands %r %r" % (bad.value, first[0], first[1])


def test_arith_selftest_rejects_unknown_kind():
    L = alsvinn_amd.load()
    ctx = default_context()
    bad = ctypes.c_int64(0)
    with pytest.raises(_abi.InvalidArgument):
        _abi.check(L.avh_arith_selftest(ctx._h, 7, 1, 10, ctypes.byref(bad), None))
