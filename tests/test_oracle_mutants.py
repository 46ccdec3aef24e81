"""The oracle's pins catch plausible mistakes.  Each mutant is a copy of
oracle/oracle.cpp with one planted error (a dropped term, a wrong sign,
index or coefficient), compiled on the CPU; the independent pin suites
(test_oracle_known_answers: the reference's own known answers;
test_oracle_closed_forms: exact solutions, textbook forms and symmetries)
must fail against every mutant.  The golden trajectories are left out: the
oracle wrote them, so they pin nothing about it."""
import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MUTANTS = {
    "weno2_right_weight": ("const real a0R = 2 / (3 * (eps + b0) * (eps + b0));",
                           "const real a0R = 1 / (3 * (eps + b0) * (eps + b0));"),
    "hll3_star_pressure_sign": ("const real pm = pl + L.rho * (L.u[DIR] - sm) * udl;",
                                "const real pm = pl - L.rho * (L.u[DIR] - sm) * udl;"),
    "hll3_star_pressure_sign_right": ("const real pm = pr + R.rho * (R.u[DIR] - sm) * udr;",
                                      "const real pm = pr - R.rho * (R.u[DIR] - sm) * udr;"),
    "hll3_contact_pressure_sign": ("const real sm = (R.m[DIR] * udr - L.m[DIR] * udl + pr - pl) / aa;",
                                   "const real sm = (R.m[DIR] * udr - L.m[DIR] * udl + pl - pr) / aa;"),
    "hll3_us_swapped_speeds": ("us[c] = (fr[1 + c] - fl[1 + c] - R.m[c] * sr + L.m[c] * sl) / aa;",
                               "us[c] = (fr[1 + c] - fl[1 + c] - R.m[c] * sl + L.m[c] * sr) / aa;"),
    "hll_jump_sign": ("F[i] = (sr * fl[i] - sl * fr[i] + srsl * (ur[i] - ul[i])) / den;",
                      "F[i] = (sr * fl[i] - sl * fr[i] + srsl * (ul[i] - ur[i])) / den;"),
    "rk3_stage2_coefficient": ("vol_mul(out, 3. / 4.);", "vol_mul(out, 1. / 4.);"),
    "periodic_off_by_one": ("at(x + sign * (-k + inx) * xDir,", "at(x + sign * (-k + inx + 1) * xDir,"),
    "neumann_mirror_index": ("at(x + sign * (k - 1) * xDir, y + sign * (k - 1) * yDir,",
                             "at(x + sign * k * xDir, y + sign * (k - 1) * yDir,"),
    "dot_dropped_term": ("for (int c = 1; c < NSD_; ++c) s = s + m[c] * m[c];",
                         "for (int c = 1; c < NSD_; ++c) s = s + m[c] * m[0];"),
    "pressure_gamma": ("a.p = (gamma - 1) * ie;", "a.p = gamma * ie;"),
    "net_flux_sign": ("const real flm = (-1.0) * temp.var(v)[im];", "const real flm = (1.0) * temp.var(v)[im];"),
    "eno3_coefficient": ("{real(1.0 / 3.0), real(5.0 / 6.0), real(-1.0 / 6.0)}",
                         "{real(1.0 / 3.0), real(5.0 / 6.0), real(1.0 / 6.0)}"),
    "cfl_length_inverted": ("total += w / cell[d];", "total += w * cell[d];"),
    "godunov_left_clip": ("eq.point_flux<DIR>(Burgers::All{std::fmax(L.u, real(0.0))}, fl);",
                          "eq.point_flux<DIR>(Burgers::All{std::fmin(L.u, real(0.0))}, fl);"),
    "eno_shift_comparison": ("if (std::fabs(ddl) < std::fabs(ddr)) shift++;", "if (std::fabs(ddl) > std::fabs(ddr)) shift++;"),
}


def _build(name, tmp):
    old, new = MUTANTS[name]
    src = open(os.path.join(ROOT, "oracle", "oracle.cpp")).read()
    assert src.count(old) == 1, name
    d = os.path.join(tmp, name)
    os.makedirs(os.path.join(d, "oracle"))
    os.makedirs(os.path.join(d, "include"))
    shutil.copy(os.path.join(ROOT, "include", "alsvinn_hip.h"), os.path.join(d, "include"))
    with open(os.path.join(d, "oracle", "oracle.cpp"), "w") as f:
        f.write(src.replace(old, new))
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-o", lib,
                    os.path.join(d, "oracle", "oracle.cpp")], check=True, capture_output=True)
    return name, lib


@pytest.fixture(scope="module")
def mutant_libs():
    tmp = tempfile.mkdtemp(prefix="oracle_mutants_")
    try:
        with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            libs = dict(ex.map(lambda n: _build(n, tmp), MUTANTS))
        yield libs
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_unmutated_source_matches_patterns():
    src = open(os.path.join(ROOT, "oracle", "oracle.cpp")).read()
    for name, (old, _) in MUTANTS.items():
        assert src.count(old) == 1, name


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_mutant(mutant_libs, name):
    env = dict(os.environ, ORACLE_LIB=mutant_libs[name], OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "tests/test_oracle_known_answers.py", "tests/test_oracle_closed_forms.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode != 0, "mutant %s passed every pin:\n%s" % (name, r.stdout[-2000:])
