"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host
code only): oracle/oracle.cpp built with -fsanitize=address,undefined, the
independent pin suites and the golden trajectories run against that build
with the sanitizer runtimes preloaded into the Python process, and must
pass with no report.  A copy with a planted out-of-bounds access must fail,
so that a silent sanitizer setup cannot pass for a clean one."""
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SUITES = ["tests/test_oracle_known_answers.py", "tests/test_oracle_closed_forms.py", "tests/test_golden.py"]
# NoReconstruction's copy loop one element too far: the last variable's
# left/right arrays are written one past their end
PLANTED = ("for (size_t i = 0; i < in.cells(); ++i) { pl[i] = pi[i]; pr[i] = pi[i]; }",
           "for (size_t i = 0; i <= in.cells(); ++i) { pl[i] = pi[i]; pr[i] = pi[i]; }")


def _runtime(name):
    p = subprocess.run(["g++", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


ASAN, UBSAN = _runtime("libasan.so"), _runtime("libubsan.so")
pytestmark = pytest.mark.skipif(not (ASAN and UBSAN and shutil.which("g++")),
                                reason="g++ sanitizer runtimes not available")


def _build(src_text, out):
    d = os.path.dirname(out)
    os.makedirs(os.path.join(d, "oracle"), exist_ok=True)
    os.makedirs(os.path.join(d, "include"), exist_ok=True)
    shutil.copy(os.path.join(ROOT, "include", "alsvinn_hip.h"), os.path.join(d, "include"))
    src = os.path.join(d, "oracle", "oracle.cpp")
    with open(src, "w") as f:
        f.write(src_text)
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
                    "-shared", "-o", out, src], check=True, capture_output=True)


def _run_suites(lib):
    env = dict(os.environ, ORACLE_LIB=lib, OMP_NUM_THREADS="2", LD_PRELOAD=ASAN + ":" + UBSAN,
               ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    # -s: a sanitizer report goes to the process's own stderr, not into pytest's capture
    return subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-s", "-p", "no:cacheprovider"] + SUITES,
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)


@pytest.fixture(scope="module")
def builds():
    tmp = tempfile.mkdtemp(prefix="oracle_san_")
    try:
        src = open(os.path.join(ROOT, "oracle", "oracle.cpp")).read()
        assert src.count(PLANTED[0]) == 1
        clean = os.path.join(tmp, "clean", "liboracle.so")
        bad = os.path.join(tmp, "planted", "liboracle.so")
        _build(src, clean)
        _build(src.replace(*PLANTED), bad)
        yield clean, bad
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_oracle_clean_under_address_and_undefined_behaviour_sanitizers(builds):
    r = _run_suites(builds[0])
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-3000:]


def test_sanitized_build_catches_a_planted_overflow(builds):
    r = _run_suites(builds[1])
    assert r.returncode != 0
    assert "AddressSanitizer" in r.stdout + r.stderr
