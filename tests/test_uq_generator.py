"""Sample parameters of the hip platform's UQ layer (host C++ in
libalsvinn_hip.so; no GPU needed) against the C++ standard library's
std::mt19937_64 / uniform_real_distribution golden vectors and the
reference's STLMersenne fast-forward rules."""
import os

import numpy as np

from alsvinn_amd import generate_parameters

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "stl_mt19937_64_uniform.npz"))


def test_mt19937_64_fixture_standard_value(oracle):
    # C++ [rand.predef]: the 10000th output of a default mt19937_64
    assert int(oracle.mt19937_64(5489, 9999, 1)[0]) == 9981545732273789042
    np.testing.assert_array_equal(GOLD["raw"][:10], oracle.mt19937_64(5489, 0, 10))


def test_uniform_is_generate_canonical():
    """libstdc++ generate_canonical<double,53> on a 64-bit engine: one draw,
    RN(x) / 2^64."""
    u = GOLD["raw"].astype(np.float64) / 2.0 ** 64
    np.testing.assert_array_equal(u, GOLD["uniform"])


def test_kh_parameters_are_consecutive_draws():
    """One 40-long parameter: component c of sample s is draw s*40 + c
    (alsuq/src/generator/STLMersenne.cpp:38-85)."""
    a = generate_parameters([("a", 40, 0.0, 1.0)], 0, 25)["a"]
    np.testing.assert_array_equal(a.ravel(), GOLD["uniform"][:1000])


def test_sample_start_fast_forwards():
    """<sampleStart> skips to draw sampleStart*40 (alsuq/src/config/Setup.cpp:70-75)."""
    a = generate_parameters([("a", 40, 0.0, 1.0)], 7, 3)["a"]
    np.testing.assert_array_equal(a.ravel(), GOLD["uniform"][280:400])


def test_uniform_scaling():
    """Uniform::scale: x*(b-a)+a (alsuq/src/distribution/Uniform.cpp:27-34)."""
    a = generate_parameters([("a", 4, -2.0, 3.0)], 0, 2)["a"]
    np.testing.assert_array_equal(a.ravel(), GOLD["uniform"][:8] * (3.0 - -2.0) + -2.0)


def stl_mersenne_replay(raw, params, sample_start, nsamples):
    """The reference's shared-counter fast-forward, restated in Python over
    the raw engine outputs (generate_canonical = RN(raw)/2^64)."""
    counter = 0
    pos = 0
    out = []
    for s in range(sample_start, sample_start + nsamples):
        row = []
        for name, dim, lo, hi in sorted(params, key=lambda p: p[0]):
            for c in range(dim):
                if counter // dim < s:
                    add = s * dim - counter
                    pos += add
                    counter += add
                if counter % dim < c:
                    add = c - counter % dim
                    pos += add
                    counter += add
                counter += 1
                x = float(np.float64(raw[pos])) / 2.0 ** 64
                pos += 1
                row.append(x * (hi - lo) + lo)
        out.append(row)
    return np.array(out)


def test_several_parameters_share_one_engine():
    """Several parameters draw from the same engine in name order with the
    reference's counter rules (the generator map iterates sorted names)."""
    params = [("b", 3, 0.0, 1.0), ("a", 5, 1.0, 2.0), ("c", 2, -1.0, 1.0)]
    got = generate_parameters(params, 2, 4)
    flat = np.concatenate([got["a"], got["b"], got["c"]], axis=1)
    np.testing.assert_array_equal(flat, stl_mersenne_replay(GOLD["raw"], params, 2, 4))


def test_monte_carlo_rate():
    """MeanVarTest (test/library_tests/src/uq/generators/MeanVarTest.cpp:85-149):
    the Monte Carlo error of the mean and variance of U[0,1) draws (D = 40)
    decays with rate >= 0.49."""
    errs_m, errs_v, ms = [], [], []
    big = generate_parameters([("a", 40, 0.0, 1.0)], 0, 1 << 14)["a"]
    for k in range(6, 15):
        m = 1 << k
        x = big[:m]
        errs_m.append(np.abs(x.mean(axis=0) - 0.5).mean())
        errs_v.append(np.abs(x.var(axis=0) - 1.0 / 12.0).mean())
        ms.append(m)
    rate_m = -np.polyfit(np.log(ms), np.log(errs_m), 1)[0]
    rate_v = -np.polyfit(np.log(ms), np.log(errs_v), 1)[0]
    assert rate_m >= 0.49 and rate_v >= 0.49


def test_monte_carlo_rate_of_the_generator():
    """MeanVarTest (test/library_tests/src/uq/generators/MeanVarTest.cpp:85-149):
    40 uniform[0,1] components from a fresh generator for M = 2^4 .. 2^19
    samples; the errors of the sample mean and the (N-1) sample variance
    against 1/2 and 1/12 decay at rate >= 0.49 in M."""
    from alsvinn_amd.uq import generate_parameters
    logM, logEm, logEv = [], [], []
    for k in range(4, 20):
        M = 1 << k
        x = generate_parameters([("a", 40, 0.0, 1.0)], 0, M)["a"]
        em = np.sqrt(np.sum((x.mean(0) - 0.5) ** 2))
        ev = np.sqrt(np.sum((x.var(0, ddof=1) - 1.0 / 12.0) ** 2))
        logM.append(np.log(M))
        logEm.append(np.log(em))
        logEv.append(np.log(ev))
    rate_mean = -np.polyfit(logM, logEm, 1)[0]
    rate_var = -np.polyfit(logM, logEv, 1)[0]
    assert rate_mean >= 0.49 and rate_var >= 0.49, (rate_mean, rate_var)
