"""Pins of the project's pow and atan (unnatural_planets_amd/csrc/upf_math.hpp), which
the oracle and the device share so that the elevation modes and the material stack give
the same bits on both sides.  Checked here against Python's math module (float64,
correctly rounded to within an ulp of float64): every float32 result within 0.5 ulp
(float32) of the exact value, i.e. correctly rounded except at exact rounding ties, over
the arguments the library meets (bases in [0, 1] with the reference's exponents 0.3,
0.5, 0.7, 0.83, 0.85, 1.24, 1.6, 1.7, 2, 3, 4, 5), a wide range of bases and exponents,
negative bases with integer exponents, and atan over the whole float line."""
import math

import numpy as np

EXPONENTS = np.array([0.3, 0.5, 0.7, 0.83, 0.85, 1.24, 1.6, 1.7, 2.0, 3.0, 4.0, 5.0], dtype=np.float32)


def ulp_err(got, exact):
    sp = np.spacing(np.abs(exact.astype(np.float32))).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / sp


def test_pow_correctly_rounded_on_the_reference_exponents(oracle):
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, 200000).astype(np.float32)
    e = rng.choice(EXPONENTS, len(x))
    got = oracle.powf(x, e)
    exact = np.array([math.pow(float(a), float(b)) for a, b in zip(x, e)])
    assert ulp_err(got, exact).max() <= 0.5 + 1e-9
    assert np.mean(got == exact.astype(np.float32)) > 0.9999


def test_pow_wide_range(oracle):
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-30, 30, 100000)).astype(np.float32)
    e = rng.uniform(-2.5, 2.5, len(x)).astype(np.float32)
    got = oracle.powf(x, e)
    exact = np.array([math.pow(float(a), float(b)) for a, b in zip(x, e)])
    ok = np.abs(exact) < 3e38  # finite in float32
    assert ulp_err(got[ok], exact[ok]).max() <= 0.5 + 1e-9


def test_pow_special_cases(oracle):
    x = np.array([-2, -0.5, -1.5, 0, 0, 1, 1, 0.25, -3], dtype=np.float32)
    e = np.array([3, 4, 5, 1.7, 0, 0.3, -2, 0.5, 2], dtype=np.float32)
    np.testing.assert_array_equal(oracle.powf(x, e), np.array([-8, 0.0625, -7.59375, 0, 1, 1, 1, 0.5, 9], np.float32))
    r = oracle.powf(np.array([-2.0, 0.0], np.float32), np.array([0.5, -1.0], np.float32))
    assert np.isnan(r[0]) and np.isinf(r[1])


def test_atan_correctly_rounded(oracle):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-2, 2, 100000), np.tan(rng.uniform(-1.5707, 1.5707, 100000)),
                        np.array([0.0, 1.0, -1.0, 1e30, -1e30, np.inf, -np.inf, 0.125, 0.0625])]).astype(np.float32)
    got = oracle.atanf(x)
    exact = np.arctan(x.astype(np.float64))
    assert ulp_err(got, exact).max() <= 0.5 + 1e-9
    assert np.mean(got == exact.astype(np.float32)) > 0.9999
