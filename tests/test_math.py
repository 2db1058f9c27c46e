"""hd_math.h (the transcendental functions both the kernels and the CPU
checker use) against the platform libm: within a few ulp on the domains the
ADI / sediment formulas use, and the special values C99 fixes."""
import math

import numpy as np
import pytest

import oracle


def ulps(a, b):
    if a == b:
        return 0.0
    return abs(a - b) / (np.spacing(abs(b)) or 5e-324)


@pytest.fixture(scope="module")
def L():
    return oracle.lib()


def test_log_exp_accuracy(L):
    rng = np.random.default_rng(11)
    xs = np.exp(rng.uniform(-700, 700, 20000))
    assert max(ulps(L.or_hd_log(x), math.log(x)) for x in xs) <= 1.0
    ys = rng.uniform(-700, 700, 20000)
    assert max(ulps(L.or_hd_exp(y), math.exp(y)) for y in ys) <= 1.0


def test_pow_accuracy(L):
    rng = np.random.default_rng(12)
    # bases/exponents of the formulas: depth^(-0.4), depth^(2/3,5/3), (1-FS)^4,
    # Pd^2.5, Tx^1.5, Tx^2.1, dm^1.5, Dxr^-0.3, (g(s-1))^(8/15) ...
    xs = np.exp(rng.uniform(-12, 6, 20000))
    ys = rng.choice([-0.4, 2 / 3, 5 / 3, 4.0, 2.5, 1.5, 2.1, -0.3, 8 / 15, 9 / 15, -1 / 15, 1 / 3, 3.0, 0.5], 20000)
    worst = max(ulps(L.or_hd_pow(x, y), math.pow(x, y)) for x, y in zip(xs, ys))
    assert worst <= 2.0


def test_pow_special_values(L):
    assert L.or_hd_pow(0.0, 2.1) == 0.0
    assert L.or_hd_pow(2.0, 0.0) == 1.0
    assert L.or_hd_pow(1.0, 123.4) == 1.0
    assert L.or_hd_pow(-2.0, 3.0) == -8.0
    assert L.or_hd_pow(3.0, 2.0) == 9.0
    assert math.isnan(L.or_hd_pow(-2.0, 0.5))
    assert L.or_hd_pow(0.0, -1.0) == math.inf


def test_trig_accuracy(L):
    rng = np.random.default_rng(13)
    xs = rng.uniform(-2e4, 2e4, 20000)
    assert max(abs(L.or_hd_cos(x) - math.cos(x)) for x in xs) < 2.5e-16
    assert max(abs(L.or_hd_sin(x) - math.sin(x)) for x in xs) < 2.5e-16
    assert L.or_hd_cos(0.0) == 1.0 and L.or_hd_sin(0.0) == 0.0
