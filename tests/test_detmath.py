"""Pins of the deterministic libm (include/openmc_amd/detmath.h), no GPU.

The histories consume log/sin/cos; the reference takes them from glibc.  The
library here must (a) be correctly rounded on the inputs the transport loop
produces, checked against 50-digit decimal arithmetic, and (b) differ from
glibc only where glibc itself misrounds.  The rate of those differences is
what separates this build's histories from the reference binary's.
"""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

getcontext().prec = 50


@pytest.fixture(scope="module")
def det():
    from oracle.oracle import load_library
    return load_library(glibc_math=False)


def _dsin(x):
    x = Decimal(x)
    s, term, k = Decimal(0), x, 1
    while abs(term) > Decimal(10) ** -60:
        s += term
        term = -term * x * x / ((k + 1) * (k + 2))
        k += 2
    return s


def _dcos(x):
    x = Decimal(x)
    s, term, k = Decimal(0), Decimal(1), 0
    while abs(term) > Decimal(10) ** -60:
        s += term
        term = -term * x * x / ((k + 1) * (k + 2))
        k += 2
    return s


def _correctly_rounded(value, exact):
    """value is the double nearest to exact (ties impossible at this precision)."""
    v = Decimal(value)
    up = Decimal(math.nextafter(value, math.inf))
    dn = Decimal(math.nextafter(value, -math.inf))
    err = abs(v - exact)
    return err <= abs(up - exact) and err <= abs(dn - exact)


def _samples(rng, n):
    return {
        "log": np.concatenate([rng.random(n), np.exp(rng.uniform(-700, 700, n)),
                               1.0 + (rng.random(n) - 0.5) * 1e-9, np.ldexp(rng.random(n), -1060)]),
        "sin": np.concatenate([rng.random(n) * 2 * np.pi, rng.random(n) * np.pi / 2, rng.random(n) * 1e-6]),
    }


def test_log_correctly_rounded(det):
    rng = np.random.default_rng(1)
    xs = _samples(rng, 600)["log"]
    for x in xs:
        x = float(x)
        assert _correctly_rounded(det.oracle_log(x), Decimal(x).ln()), x


def test_sin_cos_correctly_rounded(det):
    rng = np.random.default_rng(2)
    xs = _samples(rng, 500)["sin"]
    for x in xs:
        x = float(x)
        assert _correctly_rounded(det.oracle_sin(x), _dsin(x)), x
        assert _correctly_rounded(det.oracle_cos(x), _dcos(x)), x


def test_special_values(det):
    assert det.oracle_log(1.0) == 0.0
    assert det.oracle_log(0.0) == -math.inf
    assert math.isnan(det.oracle_log(-1.0))
    assert det.oracle_log(math.inf) == math.inf
    assert det.oracle_log(2.0) == math.log(2.0)
    assert det.oracle_sin(0.0) == 0.0 and det.oracle_cos(0.0) == 1.0
    assert math.isnan(det.oracle_sin(math.inf))


def test_differences_from_glibc_are_glibc_misroundings(det):
    """Where this library and glibc disagree, this library is the closer one."""
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.random(200000), np.exp(rng.uniform(-30, 30, 200000))])
    diffs = 0
    checked = 0
    for x in xs:
        x = float(x)
        a = det.oracle_log(x)
        b = math.log(x)
        if a != b:
            diffs += 1
            if checked < 60:
                exact = Decimal(x).ln()
                assert abs(Decimal(a) - exact) < abs(Decimal(b) - exact)
                checked += 1
    # glibc's log misrounds rarely; record the rate so a regression shows up
    assert diffs / xs.size < 2e-3
