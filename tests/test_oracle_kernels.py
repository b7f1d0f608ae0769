"""SPH kernels (sim/sph.go:258-304) against closed forms."""
import math

import numpy as np
import pytest
from scipy import integrate

MON, WEN = 0, 1


@pytest.mark.parametrize("kern", [MON, WEN])
def test_normalisation(oracle_lib, kern):
    # W(r,h) = Fpref/h^2 F(r/h) integrates to 1 over the plane:
    # Fpref * 2 pi * int_0^1 q F(q) dq = 1
    F = lambda q: oracle_lib.kernel_F(kern, q)
    fp, _ = oracle_lib.kernel_prefactors(kern)
    pts = [0.5] if kern == MON else None
    val, _ = integrate.quad(lambda q: q * F(q), 0, 1, points=pts, epsabs=1e-14, epsrel=1e-14)
    assert abs(fp * 2 * math.pi * val - 1) < 1e-12


@pytest.mark.parametrize("kern", [MON, WEN])
def test_derivative_identity(oracle_lib, kern):
    # DFpref * DF(q) = Fpref * F'(q)
    fp, dfp = oracle_lib.kernel_prefactors(kern)
    for q in np.linspace(0.01, 0.99, 57):
        if abs(q - 0.5) < 1e-3:
            continue
        h = 1e-6
        num = (oracle_lib.kernel_F(kern, q + h) - oracle_lib.kernel_F(kern, q - h)) / (2 * h)
        assert abs(dfp * oracle_lib.kernel_DF(kern, q) - fp * num) < 1e-7


@pytest.mark.parametrize("kern", [MON, WEN])
def test_vanish_at_support(oracle_lib, kern):
    assert oracle_lib.kernel_F(kern, 1.0) == 0.0
    assert oracle_lib.kernel_DF(kern, 1.0) == 0.0


def test_monaghan_continuity(oracle_lib):
    a = oracle_lib.kernel_F(MON, 0.5 - 1e-12)
    b = oracle_lib.kernel_F(MON, 0.5)
    assert abs(a - b) < 1e-11
    a = oracle_lib.kernel_DF(MON, 0.5 - 1e-12)
    b = oracle_lib.kernel_DF(MON, 0.5)
    assert abs(a - b) < 1e-11


def test_prefactors_are_go_constant_folds(oracle_lib):
    # Go folds 6*40/(Pi*7), 4*7/(Pi*4), 8*7/(Pi*4) exactly and rounds once;
    # compare with the correctly rounded quotient at 300 bits.
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 300
    pi = mpmath.mpf("3.14159265358979323846264338327950288419716939937510582097494459")
    fp, dfp = oracle_lib.kernel_prefactors(MON)
    assert fp == float(mpmath.mpf(240) / (pi * 7)) == dfp
    fp, dfp = oracle_lib.kernel_prefactors(WEN)
    assert fp == float(mpmath.mpf(28) / (pi * 4))
    assert dfp == float(mpmath.mpf(56) / (pi * 4))
