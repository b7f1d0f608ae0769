"""The oracle's Go math.Hypot restatement (sim/linear-algebra.go:64-71 calls
it for the tree's pruning distances) against what the mathematics fixes:
exact Pythagorean triples, symmetry and sign, exact scaling by powers of
two, no overflow or underflow where p^2 + q^2 would leave the double range,
Go's special cases, and a few-ulp accuracy against an extended-precision
reference.  Go's own rounding (an assembly form on some architectures) is
not pinned: no Go toolchain is available to produce vectors."""
import math

import numpy as np


def test_exact_triples_and_zero(oracle_lib):
    h = oracle_lib.hypot
    assert h(3.0, 4.0) == 5.0
    assert h(4.0, 3.0) == 5.0
    assert h(0.0, 0.0) == 0.0
    assert h(0.0, -7.5) == 7.5
    assert h(-2.25, 0.0) == 2.25


def test_symmetry_sign_and_power_of_two_scaling(oracle_lib):
    h = oracle_lib.hypot
    rng = np.random.default_rng(11)
    for p, q in rng.standard_normal((200, 2)):
        v = h(p, q)
        assert v == h(q, p) == h(-p, q) == h(p, -q) == h(-q, -p)
        for e in (-600, -40, 7, 300):
            assert h(math.ldexp(p, e), math.ldexp(q, e)) == math.ldexp(v, e)


def test_extremes_do_not_overflow_or_underflow(oracle_lib):
    h = oracle_lib.hypot
    big = 1e300
    assert math.isfinite(h(big, big)) and abs(h(big, big) / (big * math.sqrt(2)) - 1) < 1e-15
    tiny = 1e-300
    assert h(tiny, tiny) > 0 and abs(h(tiny, tiny) / (tiny * math.sqrt(2)) - 1) < 1e-15


def test_special_cases(oracle_lib):
    h = oracle_lib.hypot
    inf, nan = math.inf, math.nan
    assert h(inf, 1.0) == inf and h(-inf, 1.0) == inf and h(1.0, -inf) == inf
    assert h(inf, nan) == inf and h(nan, -inf) == inf   # Go: Inf wins over NaN
    assert math.isnan(h(nan, 1.0)) and math.isnan(h(2.0, nan))


def test_accuracy_against_extended_precision(oracle_lib):
    h = oracle_lib.hypot
    rng = np.random.default_rng(12)
    pq = rng.standard_normal((2000, 2)) * np.exp(rng.uniform(-20, 20, (2000, 1)))
    for p, q in pq:
        ref = np.sqrt(np.longdouble(p) ** 2 + np.longdouble(q) ** 2)
        v = h(p, q)
        assert abs(np.longdouble(v) - ref) <= 2 * np.spacing(np.float64(ref)), (p, q)
