"""Partition (sim/core.go:126-164) against the reference's own known-answer
tests, sim/partition_test.go:11-160, run in file order on shared arrays
(the Go tests share the package-level arrays, which Partition permutes)."""
import numpy as np

VERT, HOR = 0, 1

# sim/partition_test.go:11-16, 66-72, 128-134
PS_EVEN = [(1.0, 1.0), (0.9, 0.9), (0.8, 0.8), (0.7, 0.7)]
PS_ODD = [(0.0, 0.9), (0.5, -0.8), (1.7, 0.1), (0.7, -0.1), (-0.7, 0.1)]
PS_VAR = [(0.9, 0.0), (-0.8, 0.5), (0.1, 1.7), (-0.1, 0.7), (0.1, -0.7)]

# (array, orientation, middle, len(a), len(b)) in the order of the Go file
CASES = [
    ("even", VERT, 0.5, 0, 4),            # :18-28
    ("even", HOR, 0.5, 0, 4),             # :30-40
    ("even", VERT, 0.85, 2, 2),           # :42-52
    ("even", HOR, 0.85, 2, 2),            # :54-64
    ("odd", VERT, 0.100000000001, 4, 1),  # :74-84
    ("odd", HOR, 0.601, 3, 2),            # :86-96
    ("odd", VERT, -100, 0, 5),            # :98-108
    ("odd", HOR, 100, 5, 0),              # :110-120
    ("var", HOR, 0.100000000001, 4, 1),   # :136-139
    ("var", VERT, 0.601, 3, 2),           # :141-144
    ("var", HOR, -100, 0, 5),             # :146-149
    ("var", VERT, 100, 5, 0),             # :151-154
]


def test_partition_known_answers(oracle_lib):
    arrs = {
        "even": np.array(PS_EVEN, dtype=np.float64),
        "odd": np.array(PS_ODD, dtype=np.float64),
        "var": np.array(PS_VAR, dtype=np.float64),
    }
    for name, ori, mid, ea, eb in CASES:
        a = arrs[name]
        na, x, y = oracle_lib.partition(a[:, 0], a[:, 1], ori, mid)
        assert (na, len(a) - na) == (ea, eb), (name, ori, mid)
        arrs[name] = np.stack([x, y], axis=1)  # Go permutes in place


def test_partition_empty(oracle_lib):
    # sim/partition_test.go:156-160
    na, x, y = oracle_lib.partition(np.zeros(0), np.zeros(0), HOR, 0.85)
    assert na == 0 and len(x) == 0


def test_partition_is_a_partition(oracle_lib):
    # property: every element of a is <= middle, every element of b > middle
    rng = np.random.default_rng(7)
    for n in (2, 3, 9, 100, 1000):
        for ori in (VERT, HOR):
            x = rng.random(n)
            y = rng.random(n)
            mid = 0.5
            na, px, py = oracle_lib.partition(x, y, ori, mid)
            v = py if ori == VERT else px
            assert (v[:na] <= mid).all() and (v[na:] > mid).all()
            assert sorted(zip(px, py)) == sorted(zip(x, y))


def test_partition_single_element_asymmetry(oracle_lib):
    # sim/core.go:142-144 vs :159-161: the i==j fix-up sits inside the loop
    # for Vertical, so a one-element slice is never moved to `a` there.
    na, _, _ = oracle_lib.partition(np.array([0.1]), np.array([0.1]), VERT, 0.5)
    assert na == 0
    na, _, _ = oracle_lib.partition(np.array([0.1]), np.array([0.1]), HOR, 0.5)
    assert na == 1
