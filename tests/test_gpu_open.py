"""Open boundaries (no periodic images on an axis whose bounds are
+-MaxFloat64, sim/nearest-neighbour.go:39-55) and the reference's own
example configurations, against the oracle (GPU).

  speed test        MakeConfig defaults (both axes open, gamma 1.66666, mass 1,
                    no walls, alpha 0.75, beta 1.5, eta2 0.01), a uniform
                    rectangle spawn over [0,1]^2 with e = 0.01, DeltaTHalf
                    0.02, Acceleration (0, 0.2): examples/speed-test/
                    speed-test.go:22-28, sim/config-parser.go:131-149,51-80
  example config 2  exampleConfigSource2 (sim/config-parser.go:926-973): open
                    tube with reflection walls Left 0.2, Up 0.25, Down 0.5,
                    Wendland, gamma 4.666, mass 1e5, g (0, 0.05), DeltaTHalf
                    0.00424, rectangles of 4000 and 700 particles
  example config 1  exampleConfigSource1 (sim/config-parser.go:869-924): a
                    periodic box 0.6 x 200 (Horizontal 0.2 0.8, Vertical
                    -100 100), wall Down 0.99, Wendland, gamma 4.666, mass
                    1e6, g (0, 0.55), DeltaTHalf 0.00324, rectangles of 260
                    and 700 particles

The spawners' Go math/rand streams cannot be reproduced here, so the
rectangles are filled from a seeded NumPy generator (positions uniform in the
rectangle, as UniformRectSpawner.Spawn draws them).

Bars: bitwise form, every field of every particle bit-exact after 10 steps;
plain FP64, neighbour sets and h bit-exact, density and sound speed 1e-12
relative, acceleration and EDot 1e-12 of their terms per particle, on the
first pass and on a pass after steps (the bracketed walk); FP32C, neighbour
sets and h bit-exact and density 1e-5 relative.
"""
import os

import numpy as np
import pytest

import oracle as O
from sphugo_amd import _lib as L

from test_gpu_parity import _check_forces, _check_knn, _rel, state_pass_vs_oracle

pytestmark = pytest.mark.gpu

OPEN = (-L.DBL_MAX, L.DBL_MAX)
THREADS = min(16, os.cpu_count() or 1)


def _rects(rects, seed):
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    for n, (x0, y0), (x1, y1) in rects:
        xs.append(x0 + rng.random(n) * (x1 - x0))
        ys.append(y0 + rng.random(n) * (y1 - y0))
    x, y = np.concatenate(xs), np.concatenate(ys)
    return x, y


def speed_test(n, seed=1):
    x, y = _rects([(n, (0.0, 0.0), (1.0, 1.0))], seed)
    st = dict(x=x, y=y, vx=np.zeros(n), vy=np.zeros(n), e=np.full(n, 0.01))
    common = dict(k=32, kernel=0, gamma=1.66666, mass=1.0, dt_half=0.02, accel=(0.0, 0.2),
                  hor=OPEN, vert=OPEN, refl=(-L.DBL_MAX, L.DBL_MAX, -L.DBL_MAX, L.DBL_MAX),
                  visc_alpha=0.75, visc_beta=1.5, visc_eta2=0.01)
    return st, common


def example2(seed=2):
    x, y = _rects([(4000, (0.3, 0.3), (0.7, 0.4)), (700, (0.3, 0.3), (0.7, 0.5))], seed)
    n = len(x)
    st = dict(x=x, y=y, vx=np.zeros(n), vy=np.zeros(n), e=np.full(n, 0.01))
    common = dict(k=32, kernel=1, gamma=4.666, mass=100000.0, dt_half=0.00424,
                  accel=(0.0, 0.05), hor=OPEN, vert=OPEN,
                  refl=(0.2, L.DBL_MAX, 0.25, 0.5), visc_alpha=0.75, visc_beta=1.5,
                  visc_eta2=0.01)
    return st, common


def example1(seed=3):
    x, y = _rects([(260, (0.6, 0.2), (0.79, 0.3)), (700, (0.27, 0.3), (0.4, 0.9))], seed)
    n = len(x)
    st = dict(x=x, y=y, vx=np.zeros(n), vy=np.zeros(n), e=np.full(n, 0.01))
    common = dict(k=32, kernel=1, gamma=4.666, mass=1000000.0, dt_half=0.00324,
                  accel=(0.0, 0.55), hor=(0.2, 0.8), vert=(-100.0, 100.0),
                  refl=(-L.DBL_MAX, L.DBL_MAX, -L.DBL_MAX, 0.99), visc_alpha=0.75,
                  visc_beta=1.5, visc_eta2=0.01)
    return st, common


def _pair(st, common, precision=L.PREC_FP64, flags=0, forces=True):
    g = L.Sim(L.default_params(precision=precision, flags=flags, **common), st["x"], st["y"],
              st["vx"], st["vy"], st["e"])
    o = O.OracleSim(O.default_params(knn_mode=O.KNN_EXACT, use_cfl=0, **common), st["x"],
                    st["y"], st["vx"], st["vy"], st["e"], threads=THREADS)
    if forces:
        for s in (g, o):
            s.set("vpx", st["vx"])
            s.set("vpy", st["vy"])
            s.set("epred", st["e"])
            s.calculate_forces()
    return g, o


CONFIGS = {"speed_test": lambda: speed_test(3000), "example2": example2, "example1": example1}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_steps_bitwise(name):
    st, common = CONFIGS[name]()
    g, o = _pair(st, common, flags=L.FLAG_BITWISE, forces=False)
    g.step(10)
    o.step(10)
    for f in L.FIELDS:
        assert (g.get(f) == o.get(f)).all(), f
    # sum of e: the device reduces in a fixed tree order, the oracle serially
    assert abs(g.total_energy() - o.total_energy()) <= 1e-14 * abs(o.total_energy())


@pytest.mark.parametrize("name", list(CONFIGS))
def test_first_pass_and_walk_fp64(name):
    st, common = CONFIGS[name]()
    g, o = _pair(st, common)
    _check_knn(g, o)
    _check_forces(g, o, accel=common["accel"])
    g.step(3)
    o2, _ = state_pass_vs_oracle(g, O.default_params(knn_mode=O.KNN_EXACT, use_cfl=0, **common),
                                 THREADS)
    _check_knn(g, o2)
    _check_forces(g, o2, accel=common["accel"])


@pytest.mark.parametrize("name", list(CONFIGS))
def test_fp32c(name):
    st, common = CONFIGS[name]()
    g, o = _pair(st, common, precision=L.PREC_FP32C)
    _check_knn(g, o)
    assert _rel(g.get("rho"), o.get("rho")).max() <= 1e-5
    g.step(3)
    o2, _ = state_pass_vs_oracle(g, O.default_params(knn_mode=O.KNN_EXACT, use_cfl=0, **common),
                                 THREADS)
    _check_knn(g, o2)
    assert _rel(g.get("rho"), o2.get("rho")).max() <= 1e-5


def test_speed_test_100k_walk():
    """The reference's timing harness at its own size (100,000 particles,
    both axes open): 5 steps, then one pass from the copied state, the
    bracketed walk (and the exact search for the queries whose h moved out
    of their bracket under the large fixed step) against the oracle."""
    st, common = speed_test(100000, seed=5)
    g = L.Sim(L.default_params(**common), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    g.step(5)
    o, nfb = state_pass_vs_oracle(g, O.default_params(knn_mode=O.KNN_EXACT, use_cfl=0, **common),
                                  THREADS)
    assert nfb < 0.9 * len(st["x"]), nfb
    _check_knn(g, o)
    _check_forces(g, o, accel=common["accel"])


def test_open_box_follows_the_particles():
    """An open gas that expands: the grid box is re-derived each pass, and
    the neighbour sets stay exact as the cloud grows."""
    n = 4000
    rng = np.random.default_rng(8)
    x = 0.5 + 0.01 * rng.standard_normal(n)
    y = 0.5 + 0.01 * rng.standard_normal(n)
    st = dict(x=x, y=y, vx=(x - 0.5) * 20, vy=(y - 0.5) * 20, e=np.full(n, 0.5))
    common = dict(k=32, kernel=0, gamma=5 / 3, mass=1.0 / n, dt_half=0.002, accel=(0.0, 0.0),
                  hor=OPEN, vert=(0.0, 1.0), refl=(-L.DBL_MAX, L.DBL_MAX, -L.DBL_MAX, L.DBL_MAX),
                  visc_alpha=1.0, visc_beta=2.0, visc_eta2=0.01)
    g, o = _pair(st, common, flags=L.FLAG_BITWISE, forces=False)
    g.step(20)
    o.step(20)
    for f in L.FIELDS:
        assert (g.get(f) == o.get(f)).all(), f
    assert np.ptp(g.get("x")) > 2 * np.ptp(x)
