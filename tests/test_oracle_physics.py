"""Density, force, integrator and totals of the oracle (sim/sph.go) pinned
by closed forms, invariances and special cases."""
import math

import numpy as np
import pytest

from _np_ref import lattice_shells, square_lattice
from sphugo_amd import ics

MON, WEN = 0, 1


def _sim(O, x, y, vx=None, vy=None, e=None, k=32, kernel=MON, forces=True, **kw):
    n = len(x)
    kw.setdefault("gamma", 5 / 3)
    kw.setdefault("mass", 1.0 / n)
    kw.setdefault("hor", (0, 1))
    kw.setdefault("vert", (0, 1))
    p = O.default_params(k=k, kernel=kernel, **kw)
    vx = np.zeros(n) if vx is None else vx
    vy = np.zeros(n) if vy is None else vy
    e = np.ones(n) if e is None else e
    s = O.OracleSim(p, x, y, vx, vy, e)
    if forces:
        s.set("vpx", vx); s.set("vpy", vy); s.set("epred", e)
        s.calculate_forces()
    return s


@pytest.mark.parametrize("kernel", [MON, WEN])
@pytest.mark.parametrize("k", [32, 64])
def test_lattice_density_closed_form(oracle_lib, kernel, k):
    n = 40
    x, y = square_lattice(n)
    s = _sim(oracle_lib, x, y, k=k, kernel=kernel)
    shells, r2k = lattice_shells(k)
    F = lambda q: oracle_lib.kernel_F(kernel, q)
    fp, _ = oracle_lib.kernel_prefactors(kernel)
    # m = 1/N = a^2, h = a*sqrt(r2k): rho = Fpref/r2k * sum_shells mult*F(r/h)
    acc = sum(cnt * F(math.sqrt(rr / r2k)) for rr, cnt in shells if rr < r2k)
    expect = fp * acc / r2k
    rho = s.get("rho")
    assert np.allclose(rho, expect, rtol=1e-12, atol=0)
    assert np.allclose(s.get("h"), math.sqrt(r2k) / n, rtol=1e-12)


def test_lattice_at_rest_feels_only_gravity(oracle_lib):
    x, y = square_lattice(32)
    g = (0.05, -0.1)
    s = _sim(oracle_lib, x, y, accel=g)
    ax, ay, ed = s.get("ax"), s.get("ay"), s.get("edot")
    # pressure terms cancel to rounding; dot = 0 so edot is exactly 0
    scale = 1e-9 * (1.0 / s.get("rho").mean()) * 1e3
    assert np.abs(ax - g[0]).max() < scale
    assert np.abs(ay - g[1]).max() < scale
    assert (ed == 0).all()


def test_galilean_invariance(oracle_lib):
    st = ics.uniform(1500, seed=21)
    rng = np.random.default_rng(2)
    vx, vy = 0.1 * rng.standard_normal(1500), 0.1 * rng.standard_normal(1500)
    a = _sim(oracle_lib, st["x"], st["y"], vx, vy, visc_alpha=1.0, visc_beta=2.0)
    b = _sim(oracle_lib, st["x"], st["y"], vx + 0.375, vy - 0.25, visc_alpha=1.0, visc_beta=2.0)
    for f in ("ax", "ay", "edot"):
        A, B = a.get(f), b.get(f)
        assert np.abs(A - B).max() <= 1e-10 * np.abs(A).max()


def test_rotation_equivariance(oracle_lib):
    st = ics.uniform(1500, seed=22)
    rng = np.random.default_rng(3)
    vx, vy = 0.1 * rng.standard_normal(1500), 0.1 * rng.standard_normal(1500)
    a = _sim(oracle_lib, st["x"], st["y"], vx, vy)
    # rotate by 90 degrees about the box centre: (x, y) -> (y, 1 - x)
    xr, yr = st["y"].copy(), ics.wrap_unit(1.0 - st["x"])
    b = _sim(oracle_lib, xr, yr, vy, -vx)
    assert np.allclose(b.get("rho"), a.get("rho"), rtol=1e-12)
    sc = np.abs(a.get("ax")).max()
    assert np.abs(b.get("ax") - a.get("ay")).max() < 1e-9 * sc
    assert np.abs(b.get("ay") + a.get("ax")).max() < 1e-9 * sc
    assert np.abs(b.get("edot") - a.get("edot")).max() < 1e-9 * np.abs(a.get("edot")).max()


def test_pressure_gradient_acceleration(oracle_lib):
    # e = e0 (1 + eps sin 2 pi x) on a lattice: P = (gamma-1) rho e, so the
    # continuum acceleration is -(gamma-1) e0 eps 2 pi cos(2 pi x) along x.
    n = 64
    x, y = square_lattice(n)
    gam, e0, eps = 5 / 3, 1.0, 0.01
    e = e0 * (1 + eps * np.sin(2 * np.pi * x))
    s = _sim(oracle_lib, x, y, e=e, gamma=gam)
    ax, ay = s.get("ax"), s.get("ay")
    cont = -(gam - 1) * e0 * eps * 2 * np.pi * np.cos(2 * np.pi * x)
    corr = np.corrcoef(ax, cont)[0, 1]
    ratio = np.dot(ax, cont) / np.dot(cont, cont)
    assert corr > 0.999
    assert 0.6 < ratio < 1.4
    assert np.abs(ay).max() < 1e-6 * np.abs(ax).max()


def test_compression_heats(oracle_lib):
    # converging flow v = -beta (x - 1/2): edot > 0 everywhere away from the
    # periodic seam, and edot is linear in beta (sim/sph.go:393,400)
    n = 32
    x, y = square_lattice(n)
    inner = (np.abs(x - 0.5) < 0.3) & (np.abs(y - 0.5) < 0.3)
    r = []
    for beta in (1e-3, 2e-3):
        vx, vy = -beta * (x - 0.5), -beta * (y - 0.5)
        s = _sim(oracle_lib, x, y, vx, vy, visc_alpha=0.0, visc_beta=0.0)
        r.append(s.get("edot")[inner])
    assert (r[0] > 0).all()
    assert np.allclose(r[1], 2 * r[0], rtol=1e-9)


def test_uniformly_accelerated_lattice_is_a_parabola(oracle_lib):
    n = 24
    x0, y0 = square_lattice(n)
    g = (0.3, -0.2)
    dth = 0.01
    s = _sim(oracle_lib, x0, y0, accel=g, dt_half=dth, forces=False)
    nsteps = 10
    s.step(nsteps)
    t = nsteps * 2 * dth
    ex = (x0 + 0.5 * g[0] * t * t) % 1.0
    ey = (y0 + 0.5 * g[1] * t * t) % 1.0
    dx = s.get("x") - ex
    dy = s.get("y") - ey
    dx -= np.rint(dx)
    dy -= np.rint(dy)
    assert np.abs(dx).max() < 1e-12 and np.abs(dy).max() < 1e-12
    assert np.allclose(s.get("vx"), g[0] * t, rtol=1e-12)
    assert np.allclose(s.get("vy"), g[1] * t, rtol=1e-12)
    assert (s.get("e") == 1.0).all()


def test_wrap_continue_quirk(oracle_lib):
    # sim/sph.go:147-167: after an X wrap, `continue` skips the Y wrap
    n = 16
    x0, y0 = square_lattice(n)
    v = -0.05
    s = _sim(oracle_lib, x0, y0, np.full(n * n, v), np.full(n * n, v),
             dt_half=0.5, forces=False)
    s.step(1)
    x, y = s.get("x"), s.get("y")
    both = (x0 + v < 0) & (y0 + v < 0)
    only_y = (x0 + v >= 0) & (y0 + v < 0)
    assert both.any() and only_y.any()
    assert (y[both] < 0).all() and (x[both] >= 0).all()
    assert (y[only_y] >= 0).all()
    assert (x >= 0).all()


def test_reflection_clamps_and_flips(oracle_lib):
    # sim/sph.go:170-193
    n = 16
    x0, y0 = square_lattice(n)
    big = oracle_lib.DBL_MAX  # open boundaries (sim/config-parser.go:139-140)
    # e = 0 gives c = 0: no pressure, no viscosity, so every particle drifts
    # by -0.05 and those ending left of the wall are reflected
    s = _sim(oracle_lib, x0, y0, np.full(n * n, -0.1), np.zeros(n * n), np.zeros(n * n),
             dt_half=0.25, refl=(0.2, 1e300, -1e300, 1e300), hor=(-big, big),
             vert=(-big, big), forces=False)
    s.step(1)
    x, vx = s.get("x"), s.get("vx")
    hit = x0 - 0.05 < 0.2
    # x - (x - L) lands on L up to one rounding; the velocity flipped sign
    assert np.abs(x[hit] - 0.2).max() <= 1e-16
    assert (vx[hit] == 0.1).all() and (vx[~hit] == -0.1).all()
    assert np.allclose(x[~hit], x0[~hit] - 0.05, rtol=0, atol=1e-15)


def test_totals(oracle_lib):
    st = ics.blast(3000, seed=8)
    s = _sim(oracle_lib, st["x"], st["y"], e=st["e"], use_cfl=1, cfl=0.3, forces=False)
    s.step(2)
    red = s.reductions()
    e = s.get("e")
    assert abs(red[0] - math.fsum(e)) <= 1e-13 * abs(red[0])
    assert abs(s.total_energy() - math.fsum(e)) <= 1e-13 * abs(red[0])
    rho = s.get("rho")
    assert abs(s.total_density() - math.fsum(rho)) <= 1e-13 * math.fsum(rho)
    # TotalMomentum quirk (sim/sph.go:457-463): speed of the last particle
    # in tree order; it is the speed of *some* particle
    vx, vy = s.get("vx"), s.get("vy")
    sp = np.sqrt(vx * vx + vy * vy)
    assert (sp == s.total_momentum()).any()


def test_cfl_dt_definition(oracle_lib):
    st = ics.uniform(800, seed=1)
    s = _sim(oracle_lib, st["x"], st["y"], cfl=0.3)
    h, c = s.get("h"), s.get("c")
    v = np.hypot(s.get("vx"), s.get("vy"))
    assert s.cfl_dt_half() == 0.5 * (0.3 * np.min((0.5 * h) / (c + v)))


def _moving_particle(O, u, alpha, beta, i=137):
    """Square lattice at rest except particle i moving with vx = u; k = 20
    fills whole shells (r^2 = 1, 2, 4, 5 spacings^2)."""
    x, y = square_lattice(16)
    vx = np.zeros(len(x))
    vx[i] = u
    s = _sim(O, x, y, vx=vx, k=20, visc_alpha=alpha, visc_beta=beta)
    return s.get("ax")[i], s.get("ay")[i], s.get("edot")[i], s


@pytest.mark.parametrize("kernel", [MON])
def test_viscosity_structure(oracle_lib, kernel):
    """Viscosity (sim/sph.go:370-390) acts only on approaching pairs (dot < 0),
    is linear in the relative velocity through alpha and quadratic through
    beta, and the pressure terms and EDot cancel over the symmetric lattice."""
    u = 0.3
    # no viscosity: a lattice point feels no net pressure force
    ax0, ay0, ed0, _ = _moving_particle(oracle_lib, u, 0.0, 0.0)
    assert abs(ax0) < 1e-10 and abs(ay0) < 1e-10
    # EDot = A m sum dot DF: dot = -u (x_j - x_i) is odd over the lattice shells
    assert abs(ed0) < 1e-12
    # alpha only: linear in u, decelerating, mirror-odd
    a1, ay1, ed1, _ = _moving_particle(oracle_lib, u, 1.0, 0.0)
    a2, _, _, _ = _moving_particle(oracle_lib, 2 * u, 1.0, 0.0)
    am, _, _, _ = _moving_particle(oracle_lib, -u, 1.0, 0.0)
    assert a1 < 0 and abs(a1) > 1e3 * abs(ax0)
    assert abs(a2 - 2 * a1) <= 1e-12 * abs(a1)
    assert abs(am + a1) <= 1e-12 * abs(a1)
    assert abs(ay1) <= 1e-12 * abs(a1)
    # the viscous EDot is not in the reference's EDot (quirk Q6): still zero
    assert abs(ed1) < 1e-12
    # beta only: quadratic in u, still decelerating
    b1, _, _, _ = _moving_particle(oracle_lib, u, 0.0, 1.0)
    b2, _, _, _ = _moving_particle(oracle_lib, 2 * u, 0.0, 1.0)
    assert b1 < 0 and abs(b2 - 4 * b1) <= 1e-12 * abs(b2)
    # both: the sum of the two parts (the bracket is linear in Pi)
    ab, _, _, _ = _moving_particle(oracle_lib, u, 1.0, 1.0)
    assert abs(ab - (a1 + b1)) <= 1e-12 * abs(ab)
    # alpha scales the linear part exactly
    a3, _, _, _ = _moving_particle(oracle_lib, u, 3.0, 0.0)
    assert abs(a3 - 3 * a1) <= 1e-12 * abs(a3)
