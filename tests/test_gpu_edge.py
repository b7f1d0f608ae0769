"""Edge cases of the HIP path against the oracle (GPU): gravity and
reflection walls, the Wendland kernel over several steps, anisotropic and
shifted boxes, exact distance ties on lattices, the smallest k and N, a
batch whose simulations have different boxes, FP32C neighbourhoods wider
than half a box, a pair of particles closer than the FP32C force record's
fixed-point step, and a grid mostly empty between two patches (its cell
ranges over long empty stretches filled by the queued path)."""
import numpy as np
import pytest

import oracle as O
from sphugo_amd import _lib as L
from sphugo_amd import ics

from test_gpu_parity import _check_forces, _check_knn

pytestmark = pytest.mark.gpu


def _pair(st, k, hor=(0, 1), vert=(0, 1), precision=L.PREC_FP64, flags=0, kernel=0,
          forces=True, **kw):
    kw.setdefault("gamma", 5 / 3)
    kw.setdefault("visc_alpha", 1.0)
    kw.setdefault("visc_beta", 2.0)
    common = dict(k=k, kernel=kernel, mass=st["mass"], hor=hor, vert=vert, **kw)
    g = L.Sim(L.default_params(precision=precision, flags=flags, **common), st["x"], st["y"],
              st["vx"], st["vy"], st["e"])
    use_cfl = 1 if common.get("dt_half", 0.001) <= 0 else 0
    o = O.OracleSim(O.default_params(knn_mode=O.KNN_EXACT, use_cfl=use_cfl, **common),
                    st["x"], st["y"], st["vx"], st["vy"], st["e"])
    if forces:
        for s in (g, o):
            s.set("vpx", st["vx"])
            s.set("vpy", st["vy"])
            s.set("epred", st["e"])
            s.calculate_forces()
    return g, o


def _lattice(n, jitter=0.0):
    gl = (np.arange(n) + 0.5) / n
    x, y = [a.ravel().copy() for a in np.meshgrid(gl, gl, indexing="ij")]
    return {"x": x, "y": y, "vx": np.zeros(n * n), "vy": np.zeros(n * n), "e": np.ones(n * n),
            "mass": 1 / (n * n)}


def test_gravity_and_walls_bitwise():
    st = ics.blast(3000, seed=21)
    g, o = _pair(st, 32, flags=L.FLAG_BITWISE, forces=False, dt_half=0.0, cfl=0.3,
                 accel=(0.0, -2.0), refl=(0.05, 0.95, 0.05, 0.95))
    g.step(6)
    o.step(6)
    for f in L.FIELDS:
        assert (g.get(f) == o.get(f)).all(), f


def test_wendland_steps_bitwise():
    st = ics.clustered(4000, seed=5, k=32)
    g, o = _pair(st, 32, kernel=1, flags=L.FLAG_BITWISE, forces=False, dt_half=0.0, cfl=0.3)
    g.step(4)
    o.step(4)
    for f in L.FIELDS:
        assert (g.get(f) == o.get(f)).all(), f


@pytest.mark.parametrize("hor,vert", [((-2.0, 3.0), (1.0, 1.5)), ((0.25, 0.5), (-4.0, 4.0))])
def test_anisotropic_box(hor, vert):
    st = ics.uniform(3000, seed=7)
    st = dict(st)
    st["x"] = hor[0] + (hor[1] - hor[0]) * st["x"]
    st["y"] = vert[0] + (vert[1] - vert[0]) * st["y"]
    g, o = _pair(st, 32, hor=hor, vert=vert)
    _check_knn(g, o)
    _check_forces(g, o)
    g, o = _pair(st, 32, hor=hor, vert=vert, flags=L.FLAG_BITWISE, forces=False, dt_half=0.0,
                 cfl=0.3)
    g.step(3)
    o.step(3)
    for f in L.FIELDS:
        assert (g.get(f) == o.get(f)).all(), f


@pytest.mark.parametrize("k", [32, 40, 64])
@pytest.mark.parametrize("mode", ["fp64", "bitwise", "fp32c"])
def test_lattice_ties(k, mode):
    # a square lattice has exact squared-distance ties at every shell; these
    # k cut a shell, so the k-th neighbours are chosen by (d2, ID, image)
    st = _lattice(24)
    prec = L.PREC_FP32C if mode == "fp32c" else L.PREC_FP64
    flags = L.FLAG_BITWISE if mode == "bitwise" else 0
    g, o = _pair(st, k, precision=prec, flags=flags)
    _check_knn(g, o)
    if mode != "fp32c":
        # on a lattice the pressure terms cancel: a is rounding noise, held
        # against the pressure-gradient scale c^2 / h
        for f in ("rho", "c"):
            d = np.abs(g.get(f) - o.get(f)) / o.get(f)
            assert d.max() <= (0 if mode == "bitwise" else 1e-12), f
        scale = (o.get("c") ** 2 / o.get("h")).max()
        for f in ("ax", "ay", "edot"):
            d = np.abs(g.get(f) - o.get(f)).max()
            assert d <= (0 if mode == "bitwise" else 1e-12 * scale), f
    # a second pass starts from the previous h (the bracketed path)
    g.calculate_forces()
    o.calculate_forces()
    _check_knn(g, o)


def test_smallest_k_and_n():
    st = ics.uniform(40, seed=3)
    g, o = _pair(st, 2)
    _check_knn(g, o)
    _check_forces(g, o)
    # 9 (n - 1) >= k: five particles give 36 candidates over the 9 images,
    # all of them neighbours (h up to 2.2 box lengths)
    for n, k in ((5, 36), (10, 64)):
        st = ics.uniform(n, seed=4)
        g, o = _pair(st, k)
        _check_knn(g, o)
        _check_forces(g, o)


@pytest.mark.parametrize("n,k", [(10, 64), (52, 32), (104, 64)])
def test_fp32c_neighbourhoods_beyond_half_a_box(n, k):
    # Where h reaches 0.45 L the fixed-point difference need not be the
    # minimum image, and the FP32C density and force take the displacement
    # from the FP64 positions instead: for every query (n = 10) or for part
    # of a wave (the others), the latter in the force's mixed-wave loop.  The
    # second pass runs the bracketed walk and its finish.
    st = ics.uniform(n, seed=6)
    g, o = _pair(st, k, precision=L.PREC_FP32C)
    wide = o.get("h") >= 0.45
    assert wide.any()
    if n > 10:
        assert not wide.all()
    for _ in range(2):
        _check_knn(g, o)
        for f in ("rho", "ax", "ay", "edot"):
            a, b = g.get(f), o.get(f)
            assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max(), f
        g.calculate_forces()
        o.calculate_forces()


def test_fp32c_pair_closer_than_fixed_point_step():
    # Two approaching particles inside one L / 2^32 step of the force
    # record's fixed-point positions: their displacement there is zero, and
    # the pair term (which vanishes with the distance) must not become NaN.
    st = dict(ics.uniform(2200, seed=5))
    for f in ("x", "y", "vx", "vy"):
        st[f] = st[f].copy()
    x0 = np.floor(0.3 * 2.0**32) / 2.0**32 + 2.0**-34
    y0 = np.floor(0.6 * 2.0**32) / 2.0**32 + 2.0**-34
    st["x"][0], st["y"][0] = x0, y0
    st["x"][1], st["y"][1] = x0 + 2.0**-35, y0
    st["vx"][0], st["vx"][1] = 0.1, -0.1
    g, o = _pair(st, 32, precision=L.PREC_FP32C)
    _check_knn(g, o)
    for f in ("rho", "ax", "ay", "edot"):
        a, b = g.get(f), o.get(f)
        assert np.isfinite(a).all(), f
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max(), f
    # and the state stays finite over steps
    g.step(3)
    for f in ("x", "y", "vx", "vy", "e"):
        assert np.isfinite(g.get(f)).all(), f


def test_fp32c_compact_cluster():
    # A cluster so compact that h spans only a few hundred steps of the
    # 32-bit fixed-point positions (h < 2^-13 L): those queries take their
    # displacements from the FP64 positions, so the density stays within
    # 1e-5 of the oracle's however small the cluster.  Its background
    # neighbours see the whole cluster near q = 1.
    rng = np.random.default_rng(12)
    n_c, n_b = 3000, 1000
    x = np.concatenate([0.4 + 1e-4 * rng.standard_normal(n_c), rng.random(n_b)])
    y = np.concatenate([0.6 + 1e-4 * rng.standard_normal(n_c), rng.random(n_b)])
    n = n_c + n_b
    st = {"x": x % 1.0, "y": y % 1.0, "vx": 0.01 * rng.standard_normal(n),
          "vy": 0.01 * rng.standard_normal(n), "e": np.ones(n), "mass": 1.0 / n}
    g, o = _pair(st, 32, precision=L.PREC_FP32C)
    h = o.get("h")
    assert (h < 2.0**-13).sum() > n_c // 2 and (h > 2.0**-13).any()
    for _ in range(2):
        _check_knn(g, o)
        rel = np.abs(g.get("rho") - o.get("rho")) / o.get("rho")
        assert rel.max() <= 1e-5
        for f in ("ax", "ay"):
            a, b = g.get(f), o.get(f)
            assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max(), f
        g.calculate_forces()
        o.calculate_forces()


def test_batch_with_different_boxes():
    boxes = [((0, 1), (0, 1)), ((-1, 2), (0, 0.5)), ((10, 11), (10, 13))]
    sts, ps = [], []
    for i, (hor, vert) in enumerate(boxes):
        st = dict(ics.blast(1500 + 700 * i, seed=31 + i))
        st["x"] = hor[0] + (hor[1] - hor[0]) * st["x"]
        st["y"] = vert[0] + (vert[1] - vert[0]) * st["y"]
        sts.append(st)
        ps.append(L.default_params(k=32, mass=st["mass"], gamma=5 / 3, hor=hor, vert=vert,
                                   dt_half=0.0, cfl=0.3, visc_alpha=1.0, visc_beta=2.0,
                                   flags=L.FLAG_BITWISE))
    b = L.Batch(ps, sts)
    b.step(3)
    for i, (hor, vert) in enumerate(boxes):
        st = sts[i]
        o = O.OracleSim(O.default_params(knn_mode=O.KNN_EXACT, use_cfl=1, cfl=0.3, k=32,
                                         mass=st["mass"], gamma=5 / 3, hor=hor, vert=vert,
                                         visc_alpha=1.0, visc_beta=2.0),
                        st["x"], st["y"], st["vx"], st["vy"], st["e"])
        o.step(3)
        for f in ("x", "y", "vx", "vy", "e", "rho", "h"):
            assert (b.get_sim(i, f) == o.get(f)).all(), (i, f)


def _two_patches(n, seed):
    """n particles in two 0.06-wide patches at opposite corners of the unit
    box: with 4 cells per particle the Hilbert curve crosses stretches of
    ~10^5 empty cells before, between and after them."""
    rng = np.random.default_rng(seed)
    h = n // 2
    x = np.concatenate([0.10 + 0.06 * rng.random(h), 0.82 + 0.06 * rng.random(n - h)])
    y = np.concatenate([0.12 + 0.06 * rng.random(h), 0.80 + 0.06 * rng.random(n - h)])
    return {"x": x, "y": y, "vx": 0.05 * rng.standard_normal(n),
            "vy": 0.05 * rng.standard_normal(n), "e": np.ones(n), "mass": 1.0 / n}


def test_sparse_grid_cell_ranges():
    """Cell ranges over long empty stretches (the queued fill): bitwise steps
    equal the oracle bit for bit, and after plain FP64 steps a walk pass finds
    the oracle's neighbour sets and forces."""
    st = _two_patches(20000, seed=41)
    g, o = _pair(st, 32, flags=L.FLAG_BITWISE, forces=False, dt_half=0.0, cfl=0.3)
    g.step(3)
    o.step(3)
    for f in ("x", "y", "vx", "vy", "e", "rho", "h"):
        assert (g.get(f) == o.get(f)).all(), f
    from test_gpu_parity import state_pass_vs_oracle
    common = dict(k=32, mass=st["mass"], gamma=5 / 3, hor=(0, 1), vert=(0, 1), dt_half=0.0,
                  cfl=0.3, visc_alpha=1.0, visc_beta=2.0)
    p = L.Sim(L.default_params(**common), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    p.step(3)
    o2, _ = state_pass_vs_oracle(p, O.default_params(knn_mode=O.KNN_EXACT, use_cfl=1, **common))
    _check_knn(p, o2)
    _check_forces(p, o2)
