"""Sources of the oracle (sim/sph.go:71-86): orc_append appends a spawner's
particles with the next IDs and Go's zero value elsewhere.  Pinned by: an
append before step 0 equals creating all particles at once; the fields of
appended particles right after the append; the CFL half step ignoring
particles that have no neighbours yet; and the error and no-op cases."""
import numpy as np
import pytest

from sphugo_amd import ics

FIELDS = ("x", "y", "vx", "vy", "e", "rho", "ax", "ay", "edot", "h")


def _params(O, n, **kw):
    base = dict(k=32, gamma=5 / 3, mass=1.0 / 2200, hor=(0, 1), vert=(0, 1),
                knn_mode=O.KNN_EXACT, use_cfl=0, dt_half=1e-4)
    base.update(kw)
    return O.default_params(**base)


def _spawn(m, seed):
    g = np.random.default_rng(seed)
    return {"x": 0.4 + 0.1 * g.random(m), "y": 0.4 + 0.1 * g.random(m),
            "vx": np.zeros(m), "vy": np.zeros(m), "e": np.full(m, 0.01)}


def test_append_before_step_zero_equals_creating_all(oracle_lib):
    O = oracle_lib
    st = ics.uniform(2000, seed=3)
    sp = _spawn(200, 4)
    allp = {f: np.concatenate([st[f], sp[f]]) for f in ("x", "y", "vx", "vy", "e")}
    a = O.OracleSim(_params(O, 2200), allp["x"], allp["y"], allp["vx"], allp["vy"], allp["e"])
    b = O.OracleSim(_params(O, 2200), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    b.append(sp["x"], sp["y"], sp["vx"], sp["vy"], sp["e"])
    assert b.n == 2200
    a.step(3)
    b.step(3)
    for f in FIELDS:
        assert (a.get(f) == b.get(f)).all(), f
    ia, _ = a.neighbours()
    ib, _ = b.neighbours()
    assert (ia == ib).all()


def test_appended_particles_start_from_zero(oracle_lib):
    O = oracle_lib
    st = ics.uniform(2000, seed=5)
    s = O.OracleSim(_params(O, 2000, use_cfl=1, cfl=0.3), st["x"], st["y"], st["vx"],
                    st["vy"], st["e"])
    s.step(2)
    before = {f: s.get(f) for f in FIELDS}
    dt_before = s.cfl_dt_half()
    sp = _spawn(50, 6)
    s.append(sp["x"], sp["y"], vx=np.full(50, 0.25), e=sp["e"])
    assert s.n == 2050
    for f in FIELDS:
        v = s.get(f)
        assert (v[:2000] == before[f]).all(), f
    assert (s.get("x")[2000:] == sp["x"]).all() and (s.get("y")[2000:] == sp["y"]).all()
    assert (s.get("vx")[2000:] == 0.25).all() and (s.get("vy")[2000:] == 0).all()
    assert (s.get("e")[2000:] == 0.01).all()
    for f in ("rho", "ax", "ay", "edot", "h", "vpx", "vpy", "epred"):
        assert (s.get(f)[2000:] == 0).all(), f
    # no neighbours yet: the CFL minimum is over the particles that have them
    assert s.cfl_dt_half() == dt_before
    s.step(2)
    assert (s.get("h")[2000:] > 0).all()
    assert np.isfinite(s.get("ax")).all()


def test_append_edge_cases(oracle_lib):
    O = oracle_lib
    st = ics.uniform(500, seed=7)
    s = O.OracleSim(_params(O, 500), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    s.append(np.zeros(0), np.zeros(0))
    assert s.n == 500
    s.append(np.zeros(0), None)  # nothing to append: no y needed
    assert s.n == 500
    with pytest.raises(RuntimeError):
        s.append(np.ones(1), None)  # a particle without y
    assert s.n == 500
