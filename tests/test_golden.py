"""The committed fixtures of the small uniform case (tests/golden/c1.npz,
written by tests/golden/make_golden.py) against the oracle and against an
independent numpy brute force (CPU only)."""
import os

import numpy as np
import pytest

import oracle as O
from sphugo_amd import ics

import _np_ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLDEN))


def _params(st, **kw):
    base = dict(k=32, kernel=0, gamma=5 / 3, mass=st["mass"], hor=(0, 1), vert=(0, 1),
                visc_alpha=1.0, visc_beta=2.0, visc_eta2=0.01, knn_mode=O.KNN_EXACT)
    base.update(kw)
    return O.default_params(**base)


def test_inputs_regenerate(gold):
    st = ics.uniform(2200, seed=12345678)
    assert (st["x"] == gold["x0"]).all() and (st["y"] == gold["y0"]).all()


def test_knn_brute_force(gold):
    ids, d2 = _np_ref.brute_knn(gold["x0"], gold["y0"], 32)
    assert (ids == gold["pass0_ids"]).all()
    assert (np.sqrt(d2[:, -1]) == gold["pass0_h"]).all()


def test_oracle_pass0(oracle_lib, gold):
    st = ics.uniform(2200, seed=12345678)
    s = O.OracleSim(_params(st, use_cfl=1, cfl=0.3), st["x"], st["y"], st["vx"], st["vy"],
                    st["e"])
    s.set("vpx", st["vx"])
    s.set("vpy", st["vy"])
    s.set("epred", st["e"])
    s.calculate_forces()
    ids, _ = s.neighbours()
    assert (ids == gold["pass0_ids"]).all()
    for f in ("h", "rho", "c", "ax", "ay", "edot"):
        assert (s.get(f) == gold["pass0_" + f]).all(), f


def test_oracle_ten_steps(oracle_lib, gold):
    st = ics.uniform(2200, seed=12345678)
    s = O.OracleSim(_params(st, use_cfl=1, cfl=0.3), st["x"], st["y"], st["vx"], st["vy"],
                    st["e"])
    s.step(10)
    for f in ("x", "y", "vx", "vy", "e", "rho", "h"):
        assert (s.get(f) == gold["cfl10_" + f]).all(), f
    assert (s.reductions() == gold["cfl10_red"]).all()
    t = O.OracleSim(_params(st, use_cfl=0, dt_half=float(gold["strict_dt_half"][0]),
                            visc_alpha=0.75, visc_beta=1.5),
                    st["x"], st["y"], st["vx"], st["vy"], st["e"])
    t.step(10)
    for f in ("x", "y", "vx", "vy", "e"):
        assert (t.get(f) == gold["strict10_" + f]).all(), f
