"""The HIP path (through the C ABI) against the committed fixtures of the
small uniform case, tests/golden/c1.npz (N = 2200, k = 32).

Bars: bitwise mode reproduces every fixture bit for bit (reductions, whose
device summation order differs from the oracle's, to 1e-13 relative); the
plain FP64 form keeps the neighbour IDs and h bit-exact and density and
acceleration within 1e-12 relative (acceleration normwise); FP32C keeps the
neighbour IDs and h bit-exact with density within 1e-5 relative.
"""
import os

import numpy as np
import pytest

from sphugo_amd import _lib as L

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLDEN))


def _sim(gold, precision=L.PREC_FP64, flags=0, **kw):
    n = len(gold["x0"])
    kw.setdefault("dt_half", 0.0)
    kw.setdefault("cfl", 0.3)
    kw.setdefault("visc_alpha", 1.0)
    kw.setdefault("visc_beta", 2.0)
    p = L.default_params(k=32, kernel=0, precision=precision, flags=flags, mass=1.0 / n,
                         gamma=5 / 3, hor=(0, 1), vert=(0, 1), visc_eta2=0.01, **kw)
    return L.Sim(p, gold["x0"], gold["y0"], np.zeros(n), np.zeros(n), np.ones(n))


def _pass0(s, gold):
    n = len(gold["x0"])
    s.set("vpx", np.zeros(n))
    s.set("vpy", np.zeros(n))
    s.set("epred", np.ones(n))
    s.calculate_forces()
    ids, _ = s.neighbours(with_d2=False)
    assert (ids == gold["pass0_ids"]).all(), "neighbour IDs"
    assert (s.get("h") == gold["pass0_h"]).all(), "h"


def _rel(a, b):
    """max |a - b| / max |b|; 0 when both are identically zero (EDot of a
    state at rest)."""
    d = np.abs(a - b).max()
    m = np.abs(b).max()
    return 0.0 if d == 0 else d / m


def test_bitwise_pass0(gold):
    s = _sim(gold, flags=L.FLAG_BITWISE)
    _pass0(s, gold)
    for f in ("rho", "c", "ax", "ay", "edot"):
        assert (s.get(f) == gold["pass0_" + f]).all(), f


def test_bitwise_ten_steps(gold):
    s = _sim(gold, flags=L.FLAG_BITWISE)
    s.step(10)
    for f in ("x", "y", "vx", "vy", "e", "rho", "h"):
        assert (s.get(f) == gold["cfl10_" + f]).all(), f
    red = s.reductions()
    assert np.abs(red - gold["cfl10_red"]).max() <= 1e-13 * np.abs(gold["cfl10_red"]).max()


def test_bitwise_strict_twin(gold):
    s = _sim(gold, flags=L.FLAG_BITWISE, dt_half=float(gold["strict_dt_half"][0]),
             visc_alpha=0.75, visc_beta=1.5)
    s.step(10)
    for f in ("x", "y", "vx", "vy", "e"):
        assert (s.get(f) == gold["strict10_" + f]).all(), f


def test_fp64_pass0(gold):
    s = _sim(gold)
    _pass0(s, gold)
    for f in ("rho", "c"):
        assert (np.abs(s.get(f) - gold["pass0_" + f]) / gold["pass0_" + f]).max() <= 1e-12, f
    for f in ("ax", "ay", "edot"):
        assert _rel(s.get(f), gold["pass0_" + f]) <= 1e-12, f


def test_fp32c_pass0(gold):
    s = _sim(gold, precision=L.PREC_FP32C)
    _pass0(s, gold)
    rho = s.get("rho")
    assert (np.abs(rho - gold["pass0_rho"]) / gold["pass0_rho"]).max() <= 1e-5
    for f in ("ax", "ay"):
        assert _rel(s.get(f), gold["pass0_" + f]) <= 1e-4, f
