"""Writes tests/golden/c1.npz: the small uniform case (N = 2200, k = 32).

Contents (all FP64 unless stated):
  x0, y0                 input positions (SplitMix64 stream, seed 12345678);
                         v = 0, e = 1, m = 1/N, gamma = 5/3, Monaghan
  pass0_ids (uint16)     neighbour IDs after the first force pass, rows
                         ascending by (d2, ID)
  pass0_{h,rho,c,ax,ay,edot}
  cfl10_{x,y,vx,vy,e,rho,h}, cfl10_red
                         state after 10 steps with the CFL half step
                         (CFL 0.3, alpha 1, beta 2, eta^2 0.01)
  strict10_{x,y,vx,vy,e}, strict10_red, strict_dt_half
                         the same at the reference's viscosity constants
                         (alpha 0.75, beta 1.5) with a fixed half step equal
                         to the CFL half step of the first pass

Values come from the oracle (oracle/, exact kNN mode, a CPU restatement of
the reference's step).  Before writing, the script checks them against an
independent numpy brute force: neighbour IDs and squared distances over the
9 periodic images, and the density sum in the reference's order.

    python3 tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from sphugo_amd import ics  # noqa: E402
import _np_ref  # noqa: E402

OUT = os.path.join(HERE, "c1.npz")
K = 32
STATE = ("x", "y", "vx", "vy", "e")


def params(st, **kw):
    base = dict(k=K, kernel=0, gamma=5 / 3, mass=st["mass"], hor=(0, 1), vert=(0, 1),
                visc_alpha=1.0, visc_beta=2.0, visc_eta2=0.01, knn_mode=O.KNN_EXACT)
    base.update(kw)
    return O.default_params(**base)


def numpy_density(d2, h, mass):
    """Monaghan density summed farthest first (sim/sph.go:306-323)."""
    pref = float.fromhex("0x1.5d3b3e3583243p+3")
    acc = np.zeros(len(h))
    for n in range(d2.shape[1] - 1, -1, -1):
        q = np.sqrt(d2[:, n]) / h
        f = np.where(q < 0.5, q * q * q - q * q + 1.0 / 6, (1 - q) * (1 - q) * (1 - q) / 3)
        acc = acc + f
    return (pref * mass) * acc / (h * h)


def main():
    O.build()
    st = ics.uniform(2200, seed=12345678)
    out = {"x0": st["x"], "y0": st["y"]}

    # first force pass
    s = O.OracleSim(params(st, use_cfl=1, cfl=0.3), st["x"], st["y"], st["vx"], st["vy"],
                    st["e"])
    s.set("vpx", st["vx"])
    s.set("vpy", st["vy"])
    s.set("epred", st["e"])
    s.calculate_forces()
    ids, d2 = s.neighbours()
    bi, bd = _np_ref.brute_knn(st["x"], st["y"], K)
    assert (ids == bi).all() and (d2 == bd).all(), "oracle kNN differs from numpy brute force"
    h = s.get("h")
    assert (h == np.sqrt(d2[:, -1])).all()
    rho = s.get("rho")
    rho_np = numpy_density(d2, h, st["mass"])
    assert np.abs(rho - rho_np).max() <= 1e-14 * rho.max(), "density differs from numpy"
    out["pass0_ids"] = ids.astype(np.uint16)
    for f in ("h", "rho", "c", "ax", "ay", "edot"):
        out["pass0_" + f] = s.get(f)
    dt_first = s.cfl_dt_half()

    # 10 steps with the CFL half step
    s = O.OracleSim(params(st, use_cfl=1, cfl=0.3), st["x"], st["y"], st["vx"], st["vy"],
                    st["e"])
    s.step(10)
    for f in STATE + ("rho", "h"):
        out["cfl10_" + f] = s.get(f)
    out["cfl10_red"] = s.reductions()

    # strict twin: reference viscosity constants, fixed half step
    s = O.OracleSim(params(st, use_cfl=0, dt_half=dt_first, visc_alpha=0.75, visc_beta=1.5),
                    st["x"], st["y"], st["vx"], st["vy"], st["e"])
    s.step(10)
    for f in STATE:
        out["strict10_" + f] = s.get(f)
    out["strict10_red"] = s.reductions()
    out["strict_dt_half"] = np.array([dt_first])
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {sorted(out)}")


if __name__ == "__main__":
    main()
