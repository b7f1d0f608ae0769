"""Parity at the full sizes of the benchmark workloads (GPU).

blast 262,144 (k = 32): the oracle runs a pass in about a second, so the HIP
path is held to it directly.  clustered 4,194,304 (k = 64): one oracle pass
(about 20-40 s on the host cores), plus size-independent properties: the
bitwise form is invariant under relabelling the particles, and the FP32C
and FP64 forms select identical neighbour sets.

The first pass of a simulation searches every query exactly; every later
pass searches a bracket around each particle's h of the pass before (the
tile walk, with the exact search only for the queries it cannot resolve).
The *_walk tests step the simulation, copy its state into the oracle and
hold one more pass of both to the same bars: neighbour IDs and h bit-exact,
density and sound speed 1e-12 relative per particle, acceleration and EDot
1e-12 of the size of their terms per particle (FP64); FP32C: IDs and h
bit-exact, density 1e-5 relative."""
import os

import numpy as np
import pytest

import oracle as O
from sphugo_amd import _lib as L
from sphugo_amd import ics

from test_gpu_parity import _check_forces, state_pass_vs_oracle

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _params(st, k, **kw):
    kw.setdefault("gamma", 5 / 3)
    kw.setdefault("visc_alpha", 1.0)
    kw.setdefault("visc_beta", 2.0)
    kw.setdefault("dt_half", 0.0)
    kw.setdefault("cfl", 0.3)
    return dict(k=k, mass=st["mass"], hor=(0, 1), vert=(0, 1), **kw)


def _gpu(st, k, precision=L.PREC_FP64, flags=0, perm=None):
    x, y, vx, vy, e = (st[c] if perm is None else st[c][perm] for c in ("x", "y", "vx", "vy", "e"))
    return L.Sim(L.default_params(precision=precision, flags=flags, **_params(st, k)),
                 x, y, vx, vy, e)


def _oracle(st, k):
    return O.OracleSim(O.default_params(knn_mode=O.KNN_EXACT, use_cfl=1, **_params(st, k)),
                       st["x"], st["y"], st["vx"], st["vy"], st["e"], threads=THREADS)


def _first_pass(s, st):
    s.set("vpx", st["vx"])
    s.set("vpy", st["vy"])
    s.set("epred", st["e"])
    s.calculate_forces()


@pytest.fixture(scope="module")
def blast():
    return ics.blast(262144, seed=2, k=32)


@pytest.fixture(scope="module")
def blast_oracle(blast):
    o = _oracle(blast, 32)
    _first_pass(o, blast)
    ids, d2 = o.neighbours()
    return {"ids": ids, "d2": d2, **{f: o.get(f) for f in ("h", "rho", "c", "ax", "ay", "edot")},
            "scale": {f: o.term_scale(f) for f in ("ax", "ay", "edot")}}


class _Frozen:
    """Oracle results of one pass, read like an OracleSim."""

    def __init__(self, d):
        self.d = d

    def neighbours(self):
        return self.d["ids"], self.d["d2"]

    def get(self, f):
        return self.d[f]

    def term_scale(self, f):
        return self.d["scale"][f]


def test_blast_pass_fp64(blast, blast_oracle):
    g = _gpu(blast, 32)
    _first_pass(g, blast)
    gi, gd = g.neighbours()
    assert (gi.astype(np.int64) == blast_oracle["ids"]).all()
    assert (gd == blast_oracle["d2"]).all()
    assert (g.get("h") == blast_oracle["h"]).all()
    _check_forces(g, _Frozen(blast_oracle))


def test_blast_pass_bitwise(blast, blast_oracle):
    g = _gpu(blast, 32, flags=L.FLAG_BITWISE)
    _first_pass(g, blast)
    for f in ("h", "rho", "c", "ax", "ay", "edot"):
        assert (g.get(f) == blast_oracle[f]).all(), f


def test_blast_pass_fp32c(blast, blast_oracle):
    g = _gpu(blast, 32, precision=L.PREC_FP32C)
    _first_pass(g, blast)
    gi, _ = g.neighbours(with_d2=False)
    assert (gi.astype(np.int64) == blast_oracle["ids"]).all()
    rho = g.get("rho")
    assert (np.abs(rho - blast_oracle["rho"]) / blast_oracle["rho"]).max() <= 1e-5


def test_blast_steps_bitwise(blast):
    g = _gpu(blast, 32, flags=L.FLAG_BITWISE)
    o = _oracle(blast, 32)
    g.step(3)
    o.step(3)
    for f in L.FIELDS:
        assert (g.get(f) == o.get(f)).all(), f


@pytest.fixture(scope="module")
def clustered():
    return ics.clustered(4194304, seed=3, k=64)


def test_clustered_pass_fp64_vs_oracle(clustered):
    o = _oracle(clustered, 64)
    _first_pass(o, clustered)
    oh, orho = o.get("h"), o.get("rho")
    oi, _ = o.neighbours()
    oa = {f: o.get(f) for f in ("c", "ax", "ay", "edot")}
    oa.update({"s" + f: o.term_scale(f) for f in ("ax", "ay", "edot")})
    del o
    g = _gpu(clustered, 64)
    _first_pass(g, clustered)
    assert (g.get("h") == oh).all()
    gi, _ = g.neighbours(with_d2=False)
    assert (gi.astype(np.int64) == oi).all()
    del gi, oi
    assert (np.abs(g.get("rho") - orho) / orho).max() <= 1e-12
    assert (np.abs(g.get("c") - oa["c"]) / oa["c"]).max() <= 1e-12
    for f in ("ax", "ay", "edot"):
        a, b = g.get(f), oa[f]
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), f
        assert (np.abs(a - b) <= 1e-12 * oa["s" + f]).all(), f


def test_clustered_relabel_invariance_bitwise(clustered):
    n = clustered["n"]
    perm = np.random.default_rng(1).permutation(n)
    a = _gpu(clustered, 64, flags=L.FLAG_BITWISE)
    _first_pass(a, clustered)
    ref = {f: a.get(f) for f in ("h", "rho", "ax", "ay", "edot")}
    a.close()
    b = _gpu(clustered, 64, flags=L.FLAG_BITWISE, perm=perm)
    b.set("vpx", clustered["vx"][perm])
    b.set("vpy", clustered["vy"][perm])
    b.set("epred", clustered["e"][perm])
    b.calculate_forces()
    for f, v in ref.items():
        assert (b.get(f) == v[perm]).all(), f


def test_clustered_forms_agree(clustered):
    g64 = _gpu(clustered, 64)
    _first_pass(g64, clustered)
    h64, rho64 = g64.get("h"), g64.get("rho")
    i64, _ = g64.neighbours(with_d2=False)
    g64.close()
    g32 = _gpu(clustered, 64, precision=L.PREC_FP32C)
    _first_pass(g32, clustered)
    assert (g32.get("h") == h64).all()
    i32, _ = g32.neighbours(with_d2=False)
    assert (i32 == i64).all()
    del i32, i64
    assert (np.abs(g32.get("rho") - rho64) / rho64).max() <= 1e-5
    g32.close()
    # five steps: total energy of the two forms within 1e-8
    e = []
    for prec in (L.PREC_FP64, L.PREC_FP32C):
        s = _gpu(clustered, 64, precision=prec)
        s.step(5)
        e.append(s.reductions()[2])
        s.close()
    assert abs(e[1] - e[0]) <= 1e-8 * abs(e[0])


def _walk_vs_oracle(st, k, precision, steps=3):
    """Step the HIP path, then one pass of it and of the oracle from its
    copied state (the bracketed walk on the GPU)."""
    g = _gpu(st, k, precision=precision)
    g.step(steps)
    o, nfb = state_pass_vs_oracle(g, O.default_params(knn_mode=O.KNN_EXACT, use_cfl=1,
                                                      **_params(st, k)), THREADS)
    # the walk resolved the queries (only a few per cent take the exact search)
    assert nfb < 0.1 * st["n"], nfb
    oi, od = o.neighbours()
    gi, gd = g.neighbours()
    assert (gi.astype(np.int64) == oi).all()
    assert (gd == od).all()
    del gi, gd, oi, od
    assert (g.get("h") == o.get("h")).all()
    if precision == L.PREC_FP64:
        _check_forces(g, o)
    else:
        rho, orho = g.get("rho"), o.get("rho")
        assert (np.abs(rho - orho) / orho).max() <= 1e-5


def test_blast_walk_fp64(blast):
    _walk_vs_oracle(blast, 32, L.PREC_FP64)


def test_blast_walk_fp32c(blast):
    _walk_vs_oracle(blast, 32, L.PREC_FP32C)


def test_clustered_walk_fp64(clustered):
    _walk_vs_oracle(clustered, 64, L.PREC_FP64)


def test_clustered_walk_fp32c(clustered):
    _walk_vs_oracle(clustered, 64, L.PREC_FP32C)
