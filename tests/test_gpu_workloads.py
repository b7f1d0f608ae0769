"""The batched and served workloads at their own sizes (GPU).

batch   512 blast simulations, N log-uniform in [1024, 65536] (seed 4, about
        7.9M particles in all), k = 32: stepped in one batch (FP32C with
        graph replay, the measured form; and the bitwise FP64 form).  A
        seeded sample of 8 simulations must equal their standalone runs bit
        for bit, and in the bitwise form the oracle's as well.
serve   64 jobs of the serving mix (N log-uniform in [1024, 2^20], k in {32,
        64}, steps uniform in [10, 200], clustered initial conditions per
        job) submitted with staggered arrivals into a pool smaller than
        their sum, so that jobs are admitted while others run and share
        passes at different steps of their lives: every job equals its
        standalone run bit for bit.  A bitwise FP64 pool with one small job
        per k among others: those jobs equal the oracle bit for bit.
"""
import os
import time

import numpy as np
import pytest

import oracle as O
from sphugo_amd import _lib as L
from sphugo_amd import ics

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _p(st, precision=L.PREC_FP64, flags=0):
    return L.default_params(k=st["k"], precision=precision, flags=flags, mass=st["mass"],
                            gamma=5 / 3, dt_half=0.0, cfl=0.3, hor=(0, 1), vert=(0, 1),
                            visc_alpha=1.0, visc_beta=2.0, visc_eta2=0.01)


def _op(st):
    return O.default_params(k=st["k"], knn_mode=O.KNN_EXACT, use_cfl=1, mass=st["mass"],
                            gamma=5 / 3, cfl=0.3, hor=(0, 1), vert=(0, 1), visc_alpha=1.0,
                            visc_beta=2.0, visc_eta2=0.01)


@pytest.fixture(scope="module")
def batch512():
    return ics.batch(512, seed=4, k=32)


def _sample(count, m, seed):
    return sorted(np.random.default_rng(seed).choice(count, m, replace=False).tolist())


def test_batch512_fp32c_graph_equals_alone(batch512):
    sts = batch512
    assert len(sts) == 512 and 7_000_000 < sum(st["n"] for st in sts) < 9_000_000
    b = L.Batch([_p(st, L.PREC_FP32C) for st in sts], sts)
    L.set_graph(b, True)
    b.step(2)
    for i in _sample(512, 8, 41):
        st = sts[i]
        a = L.Sim(_p(st, L.PREC_FP32C), st["x"], st["y"], st["vx"], st["vy"], st["e"])
        a.step(2)
        for f in L.FIELDS:
            assert (b.get_sim(i, f) == a.get(f)).all(), (i, f)
        assert (b.reductions_sim(i) == a.reductions()).all(), i
        a.close()
    b.close()


def test_batch512_bitwise_matches_oracle(batch512):
    sts = batch512
    b = L.Batch([_p(st, flags=L.FLAG_BITWISE) for st in sts], sts)
    b.step(2)
    for i in _sample(512, 8, 42):
        st = sts[i]
        o = O.OracleSim(_op(st), st["x"], st["y"], st["vx"], st["vy"], st["e"], threads=THREADS)
        o.step(2)
        a = L.Sim(_p(st, flags=L.FLAG_BITWISE), st["x"], st["y"], st["vx"], st["vy"], st["e"])
        a.step(2)
        for f in L.FIELDS:
            got = b.get_sim(i, f)
            assert (got == a.get(f)).all(), (i, f, "alone")
            assert (got == o.get(f)).all(), (i, f, "oracle")
        a.close()
    b.close()


@pytest.fixture(scope="module")
def serve_make():
    return ics.serve_ics(seed=5)


def _alone(p, st, nsteps):
    a = L.Sim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"])
    a.step(nsteps)
    out = {f: a.get(f) for f in L.RESULT_FIELDS}
    red = np.concatenate([a.reductions(), [a.total_momentum()]])
    a.close()
    return out, red


def test_serve_mix_64_jobs_equal_alone(serve_make):
    jobs = ics.serve_jobs(64, seed=5)
    assert max(n for n, _, _ in jobs) > 500_000 and {k for _, k, _ in jobs} == {32, 64}
    total = sum(n for n, _, _ in jobs)
    cap = 1 << 22
    assert cap < total  # later jobs wait for ranges that earlier jobs free
    srv = L.Server(cap, max_jobs=16, kmax=64, precision=L.PREC_FP32C)
    ids = []
    for i, (n, k, s) in enumerate(jobs):
        st = serve_make(i, n, k)
        ids.append(srv.submit(_p(st, L.PREC_FP32C), st, s))
        time.sleep(0.002)  # staggered arrivals: the pool is already running
    srv.drain()
    stats = srv.stats()
    assert stats["jobs_done"] == 64 and stats["jobs_failed"] == 0
    for i, (n, k, s) in enumerate(jobs):
        info = srv.info(ids[i])
        assert info["state"] == "done" and info["steps"] == s, (i, info)
        st = serve_make(i, n, k)
        ref, red = _alone(_p(st, L.PREC_FP32C), st, s)
        for f in L.RESULT_FIELDS:
            assert (srv.result(ids[i], f) == ref[f]).all(), (i, f)
        assert (srv.reductions(ids[i]) == red).all(), i
        srv.release(ids[i])
    srv.close()


def test_serve_bitwise_jobs_match_oracle(serve_make):
    jobs = ics.serve_jobs(64, seed=5)
    small = sorted(range(64), key=lambda i: jobs[i][0])
    # the smallest job of each k is checked against the oracle; six more
    # small jobs share its passes
    checked = [next(i for i in small if jobs[i][1] == k) for k in (32, 64)]
    others = [i for i in small if i not in checked][:6]
    srv = L.Server(1 << 18, max_jobs=4, kmax=64, precision=L.PREC_FP64, flags=L.FLAG_BITWISE)
    ids = {}
    for i in others[:3] + checked[:1] + others[3:] + checked[1:]:
        n, k, s = jobs[i]
        st = serve_make(i, n, k)
        ids[i] = srv.submit(_p(st, flags=L.FLAG_BITWISE), st, s)
        time.sleep(0.002)
    srv.drain()
    for i in checked:
        n, k, s = jobs[i]
        st = serve_make(i, n, k)
        o = O.OracleSim(_op(st), st["x"], st["y"], st["vx"], st["vy"], st["e"], threads=THREADS)
        o.step(s)
        for f in L.RESULT_FIELDS:
            assert (srv.result(ids[i], f) == o.get(f)).all(), (i, f)
        red = srv.reductions(ids[i])
        assert abs(red[0] - o.total_energy()) <= 1e-12 * abs(o.total_energy()), i
    for i in others:
        n, k, s = jobs[i]
        st = serve_make(i, n, k)
        ref, _ = _alone(_p(st, flags=L.FLAG_BITWISE), st, s)
        for f in L.RESULT_FIELDS:
            assert (srv.result(ids[i], f) == ref[f]).all(), (i, f)
    srv.close()
