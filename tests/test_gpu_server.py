"""Continuous-batching serve loop (GPU).

Bar: a job's results are bit-identical to the same simulation stepped alone
through sph_create + sph_step, whatever else shares the pool and whenever it
was admitted (pool ranges reused, jobs at different steps in one pass).
"""
import threading

import numpy as np
import pytest

from sphugo_amd import _lib as L
from sphugo_amd import ics

pytestmark = pytest.mark.gpu


def _p(st, k=32, precision=L.PREC_FP64, flags=0, kernel=0, **kw):
    kw.setdefault("dt_half", 0.0)
    kw.setdefault("cfl", 0.3)
    return L.default_params(k=k, kernel=kernel, precision=precision, flags=flags,
                            mass=st["mass"], gamma=5 / 3, hor=(0, 1), vert=(0, 1),
                            visc_alpha=1.0, visc_beta=2.0, **kw)


JOBS = [  # (n, k, kernel, nsteps, ic)
    (1500, 32, 0, 3, "blast"),
    (3100, 64, 1, 1, "blast"),
    (2048, 32, 0, 5, "clustered"),
    (5000, 48, 0, 2, "blast"),
    (4000, 32, 1, 4, "clustered"),
    (1200, 16, 0, 6, "blast"),
    (2600, 32, 0, 2, "clustered"),
]


def _state(i, n, ic):
    if ic == "blast":
        return ics.blast(n, seed=7 * 10**6 + i)
    return ics.clustered(n, seed=8 * 10**6 + i, k=32)


def _alone(p, st, nsteps):
    a = L.Sim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"])
    a.step(nsteps)
    out = {f: a.get(f) for f in L.RESULT_FIELDS}
    red = np.concatenate([a.reductions(), [a.total_momentum()]])
    a.close()
    return out, red


@pytest.mark.parametrize("precision,flags", [(L.PREC_FP64, 0), (L.PREC_FP32C, 0),
                                             (L.PREC_FP64, L.FLAG_BITWISE)])
def test_server_jobs_equal_alone(precision, flags):
    # capacity below the total: later jobs wait for ranges of earlier ones
    srv = L.Server(capacity=9000, max_jobs=3, kmax=64, precision=precision, flags=flags)
    sts, ps, ids = [], [], []
    for i, (n, k, kern, nsteps, ic) in enumerate(JOBS):
        st = _state(i, n, ic)
        p = _p(st, k=k, kernel=kern, precision=precision, flags=flags)
        sts.append(st)
        ps.append(p)
        ids.append(srv.submit(p, st, nsteps))
    srv.drain()
    for i, (n, k, kern, nsteps, ic) in enumerate(JOBS):
        info = srv.info(ids[i])
        assert info["state"] == "done" and info["steps"] == nsteps, (i, info)
        ref, red = _alone(ps[i], sts[i], nsteps)
        for f in L.RESULT_FIELDS:
            assert (srv.result(ids[i], f) == ref[f]).all(), (i, f)
        assert (srv.reductions(ids[i]) == red).all(), i
    s = srv.stats()
    assert s["jobs_done"] == len(JOBS) and s["jobs_failed"] == 0
    assert s["particle_steps"] == sum(n * st for n, _, _, st, _ in JOBS)
    assert 0 < s["occupancy"] <= 1 and s["p99_ms"] >= s["p50_ms"] > 0
    srv.close()


def test_server_concurrent_clients():
    srv = L.Server(capacity=20000, max_jobs=8, kmax=32)
    sts = [ics.blast(1000 + 300 * i, seed=9 * 10**6 + i) for i in range(8)]
    ps = [_p(st) for st in sts]
    got = {}
    errs = []

    def client(i):
        try:
            jid = srv.submit(ps[i], sts[i], 2 + i % 3)
            srv.wait(jid, timeout=120)
            got[i] = (jid, srv.result(jid, "x"), srv.result(jid, "e"))
        except Exception as exc:  # noqa: BLE001 - reported below
            errs.append(exc)

    th = [threading.Thread(target=client, args=(i,)) for i in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for i in range(8):
        ref, _ = _alone(ps[i], sts[i], 2 + i % 3)
        assert (got[i][1] == ref["x"]).all() and (got[i][2] == ref["e"]).all(), i
        srv.release(got[i][0])
    srv.close()


def test_server_rejects():
    srv = L.Server(capacity=4000, max_jobs=2, kmax=32)
    st = ics.blast(2000, seed=3)
    with pytest.raises(L.SphError):
        srv.submit(_p(st, k=48), st, 1)  # k above kmax
    with pytest.raises(L.SphError):
        srv.submit(_p(st, precision=L.PREC_FP32C), st, 1)  # precision of the server
    with pytest.raises(L.SphError):
        srv.submit(_p(st), st, 0)  # no steps
    big = ics.blast(5000, seed=4)
    with pytest.raises(L.SphError):
        srv.submit(_p(big), big, 1)  # above capacity
    jid = srv.submit(_p(st), st, 1)
    with pytest.raises(L.SphError):
        srv.result(jid, "c")  # not kept for finished jobs
    srv.wait(jid)
    with pytest.raises(L.SphError):
        srv.result(jid, "c")
    srv.release(jid)
    with pytest.raises(L.SphError):
        srv.info(jid)
    srv.close()


def test_server_failure_stays_with_its_job():
    # A job whose particle leaves the periodic images fails alone (its slot's
    # flag names it); the jobs sharing its passes finish with the results
    # they have alone.
    srv = L.Server(capacity=12000, max_jobs=4, kmax=32)
    good = []
    for i in range(3):
        st = _state(i, 2500, "blast")
        p = _p(st)
        good.append((srv.submit(p, st, 6), p, st))
    bad_st = dict(_state(9, 2000, "blast"))
    bad_st["vx"] = bad_st["vx"].copy()
    bad_st["vx"][17] = 1.0e5  # with a fixed half step of 1e-3 it leaves at once
    bad = srv.submit(_p(bad_st, dt_half=1e-3), bad_st, 6)
    srv.drain()
    info = srv.info(bad)
    assert info["state"] == "failed", info
    with pytest.raises(L.SphError, match="left the periodic images"):
        srv.wait(bad)
    for jid, p, st in good:
        assert srv.info(jid)["state"] == "done"
        ref, red = _alone(p, st, 6)
        for f in L.RESULT_FIELDS:
            assert (srv.result(jid, f) == ref[f]).all(), f
    s = srv.stats()
    assert s["jobs_done"] == 3 and s["jobs_failed"] == 1
    srv.close()
