"""Batches of simulations, graph replay and the grid autotuner (GPU).

Bars: a simulation stepped inside a batch is bit-identical to the same
simulation stepped alone; graph replays are bit-identical to launching the
kernels one by one; the autotuner changes the cell grid but not the
neighbour sets, and in the bitwise form not a single result bit.
"""
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


def _states():
    return [ics.blast(n, seed=4 * 10**6 + i) for i, n in enumerate((1500, 3100, 2048, 5000))]


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_FP32C])
def test_batch_equals_alone(precision):
    sts = _states()
    ks = [32, 64, 32, 48]
    kerns = [0, 1, 0, 0]
    ps = [_p(st, k=k, kernel=kn, precision=precision) for st, k, kn in zip(sts, ks, kerns)]
    b = L.Batch(ps, sts)
    b.step(4)
    assert L.lib().sph_batch_count(b._h) == 4
    for i, st in enumerate(sts):
        a = L.Sim(ps[i], st["x"], st["y"], st["vx"], st["vy"], st["e"])
        a.step(4)
        for f in L.FIELDS:
            assert (b.get_sim(i, f) == a.get(f)).all(), (i, f)
        assert (b.reductions_sim(i) == a.reductions()).all(), i
    # the whole-handle readback is the concatenation of the simulations
    allx = b.get("x")
    assert (allx[:len(sts[0]["x"])] == b.get_sim(0, "x")).all()


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_FP32C])
def test_graph_replay_matches_eager(precision):
    st = ics.blast(6000, seed=12)
    g = L.Sim(_p(st, precision=precision), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    e = L.Sim(_p(st, precision=precision), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    L.set_graph(g, True)
    for step in range(6):
        g.step(1)
        e.step(1)
        # eager work and synchronisation between the replays
        for f in ("x", "vx", "e", "rho", "ax", "h"):
            assert (g.get(f) == e.get(f)).all(), (step, f)
        assert (g.reductions() == e.reductions()).all(), step


def test_graph_batch_replay():
    sts = _states()
    ps = [_p(st) for st in sts]
    g = L.Batch(ps, sts)
    e = L.Batch(ps, sts)
    L.set_graph(g, True)
    g.step(5)
    e.step(5)
    for i in range(len(sts)):
        for f in ("x", "y", "vx", "vy", "e"):
            assert (g.get_sim(i, f) == e.get_sim(i, f)).all(), (i, f)


def test_autotune_keeps_results():
    st = ics.clustered(20000, seed=9, k=64)
    pw = _p(st, k=64, flags=L.FLAG_BITWISE)
    a = L.Sim(pw, st["x"], st["y"], st["vx"], st["vy"], st["e"])
    ms, cpp = L.autotune(a)
    assert ms > 0 and cpp in (0.5, 1.0, 2.0, 4.0, 8.0, 16.0)
    b = L.Sim(pw, st["x"], st["y"], st["vx"], st["vy"], st["e"])
    a.step(3)
    b.step(3)
    for f in L.FIELDS:
        assert (a.get(f) == b.get(f)).all(), f
    ia, _ = a.neighbours(with_d2=False)
    ib, _ = b.neighbours(with_d2=False)
    assert (ia == ib).all()


def test_autotune_only_at_step_zero():
    st = ics.uniform(3000, seed=1)
    a = L.Sim(_p(st), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    a.step(1)
    with pytest.raises(L.SphError):
        L.autotune(a)


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_FP32C])
def test_autotune_plain_forms_within_contract(precision):
    """Outside the bitwise form a tuned grid changes the order in which
    neighbours are found and summed: neighbour sets and h stay identical,
    density and acceleration within the form's contract."""
    st = ics.clustered(20000, seed=19, k=64)
    p = _p(st, k=64, precision=precision)
    a = L.Sim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"])
    L.autotune(a)
    b = L.Sim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"])
    for s in (a, b):
        s.step(2)
    # one more pass from the same state on both
    for f in ("x", "y", "vx", "vy", "e", "vpx", "vpy", "epred"):
        a.set(f, b.get(f))
    a.calculate_forces()
    b.calculate_forces()
    ia, _ = a.neighbours(with_d2=False)
    ib, _ = b.neighbours(with_d2=False)
    assert (ia == ib).all()
    assert (a.get("h") == b.get("h")).all()
    tol = 1e-12 if precision == L.PREC_FP64 else 1e-5
    ra, rb = a.get("rho"), b.get("rho")
    assert (np.abs(ra - rb) / rb).max() <= tol
    if precision == L.PREC_FP64:
        for f in ("ax", "ay", "edot"):
            fa, fb = a.get(f), b.get(f)
            assert np.abs(fa - fb).max() <= 1e-12 * np.abs(fb).max(), f


@pytest.mark.parametrize("tuning", [dict(subtiles=4), dict(group=32), dict(split=16),
                                    dict(delta=0.02), dict(cells_per_particle=1.0, subtiles=4,
                                                           group=32, split=32, delta=0.05)])
def test_walk_shapes_match_oracle(tuning):
    """Every tile shape the autotuner (or a caller) may set finds the same
    exact neighbour sets: one pass after steps (the walk) against the oracle."""
    import oracle as O
    from test_gpu_parity import _check_forces, _check_knn, state_pass_vs_oracle
    st = ics.clustered(20000, seed=23, k=64)
    p = _p(st, k=64)
    g = L.Sim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"])
    L.set_tuning(g, **tuning)
    got = L.get_tuning(g)
    for key, v in tuning.items():
        assert got[key] == v, key
    g.step(3)
    op = O.default_params(knn_mode=O.KNN_EXACT, use_cfl=1, k=64, mass=st["mass"], gamma=5 / 3,
                          cfl=0.3, hor=(0, 1), vert=(0, 1), visc_alpha=1.0, visc_beta=2.0)
    o, nfb = state_pass_vs_oracle(g, op)
    # the walk resolves most queries (a wide bracket overflows more annuli)
    assert nfb < 0.5 * st["n"], nfb
    _check_knn(g, o)
    _check_forces(g, o)


def test_set_tuning_rejects():
    st = ics.uniform(3000, seed=2)
    a = L.Sim(_p(st), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    for bad in (dict(subtiles=6), dict(group=48), dict(split=4), dict(delta=0.5),
                dict(cells_per_particle=3.0)):
        with pytest.raises(L.SphError):
            L.set_tuning(a, **bad)


def test_bracket_adapts_to_motion():
    """A blast whose h moves faster than the default bracket: the bracket
    widens after passes where more than 1/128 of the queries missed, the
    share the walk leaves to the exact search falls, and a pass with the
    adapted bracket still finds the oracle's neighbour sets and forces."""
    import oracle as O
    from test_gpu_parity import _check_forces, _check_knn, state_pass_vs_oracle
    st = ics.blast(20000, seed=9, k=32)
    a = L.Sim(_p(st, k=32), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    moved = []
    for _ in range(8):
        c0 = a.knn_counters()
        a.step(1)
        c1 = a.knn_counters()
        moved.append((c1["miss_lo"] + c1["miss_hi"]) - (c0["miss_lo"] + c0["miss_hi"]))
    d = L.get_tuning(a)["delta"]
    assert 0.0125 <= d <= 0.1
    if max(moved[:3]) > st["n"] / 128:
        assert d > 0.025, (moved, d)
        assert moved[-1] < max(moved[:3]), (moved, d)
    op = O.default_params(knn_mode=O.KNN_EXACT, use_cfl=1, k=32, mass=st["mass"], gamma=5 / 3,
                          cfl=0.3, hor=(0, 1), vert=(0, 1), visc_alpha=1.0, visc_beta=2.0)
    o, _ = state_pass_vs_oracle(a, op)
    _check_knn(a, o)
    _check_forces(a, o)
    # a bracket set by the caller restarts every simulation from that value
    L.set_tuning(a, delta=0.02)
    assert L.get_tuning(a)["delta"] == 0.02
