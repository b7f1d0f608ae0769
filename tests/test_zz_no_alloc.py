"""sph_read_neighbours_noalloc: the neighbour readback staged through the
library's fixed device buffer instead of buffers it allocates (GPU).

Bar: bit-identical to sph_read_neighbours (IDs and squared distances) on
two calls in a row, the second on a smaller simulation, with and without
the squared distances, on a batch of mixed k (sentinel entries past each
simulation's k), and after the buffer held larger rows; sizes beyond the
buffer are refused with SPH_EINVAL while sph_read_neighbours still serves
them."""
import ctypes

import numpy as np
import pytest

from sphugo_amd import _lib as L
from sphugo_amd import ics

pytestmark = pytest.mark.gpu


def _sim(n, k, seed):
    st = ics.uniform(n, seed=seed, k=k)
    p = L.default_params(k=k, mass=st["mass"], gamma=5 / 3, hor=(0, 1), vert=(0, 1),
                         visc_alpha=1.0, visc_beta=2.0)
    s = L.Sim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"])
    s.step(1)
    return s


def _noalloc(s, with_d2=True):
    ids = np.full((s.n, s.k), 0xA5A5A5A5, np.uint32)
    d2 = np.full((s.n, s.k), -1.0) if with_d2 else None
    rc = L.lib().sph_read_neighbours_noalloc(s._h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                             L.dptr(d2) if with_d2 else None)
    return rc, ids, d2


def test_matches_the_allocating_readback_on_two_calls():
    big = _sim(4000, 64, 3)
    rc, ids, d2 = _noalloc(big)
    assert rc == 0, L.lib().sph_last_error()
    ri, rd = big.neighbours()
    assert (ids == ri).all() and (d2 == rd).all()
    # second call: a smaller simulation, without squared distances; the
    # buffer still holds the larger rows of the first call
    small = _sim(900, 16, 4)
    rc, ids2, _ = _noalloc(small, with_d2=False)
    assert rc == 0, L.lib().sph_last_error()
    ri2, _ = small.neighbours(with_d2=False)
    assert (ids2 == ri2).all()
    rc, ids3, d23 = _noalloc(small)
    assert rc == 0
    ri3, rd3 = small.neighbours()
    assert (ids3 == ri3).all() and (d23 == rd3).all()


def test_batch_of_mixed_k():
    sts = [ics.uniform(n, seed=10 + i) for i, n in enumerate((700, 1200, 500))]
    ks = [16, 32, 24]
    ps = [L.default_params(k=k, mass=st["mass"], gamma=5 / 3, hor=(0, 1), vert=(0, 1))
          for st, k in zip(sts, ks)]
    b = L.Batch(ps, sts)
    b.step(1)
    rc, ids, d2 = _noalloc(b)
    assert rc == 0, L.lib().sph_last_error()
    ri, rd = b.neighbours()
    assert (ids == ri).all()
    assert np.array_equal(d2, rd, equal_nan=True)
    # rows of the k = 16 simulation end in sentinels
    assert (ids[:700, 16:] == 0xFFFFFFFF).all() and np.isnan(d2[:700, 16:]).all()


def test_refuses_sizes_beyond_the_buffer():
    s = _sim(20000, 64, 5)  # 1,280,000 entries > 2^20
    rc, _, _ = _noalloc(s)
    assert rc == -1  # SPH_EINVAL
    assert "exceed" in L.lib().sph_last_error().decode()
    ri, _ = s.neighbours(with_d2=False)  # the allocating readback serves it
    assert ri.shape == (20000, 64)
