"""Checkpoint and resume (GPU): a handle saved after some steps and loaded
again continues bit for bit as an uninterrupted run, in every precision
form, from step 0 and after an autotuned grid, through a file as well."""
import numpy as np
import pytest

from sphugo_amd import _lib as L
from sphugo_amd import ics

pytestmark = pytest.mark.gpu


def _params(st, precision=L.PREC_FP64, flags=0, k=32):
    return L.default_params(k=k, mass=st["mass"], gamma=5 / 3, hor=(0, 1), vert=(0, 1),
                            dt_half=0.0, cfl=0.3, visc_alpha=1.0, visc_beta=2.0,
                            precision=precision, flags=flags)


def _sim(st, **kw):
    return L.Sim(_params(st, **kw), st["x"], st["y"], st["vx"], st["vy"], st["e"])


def _same(a, b):
    for f in L.FIELDS:
        assert (a.get(f) == b.get(f)).all(), f
    assert (a.reductions() == b.reductions()).all()
    assert a.current_step == b.current_step


@pytest.mark.parametrize("precision,flags", [(L.PREC_FP64, 0), (L.PREC_FP32C, 0),
                                             (L.PREC_FP64, L.FLAG_BITWISE)])
def test_resume_is_bit_exact(precision, flags):
    st = ics.clustered(6000, seed=17, k=32)
    ref = _sim(st, precision=precision, flags=flags)
    ref.step(7)
    a = _sim(st, precision=precision, flags=flags)
    a.step(4)
    data = a.checkpoint()
    a.close()
    b = L.Sim.from_checkpoint(data)
    assert b.n == 6000 and b.k == 32
    b.step(3)
    _same(ref, b)


def test_resume_from_step_zero_and_through_a_file(tmp_path):
    st = ics.blast(3000, seed=23)
    ref = _sim(st)
    ref.step(3)
    a = _sim(st)
    path = tmp_path / "c.sphckpt"
    a.save(path)
    b = L.Sim.load(path)
    assert b.current_step == 0
    b.step(3)
    _same(ref, b)


def test_resume_keeps_the_autotuned_grid():
    st = ics.clustered(20000, seed=29, k=32)
    a = _sim(st)
    L.autotune(a)
    a.step(2)
    data = a.checkpoint()
    a.step(3)
    b = L.Sim.from_checkpoint(data)
    b.step(3)
    _same(a, b)


def test_checkpoint_rejects_batches_and_bad_images():
    sts = [ics.blast(1000 + 500 * i, seed=3 + i) for i in range(2)]
    batch = L.Batch([_params(s) for s in sts], sts)
    with pytest.raises(L.SphError, match="batch"):
        batch.checkpoint()
    st = ics.uniform(2200, seed=5)
    a = _sim(st)
    a.step(1)
    data = a.checkpoint()
    with pytest.raises(L.SphError, match="truncated"):
        L.Sim.from_checkpoint(data[:-1])
    bad = bytearray(data)
    bad[0:8] = b"XXXXXXXX"
    with pytest.raises(L.SphError, match="magic"):
        L.Sim.from_checkpoint(bytes(bad))


def test_resume_keeps_the_walk_shape():
    st = ics.clustered(6000, seed=14, k=64)
    a = _sim(st, k=64)
    L.set_tuning(a, subtiles=4, group=32, split=16, delta=0.02)
    a.step(2)
    b = L.Sim.from_checkpoint(a.checkpoint())
    assert L.get_tuning(b) == L.get_tuning(a)
    a.step(2)
    b.step(2)
    _same(a, b)


def test_load_refuses_an_edited_simulation_record():
    """The simulation record in an image must be the one its parameters
    give: an edited grid size (gx) is refused with SPH_EINVAL before any
    device work, as is an edited box."""
    import ctypes
    import struct
    st = ics.uniform(3000, seed=4)
    a = _sim(st)
    a.step(1)
    img = bytearray(a.checkpoint())
    base = 64 + ctypes.sizeof(L.Params)  # header, parameters, then the record
    gx_at = base + 8 + 8 + 6 * 4         # off, coff, n k kernel prec flags use_cfl
    bad = bytearray(img)
    struct.pack_into("<i", bad, gx_at, struct.unpack_from("<i", img, gx_at)[0] * 2)
    with pytest.raises(L.SphError, match="simulation record"):
        L.Sim.from_checkpoint(bytes(bad))
    lx_at = gx_at + 5 * 4 + 4 + 2 * 8    # gx gy lg pmode fmode open, x0 y0 -> lx
    bad = bytearray(img)
    struct.pack_into("<d", bad, lx_at, 2.0)
    with pytest.raises(L.SphError, match="simulation record"):
        L.Sim.from_checkpoint(bytes(bad))
    b = L.Sim.from_checkpoint(bytes(img))  # the unedited image still loads
    b.step(1)
    a.step(1)
    _same(a, b)
