"""The C ABI library on the CPU: it loads, exports every entry point that
include/sphugo.h declares, and rejects invalid arguments before touching a
device (no compute call is made here)."""
import ctypes

import numpy as np
import pytest

from sphugo_amd import _lib as L


def test_exports_every_declared_symbol():
    lib = L.lib()
    declared = L.declared_symbols()
    assert len(declared) >= 40
    for name in declared:
        assert hasattr(lib, name), name
        assert name in L.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_defaults():
    assert L.lib().sph_version().decode()
    p = L.default_params()
    # reference defaults: MakeConfig (sim/config-parser.go:131-149), NN_SIZE,
    # viscosity constants (sim/sph.go:379-383)
    assert p.k == 32 and p.gamma == 1.66666 and p.mass == 1.0 and p.dt_half == 0.001
    assert (p.visc_alpha, p.visc_beta, p.visc_eta2) == (0.75, 1.5, 0.01)
    assert p.hor[0] == -L.DBL_MAX and p.hor[1] == L.DBL_MAX


def _create(**kw):
    n = kw.pop("n", 100)
    x = np.linspace(0.01, 0.99, n)
    base = dict(hor=(0, 1), vert=(0, 1), mass=1.0 / n)
    base.update(kw)
    p = L.default_params(**base)
    h = ctypes.c_void_p()
    rc = L.lib().sph_create(ctypes.byref(p), n, L.dptr(x), L.dptr(x), None, None, None, None,
                            ctypes.byref(h))
    return rc, L.lib().sph_last_error().decode()


@pytest.mark.parametrize("kw,needle", [
    (dict(k=1), "k must be"),
    (dict(k=65), "k must be"),
    (dict(kernel=7), "unknown kernel"),
    (dict(precision=9), "unknown precision"),
    (dict(precision=L.PREC_FP32C, flags=L.FLAG_BITWISE), "bitwise"),
    (dict(hor=(-L.DBL_MAX, 1.0)), "open and periodic"),
    (dict(vert=(0.0, L.DBL_MAX)), "open and periodic"),
    (dict(n=20, k=32, hor=(-L.DBL_MAX, L.DBL_MAX), vert=(-L.DBL_MAX, L.DBL_MAX)),
     "fewer than k"),
    (dict(hor=(1, 0)), "positive size"),
    (dict(n=1), "at least 2"),
    (dict(n=3, k=32), "fewer than k"),
    (dict(gamma=1.0), "gamma"),
    (dict(dt_half=0.0, cfl=0.0), "CFL"),
    (dict(visc_eta2=0.0), "visc_eta2"),
    (dict(visc_eta2=-0.01), "visc_eta2"),
])
def test_create_rejects(kw, needle):
    rc, msg = _create(**kw)
    assert rc == -1 and needle in msg, (rc, msg)


def test_server_create_rejects():
    h = ctypes.c_void_p()
    lib = L.lib()
    assert lib.sph_server_create(1, 4, 32, 0, 0, ctypes.byref(h)) == -1
    assert lib.sph_server_create(1 << 20, 0, 32, 0, 0, ctypes.byref(h)) == -1
    assert lib.sph_server_create(1 << 20, 4, 65, 0, 0, ctypes.byref(h)) == -1
    assert lib.sph_server_create(1 << 20, 4, 32, L.PREC_FP32C, L.FLAG_BITWISE,
                                 ctypes.byref(h)) == -1
    assert lib.sph_server_create(1 << 20, 4, 32, 0, L.FLAG_TIMING, ctypes.byref(h)) == -1


def test_null_handles():
    lib = L.lib()
    assert lib.sph_step(None, 1) == -1
    assert lib.sph_calculate_forces(None) == -1
    assert lib.sph_current_step(None) == -1
    assert lib.sph_server_wait(None, 1, 0.0) == -1
    lib.sph_destroy(None)
    lib.sph_server_destroy(None)


def test_checkpoint_load_rejects_bad_images():
    # validated before any device call
    import struct
    lib = L.lib()
    h = ctypes.c_void_p()

    def load(data):
        rc = lib.sph_checkpoint_load(data, len(data), ctypes.byref(h))
        return rc, lib.sph_last_error().decode()

    assert load(b"SPHCKPT1")[0] == -1  # shorter than the header
    rc, msg = load(b"NOTACKPT" + bytes(120))
    assert rc == -1 and "magic" in msg
    hdr = struct.pack("<8sIIqqiIIIdQ", b"SPHCKPT1", 2, 64, 1000, 0, 0,
                      ctypes.sizeof(L.Params), 0, 14, 4.0, 0)
    rc, msg = load(hdr + bytes(64))
    assert rc == -1 and "version" in msg
    hdr = struct.pack("<8sIIqqiIIIdQ", b"SPHCKPT1", 1, 64, -5, 0, 0,
                      ctypes.sizeof(L.Params), 0, 14, 4.0, 0)
    assert load(hdr + bytes(64))[0] == -1
    # valid header and parameters, but a simulation record that is not the
    # one the parameters give: refused before any device call
    def header(simdev_bytes):
        return struct.pack("<8sIIqqiIIIdQ", b"SPHCKPT1", 1, 64, 100, 3, 1,
                           ctypes.sizeof(L.Params), simdev_bytes, 14, 4.0, 0)

    p = L.default_params(hor=(0, 1), vert=(0, 1), mass=0.01)
    payload = bytes(100 * (14 * 8 + 4))
    # the record size of this build: the one size the layout check accepts
    sizes = [b for b in range(8, 1024, 8)
             if "layout" not in load(header(b) + bytes(p) + bytes(b) + payload)[1]]
    assert len(sizes) == 1, sizes
    simdev_bytes = sizes[0]
    hdr = header(simdev_bytes)
    body = bytes(p) + bytes(simdev_bytes) + payload
    rc, msg = load(hdr + body)
    assert rc == -1 and "simulation record" in msg, msg
    bad = L.default_params(hor=(0, 1), vert=(0, 1), mass=0.01, k=99)
    rc, msg = load(hdr + bytes(bad) + bytes(simdev_bytes) + payload)
    assert rc == -1 and "parameters" in msg, msg
    assert lib.sph_checkpoint_size(None, None) == -1
    assert lib.sph_checkpoint_save(None, None, 0) == -1
    assert h.value is None
