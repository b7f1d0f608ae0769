"""CPU oracle for the sphugo SPH step -- TEST INFRASTRUCTURE ONLY.

A ctypes front end over ``oracle/liboracle.so``, the plain-C restatement of
the reference's path (``sph_oracle.c`` cites the Go file:line of each
function).  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` may import this package, and only as the
checker: the product library ``sphugo_amd`` never loads it.

Parity pins (each function of the restatement and what checks it):
  partition           8 known-answer cases of sim/partition_test.go (13 in
                      total with the empty and variation cases)
  tree / spheres      sim/bounding-sphere_test.go property (N=1,60,6000) plus
                      a strict every-ancestor test that exposes the merge bug
  kNN (exact, brute)  independent numpy brute force; faithful tree with the
                      corrected merge; permutation / translation invariance
  kernels             closed forms: normalisation 2*pi*int q F(q) dq * pref = 1,
                      DFpref*DF = Fpref*F', F(1) = DF(1) = 0 (scipy quad)
  density             square-lattice shell sum in closed form
  force               lattice at rest gives a = g and edot = 0; Galilean and
                      rotation invariance; sign and size of a pressure-gradient
                      acceleration against -grad P / rho
  integrator          uniformly accelerated lattice follows x0 + v0 t + g t^2/2;
                      wrap and reflection special cases incl. the `continue`
                      quirk of sim/sph.go:147-167
  reductions          sums checked against math.fsum
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

KNN_EXACT, KNN_TREE, KNN_TREE_FIXED, KNN_BRUTE = 0, 1, 2, 3
MONAGHAN, WENDLAND = 0, 1
FIELDS = ["x", "y", "vx", "vy", "rho", "c", "e", "edot", "ax", "ay", "epred",
          "vpx", "vpy", "h"]
FIELD = {n: i for i, n in enumerate(FIELDS)}
DBL_MAX = np.finfo(np.float64).max


class Params(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_int32), ("kernel", ctypes.c_int32),
        ("knn_mode", ctypes.c_int32), ("use_cfl", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("mass", ctypes.c_double),
        ("dt_half", ctypes.c_double), ("cfl", ctypes.c_double),
        ("accel", ctypes.c_double * 2), ("hor", ctypes.c_double * 2),
        ("vert", ctypes.c_double * 2), ("refl", ctypes.c_double * 4),
        ("visc_alpha", ctypes.c_double), ("visc_beta", ctypes.c_double),
        ("visc_eta2", ctypes.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIBPATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE],
                       check=True, stdout=subprocess.DEVNULL)
    return LIBPATH


_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        build()
    L = ctypes.CDLL(LIBPATH)
    vp = ctypes.c_void_p
    sig = {
        "orc_default_params": (None, [ctypes.POINTER(Params)]),
        "orc_partition": (ctypes.c_int64, [_D, _D, ctypes.c_int64, ctypes.c_int32, ctypes.c_double]),
        "orc_kernel_F": (ctypes.c_double, [ctypes.c_int32, ctypes.c_double]),
        "orc_kernel_DF": (ctypes.c_double, [ctypes.c_int32, ctypes.c_double]),
        "orc_kernel_Fpref": (ctypes.c_double, [ctypes.c_int32]),
        "orc_kernel_DFpref": (ctypes.c_double, [ctypes.c_int32]),
        "orc_hypot": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "orc_tree_make": (vp, [_D, _D, ctypes.c_int64, ctypes.c_int32]),
        "orc_tree_ncells": (ctypes.c_int64, [vp]),
        "orc_tree_cells": (None, [vp, _I64, _I64, _I64, _I64, _D, _D]),
        "orc_tree_positions": (None, [vp, _D, _D]),
        "orc_tree_free": (None, [vp]),
        "orc_create": (vp, [ctypes.POINTER(Params), ctypes.c_int64, _D, _D, _D, _D, _D]),
        "orc_destroy": (None, [vp]),
        "orc_append": (ctypes.c_int32, [vp, ctypes.c_int64, _D, _D, _D, _D, _D]),
        "orc_set_threads": (None, [vp, ctypes.c_int32]),
        "orc_step": (ctypes.c_int32, [vp, ctypes.c_int32]),
        "orc_calculate_forces": (ctypes.c_int32, [vp]),
        "orc_get_field": (None, [vp, ctypes.c_int32, _D]),
        "orc_set_field": (None, [vp, ctypes.c_int32, _D]),
        "orc_get_neighbours": (ctypes.c_int64, [vp, _I64, _D]),
        "orc_reductions": (None, [vp, _D]),
        "orc_get_term_scale": (None, [vp, ctypes.c_int32, _D]),
        "orc_get_neighbour_pos": (None, [vp, _D, _D]),
        "orc_total_energy": (ctypes.c_double, [vp]),
        "orc_total_density": (ctypes.c_double, [vp]),
        "orc_total_momentum": (ctypes.c_double, [vp]),
        "orc_current_step": (ctypes.c_int64, [vp]),
        "orc_cfl_dt_half": (ctypes.c_double, [vp]),
        "orc_boundary_ties": (ctypes.c_int64, [vp]),
        "orc_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(_D)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def default_params(**kw) -> Params:
    p = Params()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k in ("accel", "hor", "vert", "refl"):
            arr = getattr(p, k)
            for i, val in enumerate(v):
                arr[i] = val
        else:
            setattr(p, k, v)
    return p


def partition(x, y, orientation: int, middle: float):
    """Partition (sim/core.go:126-164); returns (len(a), x', y')."""
    x = _c64(x).copy()
    y = _c64(y).copy()
    na = lib().orc_partition(_dp(x), _dp(y), len(x), orientation, middle)
    return int(na), x, y


def hypot(p, q):
    """Go's math.Hypot as restated for the tree's pruning distances."""
    return lib().orc_hypot(p, q)


def kernel_F(kernel, q):
    return lib().orc_kernel_F(kernel, q)


def kernel_DF(kernel, q):
    return lib().orc_kernel_DF(kernel, q)


def kernel_prefactors(kernel):
    return lib().orc_kernel_Fpref(kernel), lib().orc_kernel_DFpref(kernel)


class Tree:
    """MakeCells (sim/core.go:93-105) over the given positions."""

    def __init__(self, x, y, fixed_merge=False):
        x = _c64(x)
        y = _c64(y)
        self.n = len(x)
        self._h = lib().orc_tree_make(_dp(x), _dp(y), self.n, int(fixed_merge))
        nc = lib().orc_tree_ncells(self._h)
        self.begin = np.zeros(nc, np.int64)
        self.end = np.zeros(nc, np.int64)
        self.lower = np.zeros(nc, np.int64)
        self.upper = np.zeros(nc, np.int64)
        self.box = np.zeros((nc, 4))
        self.sphere = np.zeros((nc, 3))
        lib().orc_tree_cells(self._h, self.begin.ctypes.data_as(_I64),
                             self.end.ctypes.data_as(_I64),
                             self.lower.ctypes.data_as(_I64),
                             self.upper.ctypes.data_as(_I64),
                             _dp(self.box), _dp(self.sphere))
        self.px = np.zeros(self.n)
        self.py = np.zeros(self.n)
        lib().orc_tree_positions(self._h, _dp(self.px), _dp(self.py))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_tree_free(self._h)
            self._h = None


class OracleSim:
    """The reference Simulation (sim/sph.go:14-198) restated on the CPU."""

    def __init__(self, params: Params, x, y, vx=None, vy=None, e=None,
                 threads: int = 1):
        n = len(x)
        self.n = n
        self.k = params.k
        self.params = params
        arrs = [_c64(a) if a is not None else np.zeros(n) for a in (x, y, vx, vy, e)]
        self._h = lib().orc_create(ctypes.byref(params), n, *[_dp(a) for a in arrs])
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())
        if threads != 1:
            lib().orc_set_threads(self._h, threads)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def append(self, x, y, vx=None, vy=None, e=None):
        """Particles of a source, appended with the next IDs (sim/sph.go:71-86)."""
        m = len(x)
        arrs = [_c64(a) if a is not None else None for a in (x, y, vx, vy, e)]
        self._check(lib().orc_append(self._h, m, *[_dp(a) if a is not None else None
                                                     for a in arrs]))
        self.n += m

    def step(self, nsteps: int = 1):
        self._check(lib().orc_step(self._h, nsteps))

    def calculate_forces(self):
        self._check(lib().orc_calculate_forces(self._h))

    def get(self, name: str) -> np.ndarray:
        out = np.zeros(self.n)
        lib().orc_get_field(self._h, FIELD[name], _dp(out))
        return out

    def set(self, name: str, values):
        v = _c64(values)
        lib().orc_set_field(self._h, FIELD[name], _dp(v))

    def neighbours(self):
        ids = np.zeros((self.n, self.k), np.int64)
        d2 = np.zeros((self.n, self.k))
        lib().orc_get_neighbours(self._h, ids.ctypes.data_as(_I64), _dp(d2))
        return ids, d2

    def term_scale(self, name: str) -> np.ndarray:
        """Per ID, sum over the neighbours of |term| of ax, ay or edot (the
        size of what the sum cancels), scaled as the sum is."""
        out = np.zeros(self.n)
        lib().orc_get_term_scale(self._h, {"ax": 0, "ay": 1, "edot": 2}[name], _dp(out))
        return out

    def neighbour_pos(self):
        """NNPos of the last pass, rows ascending by (d2, ID)."""
        px = np.zeros((self.n, self.k))
        py = np.zeros((self.n, self.k))
        lib().orc_get_neighbour_pos(self._h, _dp(px), _dp(py))
        return px, py

    def reductions(self):
        out = np.zeros(6)
        lib().orc_reductions(self._h, _dp(out))
        return out

    def total_energy(self):
        return lib().orc_total_energy(self._h)

    def total_density(self):
        return lib().orc_total_density(self._h)

    def total_momentum(self):
        return lib().orc_total_momentum(self._h)

    @property
    def current_step(self):
        return int(lib().orc_current_step(self._h))

    def cfl_dt_half(self):
        return lib().orc_cfl_dt_half(self._h)

    def boundary_ties(self):
        return int(lib().orc_boundary_ties(self._h))
