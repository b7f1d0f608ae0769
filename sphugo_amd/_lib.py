"""ctypes binding of libsphugo_amd.so (the C ABI in include/sphugo.h).

There is no CPU fallback: if the HIP library is missing or the device call
fails, the error is raised.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBPATH = os.path.join(HERE, "libsphugo_amd.so")
HEADER = os.path.join(ROOT, "include", "sphugo.h")

DBL_MAX = np.finfo(np.float64).max

# constants mirrored from include/sphugo.h
KERNEL_MONAGHAN, KERNEL_WENDLAND = 0, 1
PREC_FP64, PREC_FP32C = 0, 1
FLAG_BITWISE, FLAG_TIMING = 1, 2
FIELDS = ["x", "y", "vx", "vy", "rho", "c", "e", "edot", "ax", "ay", "epred",
          "vpx", "vpy", "h"]
FIELD = {n: i for i, n in enumerate(FIELDS)}
TIMERS = ["keys", "sort", "permute", "knn_density", "force", "reduce", "knn_walk",
          "knn_finish", "knn_fallback"]
ERRORS = {-1: "EINVAL", -2: "EDEVICE", -3: "ENOMEM", -4: "ESTATE", -5: "ERANGE"}


class SphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_int32), ("kernel", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("flags", ctypes.c_uint32),
        ("gamma", ctypes.c_double), ("mass", ctypes.c_double),
        ("dt_half", ctypes.c_double), ("cfl", ctypes.c_double),
        ("accel", ctypes.c_double * 2), ("hor", ctypes.c_double * 2),
        ("vert", ctypes.c_double * 2), ("refl", ctypes.c_double * 4),
        ("visc_alpha", ctypes.c_double), ("visc_beta", ctypes.c_double),
        ("visc_eta2", ctypes.c_double),
    ]


_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p

SIGNATURES = {
    "sph_default_params": (ctypes.c_int, [ctypes.POINTER(Params)]),
    "sph_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_int64, _D, _D, _D, _D, _D,
                                  _U32, ctypes.POINTER(_VP)]),
    "sph_destroy": (None, [_VP]),
    "sph_append": (ctypes.c_int, [_VP, ctypes.c_int64, _D, _D, _D, _D, _D]),
    "sph_step": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "sph_step_async": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "sph_synchronize": (ctypes.c_int, [_VP]),
    "sph_calculate_forces": (ctypes.c_int, [_VP]),
    "sph_read_field": (ctypes.c_int, [_VP, ctypes.c_int32, _D]),
    "sph_write_field": (ctypes.c_int, [_VP, ctypes.c_int32, _D]),
    "sph_read_neighbours": (ctypes.c_int, [_VP, _U32, _D]),
    "sph_read_neighbour_pos": (ctypes.c_int, [_VP, _D, _D]),
    "sph_read_neighbours_noalloc": (ctypes.c_int, [_VP, _U32, _D]),
    "sph_reductions": (ctypes.c_int, [_VP, _D]),
    "sph_total_energy": (ctypes.c_int, [_VP, _D]),
    "sph_total_density": (ctypes.c_int, [_VP, _D]),
    "sph_total_momentum": (ctypes.c_int, [_VP, _D]),
    "sph_current_step": (ctypes.c_int64, [_VP]),
    "sph_num_particles": (ctypes.c_int64, [_VP]),
    "sph_next_dt_half": (ctypes.c_int, [_VP, _D]),
    "sph_kernel_times": (ctypes.c_int, [_VP, _D, _I64]),
    "sph_reset_kernel_times": (ctypes.c_int, [_VP]),
    "sph_knn_stats": (ctypes.c_int, [_VP, _I64, _I64]),
    "sph_set_timing": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "sph_knn_counters": (ctypes.c_int, [_VP, _I64]),
    "sph_batch_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(Params), _I64,
                                        ctypes.POINTER(_D), ctypes.POINTER(_D), ctypes.POINTER(_D),
                                        ctypes.POINTER(_D), ctypes.POINTER(_D),
                                        ctypes.POINTER(_VP)]),
    "sph_batch_count": (ctypes.c_int32, [_VP]),
    "sph_batch_num_particles": (ctypes.c_int64, [_VP, ctypes.c_int32]),
    "sph_batch_read_field": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, _D]),
    "sph_batch_write_field": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, _D]),
    "sph_batch_reductions": (ctypes.c_int, [_VP, ctypes.c_int32, _D]),
    "sph_checkpoint_size": (ctypes.c_int, [_VP, _I64]),
    "sph_checkpoint_save": (ctypes.c_int, [_VP, ctypes.c_void_p, ctypes.c_int64]),
    "sph_checkpoint_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(_VP)]),
    "sph_set_graph": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "sph_autotune": (ctypes.c_int, [_VP, _D, _D]),
    "sph_get_tuning": (ctypes.c_int, [_VP, _D]),
    "sph_set_tuning": (ctypes.c_int, [_VP, _D]),
    "sph_server_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_uint32, ctypes.POINTER(_VP)]),
    "sph_server_destroy": (None, [_VP]),
    "sph_server_submit": (ctypes.c_int, [_VP, ctypes.POINTER(Params), ctypes.c_int64, _D, _D, _D,
                                         _D, _D, ctypes.c_int32, _I64]),
    "sph_server_wait": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_double]),
    "sph_server_drain": (ctypes.c_int, [_VP]),
    "sph_server_job_info": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32), _D]),
    "sph_server_read_result": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_int32, _D]),
    "sph_server_job_reductions": (ctypes.c_int, [_VP, ctypes.c_int64, _D]),
    "sph_server_release": (ctypes.c_int, [_VP, ctypes.c_int64]),
    "sph_server_stats": (ctypes.c_int, [_VP, _D]),
    "sph_last_error": (ctypes.c_char_p, []),
    "sph_version": (ctypes.c_char_p, []),
}


def declared_symbols():
    """Entry points declared in include/sphugo.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(
            f"{LIBPATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIBPATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise SphError(rc, lib().sph_last_error().decode())
    return rc


def dptr(a):
    return a.ctypes.data_as(_D)


def default_params(**kw) -> Params:
    p = Params()
    check(lib().sph_default_params(ctypes.byref(p)))
    for k, v in kw.items():
        if k in ("accel", "hor", "vert", "refl"):
            arr = getattr(p, k)
            for i, val in enumerate(v):
                arr[i] = val
        else:
            setattr(p, k, v)
    return p


class Sim:
    """A device-resident simulation handle (one sph_sim*)."""

    def __init__(self, params: Params, x, y, vx=None, vy=None, e=None, ids=None):
        n = len(x)
        self.n = n
        self.k = params.k
        self.params = params
        arrs = []
        for a in (x, y, vx, vy, e):
            arrs.append(None if a is None else np.ascontiguousarray(a, np.float64))
        self._keep = arrs
        ptrs = [dptr(a) if a is not None else None for a in arrs]
        idp = None
        if ids is not None:
            self._ids = np.ascontiguousarray(ids, np.uint32)
            idp = self._ids.ctypes.data_as(_U32)
        h = _VP()
        check(lib().sph_create(ctypes.byref(params), n, *ptrs, idp, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_destroy(self._h)
            self._h = None

    def append(self, x, y, vx=None, vy=None, e=None):
        """A source's particles, appended with the next IDs (sph_append)."""
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64)
                for a in (x, y, vx, vy, e)]
        check(lib().sph_append(self._h, len(arrs[0]),
                               *[None if a is None else dptr(a) for a in arrs]))
        self.n = int(lib().sph_num_particles(self._h))

    # checkpoint / resume (sph_checkpoint_*)
    def checkpoint(self) -> bytes:
        size = ctypes.c_int64()
        check(lib().sph_checkpoint_size(self._h, ctypes.byref(size)))
        buf = ctypes.create_string_buffer(size.value)
        check(lib().sph_checkpoint_save(self._h, buf, size.value))
        return buf.raw

    def save(self, path):
        with open(path, "wb") as f:
            f.write(self.checkpoint())

    @classmethod
    def from_checkpoint(cls, data: bytes) -> "Sim":
        data = data if isinstance(data, bytes) else bytes(data)
        h = _VP()
        check(lib().sph_checkpoint_load(data, len(data), ctypes.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self.params = Params.from_buffer_copy(bytes(data[64:64 + ctypes.sizeof(Params)]))
        self.k = self.params.k
        self.n = int(lib().sph_num_particles(h))
        return self

    @classmethod
    def load(cls, path) -> "Sim":
        with open(path, "rb") as f:
            return cls.from_checkpoint(f.read())

    def __del__(self):
        self.close()

    def step(self, nsteps=1):
        check(lib().sph_step(self._h, nsteps))

    def step_async(self, nsteps=1):
        check(lib().sph_step_async(self._h, nsteps))

    def synchronize(self):
        check(lib().sph_synchronize(self._h))

    def calculate_forces(self):
        check(lib().sph_calculate_forces(self._h))

    def get(self, name):
        out = np.empty(self.n)
        check(lib().sph_read_field(self._h, FIELD[name], dptr(out)))
        return out

    def set(self, name, values):
        v = np.ascontiguousarray(values, np.float64)
        assert v.shape == (self.n,)
        check(lib().sph_write_field(self._h, FIELD[name], dptr(v)))

    def neighbours(self, with_d2=True):
        ids = np.empty((self.n, self.k), np.uint32)
        d2 = np.empty((self.n, self.k)) if with_d2 else None
        check(lib().sph_read_neighbours(self._h, ids.ctypes.data_as(_U32),
                                        dptr(d2) if with_d2 else None))
        return ids, d2

    def neighbour_pos(self):
        """NNPos of the last pass (p_j - o), rows as neighbours()."""
        px = np.empty((self.n, self.k))
        py = np.empty((self.n, self.k))
        check(lib().sph_read_neighbour_pos(self._h, dptr(px), dptr(py)))
        return px, py

    def reductions(self):
        out = np.empty(6)
        check(lib().sph_reductions(self._h, dptr(out)))
        return out

    def scalar(self, fn):
        out = ctypes.c_double()
        check(getattr(lib(), fn)(self._h, ctypes.byref(out)))
        return out.value

    def total_energy(self):
        return self.scalar("sph_total_energy")

    def total_density(self):
        return self.scalar("sph_total_density")

    def total_momentum(self):
        return self.scalar("sph_total_momentum")

    def next_dt_half(self):
        return self.scalar("sph_next_dt_half")

    @property
    def current_step(self):
        return int(lib().sph_current_step(self._h))

    def kernel_times(self):
        ms = np.zeros(len(TIMERS))
        cnt = np.zeros(len(TIMERS), np.int64)
        check(lib().sph_kernel_times(self._h, dptr(ms), cnt.ctypes.data_as(_I64)))
        return dict(zip(TIMERS, ms)), dict(zip(TIMERS, cnt))

    def reset_kernel_times(self):
        check(lib().sph_reset_kernel_times(self._h))

    def set_timing(self, on):
        check(lib().sph_set_timing(self._h, 1 if on else 0))

    def knn_counters(self):
        out = np.zeros(7, np.int64)
        check(lib().sph_knn_counters(self._h, out.ctypes.data_as(_I64)))
        return dict(zip(["retries", "ties", "failures", "miss_lo", "miss_hi", "miss_ann",
                         "split_fallback"], out.tolist()))

    def knn_stats(self):
        r = ctypes.c_int64()
        t = ctypes.c_int64()
        check(lib().sph_knn_stats(self._h, ctypes.byref(r), ctypes.byref(t)))
        return r.value, t.value


class Batch(Sim):
    """Several independent simulations stepped together (sph_batch_create)."""

    def __init__(self, params_list, states):
        count = len(states)
        assert count == len(params_list) and count > 0
        P = (Params * count)(*params_list)
        ns = np.array([len(st["x"]) for st in states], np.int64)
        keep = []

        def ptrs(name):
            arr = (_D * count)()
            for i, st in enumerate(states):
                a = st.get(name)
                if a is None:
                    arr[i] = None
                else:
                    a = np.ascontiguousarray(a, np.float64)
                    keep.append(a)
                    arr[i] = dptr(a)
            return arr

        xs, ys, vxs, vys, es = (ptrs(n) for n in ("x", "y", "vx", "vy", "e"))
        h = _VP()
        check(lib().sph_batch_create(count, P, ns.ctypes.data_as(_I64), xs, ys, vxs, vys, es,
                                     ctypes.byref(h)))
        self._h = h
        self.count = count
        self.sizes = ns
        self.n = int(ns.sum())
        self.k = max(p.k for p in params_list)
        self.params = params_list[0]

    def get_sim(self, i, name):
        out = np.empty(int(self.sizes[i]))
        check(lib().sph_batch_read_field(self._h, i, FIELD[name], dptr(out)))
        return out

    def set_sim(self, i, name, values):
        v = np.ascontiguousarray(values, np.float64)
        check(lib().sph_batch_write_field(self._h, i, FIELD[name], dptr(v)))

    def reductions_sim(self, i):
        out = np.empty(6)
        check(lib().sph_batch_reductions(self._h, i, dptr(out)))
        return out


def set_graph(sim, on):
    check(lib().sph_set_graph(sim._h, 1 if on else 0))


def autotune(sim):
    ms = ctypes.c_double()
    cpp = ctypes.c_double()
    check(lib().sph_autotune(sim._h, ctypes.byref(ms), ctypes.byref(cpp)))
    return ms.value, cpp.value


TUNING = ["cells_per_particle", "delta", "split", "group", "subtiles"]


def get_tuning(sim):
    out = np.zeros(5)
    check(lib().sph_get_tuning(sim._h, dptr(out)))
    return dict(zip(TUNING, out.tolist()))


def set_tuning(sim, **kw):
    cur = get_tuning(sim)
    cur.update(kw)
    v = np.array([cur[n] for n in TUNING], np.float64)
    check(lib().sph_set_tuning(sim._h, dptr(v)))


JOB_STATES = ["queued", "running", "done", "failed"]
SERVER_STATS = ["jobs_done", "jobs_failed", "particle_steps", "wall_s", "particle_steps_per_s",
                "p50_ms", "p99_ms", "occupancy", "ticks", "in_flight", "tick_ms", "tick_gpu_ms",
                "queue_p50_ms", "queue_p99_ms", "run_p50_ms", "run_p99_ms"]
RESULT_FIELDS = ["x", "y", "vx", "vy", "rho", "e", "h"]


class Server:
    """Continuous-batching serve loop (sph_server_*): jobs of independent
    simulations share one device pool; a native worker thread steps them."""

    def __init__(self, capacity, max_jobs=256, kmax=64, precision=PREC_FP64, flags=0):
        h = _VP()
        check(lib().sph_server_create(capacity, max_jobs, kmax, precision, flags,
                                      ctypes.byref(h)))
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, params: Params, st, nsteps):
        arrs = [None if st.get(k) is None else np.ascontiguousarray(st[k], np.float64)
                for k in ("x", "y", "vx", "vy", "e")]
        n = len(arrs[0])
        jid = ctypes.c_int64()
        check(lib().sph_server_submit(self._h, ctypes.byref(params), n,
                                      *[dptr(a) if a is not None else None for a in arrs],
                                      nsteps, ctypes.byref(jid)))
        self._n[jid.value] = n
        return jid.value

    def wait(self, jid, timeout=-1.0):
        check(lib().sph_server_wait(self._h, jid, float(timeout)))

    def drain(self):
        check(lib().sph_server_drain(self._h))

    def info(self, jid):
        state = ctypes.c_int32()
        out = np.zeros(4)
        check(lib().sph_server_job_info(self._h, jid, ctypes.byref(state), dptr(out)))
        return {"state": JOB_STATES[state.value], "queue_ms": out[0], "run_ms": out[1],
                "latency_ms": out[2], "steps": int(out[3])}

    def result(self, jid, name):
        out = np.empty(self._n[jid])
        check(lib().sph_server_read_result(self._h, jid, FIELD[name], dptr(out)))
        return out

    def reductions(self, jid):
        out = np.empty(7)
        check(lib().sph_server_job_reductions(self._h, jid, dptr(out)))
        return out

    def release(self, jid):
        check(lib().sph_server_release(self._h, jid))
        self._n.pop(jid, None)

    def stats(self):
        out = np.zeros(len(SERVER_STATS))
        check(lib().sph_server_stats(self._h, dptr(out)))
        return dict(zip(SERVER_STATS, out.tolist()))
