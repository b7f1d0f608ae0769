"""Seeded synthetic initial conditions for the SPH workloads.

The reference draws positions from Go's math/rand seeded with 12345678
(sim/config-parser.go:58-80, sim/core.go:76-91); that stream cannot be
reproduced outside Go, so every workload here uses its own documented
generator instead:

  * SplitMix64 (Steele, Lea & Flood 2014) on a per-stream counter, the top
    53 bits scaled by 2**-53 give uniforms in [0, 1);
  * Box-Muller on two uniforms gives normals;
  * positions are wrapped into the periodic unit box [0, 1)^2;
  * particle mass m = 1/N, gamma = 5/3, periodic box [0,1]^2, no reflection,
    no gravity.

Workloads (N, k, layout):
  uniform     N=2,200, k=32: uniform positions, v=0, e=1     (seed 12345678)
  blast       N=262,144, k=32: 70% uniform + 30% N((.5,.5), .05^2), v=0,
              e=1 except e=10 within minimum-image r<0.05 of (.5,.5) (seed 2)
  clustered   N=4,194,304, k=64: 80% in 16 Gaussians (centres uniform,
              sigma .02) + 20% uniform, v ~ N(0, .1^2) per axis, e=1 (seed 3)
  batch       512 simulations, N_i = floor(exp(U(ln 1024, ln 65536))),
              each a `blast` layout with seed 4*10**6 + i, k=32
  speed_test  N=100,000, k=32: the reference's timing harness input, uniform
              in the unit square, v=0, e=0.01, mass 1 (seed 12345678)
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class Stream:
    """Counter-based SplitMix64 stream: value i is mix(seed + (i+1)*golden)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            i = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
            z = self.seed + i * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.counter += n
        return z

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1) with 53 random bits each."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, n: int) -> np.ndarray:
        """n standard normals by Box-Muller (cos branch then sin branch)."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        t = 2.0 * math.pi * u2
        out = np.empty(2 * m)
        out[0::2] = r * np.cos(t)
        out[1::2] = r * np.sin(t)
        return out[:n]


def wrap_unit(a: np.ndarray) -> np.ndarray:
    """Wrap into [0, 1); values that round to 1.0 become 0.0."""
    w = a - np.floor(a)
    w[w >= 1.0] = 0.0
    return w


def _state(x, y, vx=None, vy=None, e=None, k=32, seed=0, name=""):
    n = len(x)
    return {
        "name": name, "seed": seed, "n": n, "k": k,
        "x": np.ascontiguousarray(x, np.float64),
        "y": np.ascontiguousarray(y, np.float64),
        "vx": np.zeros(n) if vx is None else np.ascontiguousarray(vx, np.float64),
        "vy": np.zeros(n) if vy is None else np.ascontiguousarray(vy, np.float64),
        "e": np.ones(n) if e is None else np.ascontiguousarray(e, np.float64),
        "mass": 1.0 / n,
    }


def uniform(n: int = 2200, seed: int = 12345678, k: int = 32):
    s = Stream(seed)
    x = s.uniform(n)
    y = s.uniform(n)
    return _state(x, y, k=k, seed=seed, name=f"uniform-{n}")


def blast(n: int = 262144, seed: int = 2, k: int = 32):
    s = Stream(seed)
    nb = int(0.3 * n)
    nu = n - nb
    x = np.empty(n)
    y = np.empty(n)
    x[:nu] = s.uniform(nu)
    y[:nu] = s.uniform(nu)
    x[nu:] = 0.5 + 0.05 * s.normal(nb)
    y[nu:] = 0.5 + 0.05 * s.normal(nb)
    x = wrap_unit(x)
    y = wrap_unit(y)
    dx = x - 0.5
    dy = y - 0.5
    dx -= np.rint(dx)
    dy -= np.rint(dy)
    e = np.ones(n)
    e[dx * dx + dy * dy < 0.05 * 0.05] = 10.0
    return _state(x, y, e=e, k=k, seed=seed, name=f"blast-{n}")


def clustered(n: int = 4194304, seed: int = 3, k: int = 64, nclusters: int = 16,
              sigma: float = 0.02, frac: float = 0.8, vsigma: float = 0.1):
    s = Stream(seed)
    cx = s.uniform(nclusters)
    cy = s.uniform(nclusters)
    nc = int(frac * n)
    nu = n - nc
    x = np.empty(n)
    y = np.empty(n)
    which = np.arange(nc) % nclusters
    x[:nc] = cx[which] + sigma * s.normal(nc)
    y[:nc] = cy[which] + sigma * s.normal(nc)
    x[nc:] = s.uniform(nu)
    y[nc:] = s.uniform(nu)
    x = wrap_unit(x)
    y = wrap_unit(y)
    vx = vsigma * s.normal(n)
    vy = vsigma * s.normal(n)
    return _state(x, y, vx=vx, vy=vy, k=k, seed=seed, name=f"clustered-{n}")


def speed_test(n: int = 100000, seed: int = 12345678):
    """The input of the reference's timing harness
    (examples/speed-test/speed-test.go:22-28): a UniformRectSpawner of n
    particles over the unit square (sim/config-parser.go:50-80: at rest,
    E = 0.01), with MakeConfig's ParticleMass 1 (sim/config-parser.go:131-149).
    Positions come from this module's generator, not Go's math/rand."""
    s = Stream(seed)
    x = s.uniform(n)
    y = s.uniform(n)
    st = _state(x, y, e=np.full(n, 0.01), k=32, seed=seed, name=f"speed-test-{n}")
    st["mass"] = 1.0
    return st


def batch_sizes(count: int = 512, seed: int = 4, lo: int = 1024, hi: int = 65536):
    s = Stream(seed)
    u = s.uniform(count)
    return np.floor(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int64)


def batch(count: int = 512, seed: int = 4, k: int = 32):
    sizes = batch_sizes(count, seed)
    return [blast(int(n), seed=seed * 10**6 + i, k=k) for i, n in enumerate(sizes)]


def clustered_pcg(n: int, seed: int, k: int = 64, nclusters: int = 16, sigma: float = 0.02,
                  frac: float = 0.8, vsigma: float = 0.1):
    """The clustered recipe drawn with numpy's PCG64 generator (several times
    faster than the SplitMix stream above): used where many independent
    initial conditions are made on the fly, e.g. one per served job."""
    g = np.random.default_rng(seed)
    cx = g.random(nclusters)
    cy = g.random(nclusters)
    nc = int(frac * n)
    which = np.arange(nc) % nclusters
    x = np.empty(n)
    y = np.empty(n)
    x[:nc] = cx[which] + sigma * g.standard_normal(nc)
    y[:nc] = cy[which] + sigma * g.standard_normal(nc)
    x[nc:] = g.random(n - nc)
    y[nc:] = g.random(n - nc)
    x = wrap_unit(x)
    y = wrap_unit(y)
    vx = vsigma * g.standard_normal(n)
    vy = vsigma * g.standard_normal(n)
    return _state(x, y, vx=vx, vy=vy, k=k, seed=seed, name=f"clustered-pcg-{n}")


def serve_jobs(count: int = 1000, seed: int = 5, lo: int = 1024, hi: int = 1 << 20):
    """Job mix of the serving workload: N log-uniform in [lo, hi], k in
    {32, 64}, steps uniform in [10, 200]."""
    g = np.random.default_rng(seed)
    n = np.floor(np.exp(g.uniform(math.log(lo), math.log(hi), count))).astype(np.int64)
    k = g.choice(np.array([32, 64]), count)
    steps = g.integers(10, 201, count)
    return [(int(a), int(b), int(c)) for a, b, c in zip(n, k, steps)]


SERVE_TEMPLATE_N = 1 << 20
SERVE_TEMPLATES = 16


def serve_ics(seed: int = 5):
    """Initial conditions of served jobs: 16 clustered templates of 2^20
    particles (the clustered recipe, template seeds seed*1000+t), shuffled
    once; job i takes a window of n particles of template t_i from a random
    start and shifts it by a random periodic offset (all drawn from the job's
    own seed), so every job has its own clustered layout at a few ms of
    host time.  Returns make(i, n, k) -> state."""
    base = []
    for t in range(SERVE_TEMPLATES):
        st = clustered_pcg(SERVE_TEMPLATE_N, seed=seed * 1000 + t)
        perm = np.random.default_rng(seed * 1000 + t).permutation(SERVE_TEMPLATE_N)
        base.append({c: st[c][perm] for c in ("x", "y", "vx", "vy")})

    def make(i, n, k):
        g = np.random.default_rng(seed * 10**6 + i)
        b = base[int(g.integers(SERVE_TEMPLATES))]
        idx = (int(g.integers(SERVE_TEMPLATE_N)) + np.arange(n)) % SERVE_TEMPLATE_N
        sx, sy = g.random(2)
        x = wrap_unit(b["x"][idx] + sx)
        y = wrap_unit(b["y"][idx] + sy)
        return {"x": x, "y": y, "vx": b["vx"][idx], "vy": b["vy"][idx], "e": np.ones(n),
                "mass": 1.0 / n, "k": k, "n": n}

    return make
