"""Independent numpy checks used to pin the oracle (test helpers only).

Brute-force periodic kNN written directly from the reference's distance
formula d2 = ((p_i + o) - p_j)^2 over the 9 offsets o in {-L,0,L}^2
(sim/nearest-neighbour.go:57-61,72,76), self excluded at every image
(:79), ordered by (d2, ID).  numpy evaluates each elementwise * and + as a
separate rounding, which is what the Go code does.
"""
import numpy as np


def brute_knn(x, y, k, L=1.0, chunk=128):
    n = len(x)
    offs = np.array([(i * L, j * L) for i in (-1, 0, 1) for j in (-1, 0, 1)])
    ids = np.zeros((n, k), np.int64)
    d2s = np.zeros((n, k))
    jj = np.arange(n)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        q = np.arange(a, b)
        # (nq, 9, n)
        px = (x[q][:, None] + offs[None, :, 0])[:, :, None]
        py = (y[q][:, None] + offs[None, :, 1])[:, :, None]
        dx = px - x[None, None, :]
        dy = py - y[None, None, :]
        d2 = dx * dx + dy * dy
        d2[np.arange(b - a), :, q] = np.inf  # self at every image
        d2 = d2.reshape(b - a, -1)
        cid = np.tile(jj, 9)
        for r in range(b - a):
            sel = np.argpartition(d2[r], k + 8)[: k + 8]
            order = np.lexsort((cid[sel], d2[r][sel]))[:k]
            ids[a + r] = cid[sel][order]
            d2s[a + r] = d2[r][sel][order]
    return ids, d2s


def square_lattice(n, jitter=0.0, seed=0):
    """n x n lattice in the unit box, positions (i + 1/2)/n."""
    g = (np.arange(n) + 0.5) / n
    x, y = np.meshgrid(g, g, indexing="ij")
    x = x.ravel().copy()
    y = y.ravel().copy()
    if jitter:
        rng = np.random.default_rng(seed)
        x += jitter / n * rng.standard_normal(len(x))
        y += jitter / n * rng.standard_normal(len(y))
        x %= 1.0
        y %= 1.0
    return x, y


def lattice_shells(k):
    """Shells of the square lattice in units of the spacing: list of
    (r^2, multiplicity) in increasing r, enough to hold k points, and the
    r^2 of the shell holding the k-th point."""
    m = int(np.ceil(np.sqrt(k))) + 3
    r2 = {}
    for i in range(-m, m + 1):
        for j in range(-m, m + 1):
            if i == 0 and j == 0:
                continue
            r2[i * i + j * j] = r2.get(i * i + j * j, 0) + 1
    shells = sorted(r2.items())
    cum = 0
    for s, (rr, cnt) in enumerate(shells):
        cum += cnt
        if cum >= k:
            return shells[: s + 1], rr
    raise ValueError
