"""Hand cases that pin the magnitudes of the oracle's force, viscosity and
EDot (AccelerationAndEDot2D, sim/sph.go:327-401).

The expected values come from an independent model evaluated with mpmath at
50 digits, written from the textbook SPH form rather than from the
reference's code:

  support H_i = distance to the k-th neighbour, q = r / H_i
  W(r, H)   = sigma / H^2 * w(q),   dW/dr = sigma / H^3 * w'(q)
      cubic spline (Monaghan):  sigma = 40 / (7 pi),
          w = 6q^3 - 6q^2 + 1 (q < 1/2), 2 (1 - q)^3 (q <= 1)
      Wendland C2:              sigma = 7 / pi,  w = (1 - q)^4 (1 + 4q)
  rho_i = sum_j m W(r_ij, H_i)              (neighbours only, no self term)
  P_i = (gamma - 1) rho_i e_i,  A_i = P_i / rho_i^2
  Pi_ij (Monaghan 1992) when v_ij . r_ij < 0:
      mu = hbar (v_ij . r_ij) / (r_ij^2 + eta2),   eta2 absolute,
      Pi = (-alpha cbar mu + beta mu^2) / rhobar,  bars = arithmetic means,
      c_i^2 = gamma (gamma - 1) e_i,  hbar from the supports H
  a_i = -sum_j m (A_i + A_j + Pi_ij) grad_i W(r_ij, H_i) + g
  edot_i (the reference's form, quirk Q6: the textbook term
      A_i m v_ij . grad_i W without the factor D / (H_i^3 r_ij), where the
      reference writes dW/dr = D / H^3 * DF(q) with D = 240 / (7 pi) for
      the cubic spline (DF = w' / 6) and D = 14 / pi for Wendland
      (DF = w' / 2))

Cases: three and five particles with open boundaries, a periodic box whose
nearest neighbours are images, and a square lattice whose centre particle
moves.  Both signs of v.r occur, alpha, beta, eta2 are not the reference's
defaults, and densities, sound speeds and supports differ between pair
members, so a missing 1/2 in a mean, eta2 scaled by h^2, A_j in place of A_i
in EDot, a transposed r_ij or a dropped pressure term each move some result
by many orders of magnitude more than the 1e-14 bar
(test_mutations_are_detected)."""
import math

import mpmath as mp
import numpy as np
import pytest

import oracle as O

mp.mp.dps = 50
MON, WEN = 0, 1
BAR = 1e-14


def _w(kernel, q):
    if kernel == MON:
        return 6 * q**3 - 6 * q**2 + 1 if q < mp.mpf(1) / 2 else 2 * (1 - q) ** 3
    return (1 - q) ** 4 * (1 + 4 * q)


def _dw(kernel, q):
    if kernel == MON:
        return 18 * q**2 - 12 * q if q < mp.mpf(1) / 2 else -6 * (1 - q) ** 2
    return -20 * q * (1 - q) ** 3


def _sigma(kernel):
    return 40 / (7 * mp.pi) if kernel == MON else 7 / mp.pi


def _edot_pref(kernel):
    # dW/dr = D / H^3 * DF(q) in the reference's dimensionless derivative
    return 6 * _sigma(kernel) if kernel == MON else 2 * _sigma(kernel)


def model(case, mutation=None):
    """Exact (50-digit) density, sound speed, support, acceleration and EDot
    of every particle, plus the size of the summed terms and whether the
    k-th neighbour is strictly separated from the (k+1)-th."""
    x = [mp.mpf(float(v)) for v in case["x"]]
    y = [mp.mpf(float(v)) for v in case["y"]]
    vx = [mp.mpf(float(v)) for v in case["vx"]]
    vy = [mp.mpf(float(v)) for v in case["vy"]]
    e = [mp.mpf(float(v)) for v in case["e"]]
    n, k, kern = len(x), case["k"], case["kernel"]
    m, gam = mp.mpf(case["mass"]), mp.mpf(case["gamma"])
    al, be, eta2 = (mp.mpf(case[c]) for c in ("alpha", "beta", "eta2"))
    gx, gy = (mp.mpf(v) for v in case.get("accel", (0.0, 0.0)))
    box = case.get("box")  # None: open axes; else (L, L) periodic from 0
    offs = [(0, 0)] if box is None else [(i, j) for i in (-1, 0, 1) for j in (-1, 0, 1)]
    sig = _sigma(kern)
    # neighbours of each particle: (d2, j, image position of j)
    nbrs, sep = [], []
    for i in range(n):
        cands = []
        for j in range(n):
            if j == i:
                continue
            for ox, oy in offs:
                L = (0, 0) if box is None else box
                # image of j seen from i: p_j - o with o = (ox L, oy L)
                px, py = x[j] - ox * L[0], y[j] - oy * L[1]
                cands.append(((px - x[i]) ** 2 + (py - y[i]) ** 2, j, px, py))
        cands.sort(key=lambda c: c[0])
        nbrs.append(cands[:k])
        sep.append(len(cands) == k or cands[k][0] > cands[k - 1][0] * (1 + mp.mpf(10) ** -9))
    H = [mp.sqrt(nb[-1][0]) for nb in nbrs]
    rho = [sum(m * sig / H[i] ** 2 * _w(kern, mp.sqrt(d2) / H[i]) for d2, *_ in nbrs[i])
           for i in range(n)]
    c = [mp.sqrt(gam * (gam - 1) * e[i]) for i in range(n)]
    A = [(gam - 1) * rho[i] * e[i] / rho[i] ** 2 for i in range(n)]
    if mutation == "drop_A_j":
        Aj = [mp.mpf(0)] * n
    else:
        Aj = A
    half = {"cbar": 1, "rhobar": 1, "hbar": 1}
    if mutation in half:
        half[mutation] = 2
    ax, ay, ed, sx, sy, se = [], [], [], [], [], []
    signs = {"on": 0, "off": 0}  # contributing terms (q < 1) by the sign of v.r
    for i in range(n):
        tx = ty = te = mp.mpf(0)
        asx = asy = ase = mp.mpf(0)
        for d2, j, px, py in nbrs[i]:
            r = mp.sqrt(d2)
            rx, ry = px - x[i], py - y[i]  # r_ij = r_j - r_i
            if mutation == "transpose_r":
                rx, ry = -rx, -ry
            wx, wy = vx[j] - vx[i], vy[j] - vy[i]  # v_ij = v_j - v_i
            dot = wx * rx + wy * ry
            q = r / H[i]
            dWdr = sig / H[i] ** 3 * _dw(kern, q)
            Pi = mp.mpf(0)
            if q < 1:
                signs["on" if dot < 0 else "off"] += 1
            visc_on = dot < 0 if mutation != "visc_sign" else dot > 0
            if visc_on:
                hbar = half["hbar"] * (H[i] + H[j]) / 2
                cbar = half["cbar"] * (c[i] + c[j]) / 2
                rhobar = half["rhobar"] * (rho[i] + rho[j]) / 2
                soft = eta2 * hbar**2 if mutation == "eta2_h2" else eta2
                mu = hbar * dot / (r * r + soft)
                Pi = (-al * cbar * mu + be * mu * mu) / rhobar
            # -grad_i W = -(dW/dr) (r_i - r_j) / r = dW/dr * r_ij / r
            fx = m * (A[i] + Aj[j] + Pi) * dWdr * rx / r
            fy = m * (A[i] + Aj[j] + Pi) * dWdr * ry / r
            tx += fx
            ty += fy
            asx += abs(fx)
            asy += abs(fy)
            # textbook A_i m v_ij.grad_i W = A_i m dot dWdr / r; the
            # reference's EDot drops D / (H^3 r) from each term
            Ae = A[j] if mutation == "edot_uses_B" else A[i]
            et = Ae * m * dot * dWdr / r * (H[i] ** 3 * r / _edot_pref(kern))
            te += et
            ase += abs(et)
        ax.append(tx + gx)
        ay.append(ty + gy)
        ed.append(te)
        sx.append(asx + abs(gx))
        sy.append(asy + abs(gy))
        se.append(ase)
    return dict(rho=rho, c=c, h=H, ax=ax, ay=ay, edot=ed, sax=sx, say=sy, sedot=se, sep=sep,
                signs=signs)


def run_oracle(case):
    n = len(case["x"])
    box = case.get("box")
    kw = dict(k=case["k"], kernel=case["kernel"], knn_mode=O.KNN_BRUTE, mass=case["mass"],
              gamma=case["gamma"], visc_alpha=case["alpha"], visc_beta=case["beta"],
              visc_eta2=case["eta2"], accel=case.get("accel", (0.0, 0.0)))
    if box is not None:
        kw.update(hor=(0.0, box[0]), vert=(0.0, box[1]))
    s = O.OracleSim(O.default_params(**kw), case["x"], case["y"], case["vx"], case["vy"],
                    case["e"])
    s.set("vpx", case["vx"])
    s.set("vpy", case["vy"])
    s.set("epred", case["e"])
    s.calculate_forces()
    return {f: s.get(f) for f in ("rho", "c", "h", "ax", "ay", "edot")}


def _three(kernel=MON):
    # particle 0: nearest 1 (approaching: v.r < 0), support 2
    # particle 1: nearest 0 (approaching), support 2
    # particle 2: nearest 0 (receding: v.r > 0, no viscosity), support 1
    return dict(x=[0.1, 0.23, 0.05], y=[0.2, 0.17, 0.45], vx=[0.3, -0.2, 0.1],
                vy=[-0.1, 0.05, 0.4], e=[1.0, 2.5, 0.7], k=2, kernel=kernel, mass=0.013,
                gamma=5 / 3, alpha=1.3, beta=2.7, eta2=0.02, accel=(0.05, -0.2))


def _five(kernel=WEN):
    # k = 3: two contributing neighbours per particle, of both signs of v.r
    return dict(x=[0.40, 0.47, 0.36, 0.52, 0.44], y=[0.50, 0.53, 0.58, 0.41, 0.36],
                vx=[0.2, -0.3, 0.15, 0.05, -0.1], vy=[-0.05, 0.1, -0.2, 0.3, 0.25],
                e=[1.0, 0.6, 1.8, 1.2, 0.9], k=3, kernel=kernel, mass=0.007,
                gamma=1.4, alpha=0.9, beta=1.7, eta2=0.004)


def _periodic(kernel=MON):
    # unit periodic box: the nearest neighbours of the corner particles are
    # images across the boundary (r_ij from the image position p_j - o)
    return dict(x=[0.02, 0.97, 0.04, 0.5], y=[0.03, 0.05, 0.93, 0.5],
                vx=[0.4, -0.3, 0.1, 0.0], vy=[0.1, 0.2, -0.5, 0.05], e=[1.0, 1.5, 0.8, 1.1],
                k=3, kernel=kernel, mass=0.25, gamma=5 / 3, alpha=0.75, beta=1.5, eta2=0.01,
                box=(1.0, 1.0))


def _lattice(kernel=MON):
    # 9 x 9 square lattice at rest, spacing 0.05; k = 12 is three full
    # shells (s, sqrt2 s, 2 s), so every particle within 2 s of the centre
    # has the same density and support.  The centre particle moves: its
    # pressure terms cancel, its viscous terms do not.
    g = np.arange(-4, 5) * 0.05 + 0.5
    X, Y = np.meshgrid(g, g, indexing="ij")
    x, y = X.ravel(), Y.ravel()
    n = len(x)
    vx, vy = np.zeros(n), np.zeros(n)
    ctr = n // 2
    vx[ctr], vy[ctr] = 0.7, 0.4
    return dict(x=x, y=y, vx=vx, vy=vy, e=np.full(n, 1.3), k=12, kernel=kernel,
                mass=1.0 / n, gamma=5 / 3, alpha=1.1, beta=2.2, eta2=0.003, ctr=ctr)


CASES = {"three_mon": _three(MON), "three_wen": _three(WEN), "five_wen": _five(WEN),
         "five_mon": _five(MON), "periodic_mon": _periodic(MON), "lattice_mon": _lattice(MON),
         "lattice_wen": _lattice(WEN)}


def _compare(res, exp, i):
    for f in ("rho", "c", "h"):
        assert abs(res[f][i] - float(exp[f][i])) <= BAR * abs(float(exp[f][i])), (f, i)
    for f, s in (("ax", "sax"), ("ay", "say"), ("edot", "sedot")):
        assert abs(res[f][i] - float(exp[f][i])) <= BAR * float(exp[s][i]) + 1e-300, (f, i)


@pytest.mark.parametrize("name", list(CASES))
def test_hand_case(oracle_lib, name):
    case = CASES[name]
    exp = model(case)
    res = run_oracle(case)
    idx = [case["ctr"]] if "ctr" in case else range(len(case["x"]))
    checked = 0
    for i in idx:
        if not exp["sep"][i]:
            continue
        _compare(res, exp, i)
        checked += 1
    assert checked == len(idx)


def test_cases_cover_both_signs_and_distinct_pairs():
    """The open and periodic cases hold contributing terms with viscosity on
    (v.r < 0) and off, and pair members whose densities and supports
    differ (so every mean matters)."""
    for name in ("three_mon", "five_wen", "periodic_mon"):
        case = CASES[name]
        base = model(case)
        assert base["signs"]["on"] > 0 and base["signs"]["off"] > 0, name
        assert len(set(float(v) for v in base["rho"])) == len(case["x"]), name
        assert len(set(float(v) for v in base["h"])) >= 2, name
    assert model(CASES["lattice_mon"])["signs"]["on"] > 0


MUTATIONS = ["cbar", "rhobar", "hbar", "eta2_h2", "edot_uses_B", "visc_sign", "transpose_r",
             "drop_A_j"]


@pytest.mark.parametrize("mutation", MUTATIONS)
def test_mutations_are_detected(oracle_lib, mutation):
    """Each plausible mistake moves at least one checked result of the hand
    cases by far more than the bar, so the comparison above would fail."""
    worst = 0.0
    for name, case in CASES.items():
        exp = model(case, mutation=mutation)
        res = run_oracle(case)
        idx = [case["ctr"]] if "ctr" in case else range(len(case["x"]))
        for i in idx:
            for f, s in (("ax", "sax"), ("ay", "say"), ("edot", "sedot")):
                scale = float(exp[s][i])
                if scale > 0:
                    worst = max(worst, abs(res[f][i] - float(exp[f][i])) / scale)
    assert worst > 1e6 * BAR, (mutation, worst)


def test_lattice_centre_pressure_cancels():
    """With the centre at rest the lattice's pressure terms cancel: the
    model's acceleration is zero to 40 digits, so the moving case's
    acceleration is viscous only."""
    case = dict(CASES["lattice_mon"])
    case["vx"] = np.zeros(len(case["x"]))
    case["vy"] = np.zeros(len(case["x"]))
    exp = model(case)
    i = case["ctr"]
    assert abs(exp["ax"][i]) < mp.mpf(10) ** -12 * exp["sax"][i]
    assert abs(exp["ay"][i]) < mp.mpf(10) ** -12 * exp["say"][i]
    assert exp["edot"][i] == 0
