"""The kNN walk's bracket adaptation (sph::adapt_delta, applied by k_reduce
after every pass), built for the host from the device header: hand-picked
cases of each branch, with n = 12800 queries (1/128 of them = 100 misses,
a quarter of that = 25)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from sphugo_amd import build as B  # noqa: E402

N = 12800
CASES = [
    # (d, d0, move, wide) -> next d
    ((0.025, 0.025, 101, 0), 0.03125),      # h moved for > 1/128: widen by a quarter
    ((0.025, 0.025, 100, 0), 0.025),        # exactly 1/128 is not above it; at d0: stay
    ((0.09, 0.025, 500, 10), 0.1),          # widening stops at 0.1
    ((0.1, 0.025, 500, 0), 0.1),
    ((0.025, 0.025, 0, 101), 0.02),         # annulus overflow: narrow by a fifth
    ((0.014, 0.025, 0, 300), 0.0125),       # narrowing stops at 0.0125
    ((0.025, 0.025, 150, 200), 0.02),       # both above 1/128, overflow dominates: narrow
    ((0.025, 0.025, 200, 200), 0.03125),    # a tie widens
    ((0.05, 0.025, 26, 0), 0.05),           # between a quarter and all of 1/128: hold
    ((0.05, 0.025, 0, 26), 0.05),
    ((0.05, 0.025, 25, 25), 0.045),         # at most a quarter: a tenth back towards d0
    ((0.0265, 0.025, 0, 0), 0.025),         # ... not past d0
    ((0.02, 0.025, 0, 0), 0.022),           # ... from below too
    ((0.0125, 0.0125, 0, 0), 0.0125),
]


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("adapt") / "adapt_delta")
    src = os.path.join(ROOT, "tests", "adapt_delta_main.cpp")
    r = subprocess.run([B.hipcc(), "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17",
                        "-I", B.CSRC, src, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def test_adapt_delta_cases(prog):
    lines = "".join(f"{d!r} {d0!r} {m} {w} {N}\n" for (d, d0, m, w), _ in CASES)
    r = subprocess.run([prog], input=lines, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = [float(v) for v in r.stdout.split()]
    assert len(got) == len(CASES)
    for ((args, want), g) in zip(CASES, got):
        assert g == pytest.approx(want, rel=1e-15, abs=0), (args, g, want)
