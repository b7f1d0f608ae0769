"""Small fixed workloads to run under rocprofv3.

    python3 sphugo_amd/profile_run.py step [--precision fp32c|fp64] [--steps 3]
        the clustered 4M case (k = 64): create, one untimed step, then
        `steps` steps; every pass launches each kernel once
    python3 sphugo_amd/profile_run.py copy [--mib 1024]
        one device-to-device copy of known size (torch), the reference for
        scaling the FETCH_SIZE / WRITE_SIZE counters to bytes

Counter runs (one counter group per run):
    rocprofv3 --pmc FETCH_SIZE -d OUT -o run -- python3 sphugo_amd/profile_run.py step
    rocprofv3 --pmc WRITE_SIZE -d OUT -o run -- python3 sphugo_amd/profile_run.py step
"""
from __future__ import annotations

import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def run_step(precision: str, steps: int) -> None:
    from sphugo_amd import _lib as L
    from sphugo_amd import ics

    st = ics.clustered(4194304, seed=3, k=64)
    prec = L.PREC_FP32C if precision == "fp32c" else L.PREC_FP64
    p = L.default_params(k=64, precision=prec, mass=st["mass"], gamma=5 / 3, dt_half=0.0,
                         cfl=0.3, hor=(0, 1), vert=(0, 1), visc_alpha=1.0, visc_beta=2.0,
                         visc_eta2=0.01)
    sim = L.Sim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"])
    sim.step(1)
    sim.step(steps)
    print(f"stepped {steps} steps at N={st['n']} ({precision}); E = {sim.total_energy():.12g}",
          flush=True)
    sim.close()


def run_copy(mib: int) -> None:
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB (read {mib} MiB, wrote {mib} MiB)", flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["step", "copy"])
    ap.add_argument("--precision", default="fp32c", choices=["fp32c", "fp64"])
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--mib", type=int, default=1024)
    a = ap.parse_args()
    if a.mode == "step":
        run_step(a.precision, a.steps)
    else:
        run_copy(a.mib)


if __name__ == "__main__":
    main()
