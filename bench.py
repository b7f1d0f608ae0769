"""Throughput of the SPH step on one MI355X.

Workload (default): the clustered 4M case -- N = 4,194,304 particles in the
periodic unit box, 80% in 16 Gaussians (centres uniform, sigma 0.02) and 20%
uniform, v ~ N(0, 0.1^2) per axis, e = 1, m = 1/N, gamma = 5/3, k = 64,
Monaghan kernel, viscosity alpha = 1, beta = 2, eta^2 = 0.01, CFL 0.3.

A step is one Simulation.Step (drift, kNN, density, force, kick, drift,
wrap), state resident in HBM.  `value` is particle-steps/s of the
FP32-compensated form; the plain FP64 form is timed beside it.  The dominant
kernel's roofline uses per-launch HIP event timings taken on the library's
own stream.  The CPU baseline is the oracle's faithful restatement of the
reference (tree build + 9-image tree kNN), timed on this host.

The serving workload (`serve` in the line) pushes 1000 jobs with Poisson
arrivals through the continuous-batching server and reports throughput and
p50/p99 submit-to-done latency.  `speed_test` times the reference's own
harness (examples/speed-test/speed-test.go: 100,000 uniform particles on
MakeConfig's defaults, 20 Step() calls) on the GPU, beside one thread of the
oracle on the same input in the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from sphugo_amd import _lib as L  # noqa: E402
from sphugo_amd import ics  # noqa: E402

HBM_PEAK_GBS = 8000.0

METRIC = "particle-steps/s at 4M (FP32-comp vs plain FP64); p50/p99 job latency in serve"


def algo_bytes(kernel: str, k: int, precision: str = "fp32c") -> int:
    """Algorithmic HBM bytes per particle of one launch: own-particle streams
    read or written once (state 8 B, IDs and keys 4 B), materialised
    neighbour lists 4 B per entry; neighbour gathers count as on-chip reuse."""
    rec = 32 if precision == "fp32c" else 64            # force record (RecF / RecD)
    uxy = 8 if precision == "fp32c" else 0              # fixed-point positions
    return {
        "keys": 32 + 4,                                   # x, y, vx, vy -> key
        "permute": 4 + 4 + 9 * 8 + 4 + 10 * 8 + 4 + 1 + uxy,  # perm, key, gather, write
        "knn_walk": 16 + 8 + 4 * k + 4,                   # xy, h_prev -> list, count
        "knn_finish": 4 + 16 + 3 * 8 + 4 * k + 3 * 8 + rec,  # count, xy, pred -> list in; h, rho, c, record
        "force": 4 * k + 80 + 64,                         # list, own state -> a, edot, v, e, x, y
    }[kernel]


def profiled_traffic(precision: str, slot: str):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3
    counter summary (profiles/hbm_traffic.txt: scaled FETCH_SIZE and
    WRITE_SIZE of the clustered 4M case), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == precision and f[1] == slot:
            return int(float(f[2]) + float(f[3]))
    return None


def make_state(args):
    if args.workload == "batch":
        sts = ics.batch(args.batch_count or 512, seed=4, k=32)
        return {"name": f"batch-{len(sts)}", "batch": sts, "k": 32,
                "n": int(sum(st["n"] for st in sts))}
    if args.workload == "clustered":
        return ics.clustered(args.n or 4194304, seed=3, k=64)
    if args.workload == "blast":
        return ics.blast(args.n or 262144, seed=2, k=32)
    if args.workload == "uniform":
        return ics.uniform(args.n or 2200, seed=12345678, k=32)
    raise SystemExit(f"unknown workload {args.workload}")


def params_for(st, precision):
    return L.default_params(k=st["k"], precision=precision, mass=st["mass"], gamma=5 / 3,
                            dt_half=0.0, cfl=0.3, hor=(0, 1), vert=(0, 1),
                            visc_alpha=1.0, visc_beta=2.0, visc_eta2=0.01)


def make_sim(st, precision, graph=False, tune=False):
    if "batch" in st:
        sim = L.Batch([params_for(b, precision) for b in st["batch"]], st["batch"])
    else:
        sim = L.Sim(params_for(st, precision), st["x"], st["y"], st["vx"], st["vy"], st["e"])
    tuned = None
    if tune:
        ms, _ = L.autotune(sim)
        tuned = {"pass_ms": round(ms, 4), **L.get_tuning(sim)}
    if graph:
        L.set_graph(sim, True)
    return sim, tuned


def run_variant(st, precision, steps, warmup, timing_steps=3, graph=False, tune=False):
    sim, tuned = make_sim(st, precision, graph, tune)
    sim.step(warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)  # returns after the device finished
    t1 = time.perf_counter()
    dt = t1 - t0
    # per-kernel device times (HIP events on the library's stream)
    sim.set_timing(True)
    sim.reset_kernel_times()
    sim.step(timing_steps)
    ms, cnt = sim.kernel_times()
    sim.set_timing(False)
    if graph:
        L.set_graph(sim, True)
    retries, ties = sim.knn_stats()
    counters = sim.knn_counters()
    red = sim.reductions()
    sim.close()
    return {
        "seconds": dt,
        "ms_per_step": 1e3 * dt / steps,
        "value": st["n"] * steps / dt,
        "kernel_ms": {k: round(v, 4) for k, v in ms.items()},
        "knn_retries": retries,
        "knn_counters": counters,
        "dt_half_last": red[5],
        "autotune": tuned,
    }


def run_batch(steps=10, warmup=3):
    """The batched workload: 512 blast simulations, N log-uniform in [1024,
    65536] (seed 4), k = 32, stepped in one launch per kernel, grid and walk
    shape autotuned, steps replayed from a hipGraph; FP32C and plain FP64."""
    args = argparse.Namespace(workload="batch", batch_count=512, n=0)
    st = make_state(args)
    out = {"workload": f"batch-512 (blast ICs, N log-uniform in [1024, 65536], seed 4, k=32)",
           "n_particles": st["n"], "steps": steps, "warmup": warmup, "graph": True}
    for prec, code in (("fp32c", L.PREC_FP32C), ("fp64", L.PREC_FP64)):
        r = run_variant(st, code, steps, warmup, graph=True, tune=True)
        out[prec] = {"value": round(r["value"], 1), "ms_per_step": round(r["ms_per_step"], 4),
                     "kernel_ms": r["kernel_ms"], "autotune": r["autotune"]}
        print(f"# batch {prec}: {r}", file=sys.stderr, flush=True)
    out["speedup_fp32c_over_fp64"] = round(out["fp32c"]["value"] / out["fp64"]["value"], 3)
    return out


SPEED_TEST_STEPS = 20  # Step() calls of the reference's harness


def speed_test_params(precision=L.PREC_FP64):
    """MakeConfig() (sim/config-parser.go:131-149: both axes open, no walls,
    Monaghan, gamma 1.66666, mass 1, k = NN_SIZE = 32, viscosity 0.75 / 1.5 /
    0.01) with the harness's DeltaTHalf 0.02 and Acceleration (0, 0.2)
    (examples/speed-test/speed-test.go:25-27)."""
    return L.default_params(precision=precision, dt_half=0.02, accel=(0.0, 0.2))


def run_speed_test():
    """The reference's timing harness on the GPU: a fresh simulation of its
    100,000-particle input, SPEED_TEST_STEPS Step() calls from creation (the
    first with its extra force pass), as frames per second the way the
    harness prints them, then the steady rate of the next SPEED_TEST_STEPS."""
    st = ics.speed_test()
    out = {"workload": "the reference's speed-test harness (examples/speed-test/speed-test.go): "
                       "100,000 uniform particles, MakeConfig defaults (open axes), "
                       "DeltaTHalf 0.02, g = (0, 0.2), k = 32",
           "n_particles": st["n"], "steps": SPEED_TEST_STEPS}
    for name, prec in (("fp64", L.PREC_FP64), ("fp32c", L.PREC_FP32C)):
        sim = L.Sim(speed_test_params(prec), st["x"], st["y"], st["vx"], st["vy"], st["e"])
        sim.synchronize()
        t0 = time.perf_counter()
        sim.step(SPEED_TEST_STEPS)  # returns after the device finished
        t1 = time.perf_counter()
        sim.step(SPEED_TEST_STEPS)
        t2 = time.perf_counter()
        sim.close()
        out[name] = {"fps": round(SPEED_TEST_STEPS / (t1 - t0), 1),
                     "steady_particle_steps_per_s": round(st["n"] * SPEED_TEST_STEPS / (t2 - t1), 1)}
        print(f"# speed test {name}: {out[name]}", file=sys.stderr, flush=True)
    return out


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": avail, "os_cpu_count": os.cpu_count()}


def cpu_baseline(threads_all: int, c3_steps: int = 1):
    """The oracle's faithful restatement of the reference (Treebuild +
    BoundingSpheres + 9-image tree kNN + density + force + DKD, gcc -O2
    -ffp-contract=off) on this host: at the headline size (clustered 4M,
    k = 64) on all cores (OpenMP over particles in the kNN, density and force
    loops; the tree build stays serial, as in the reference), and one thread
    on a 32,768-particle sample of the same layout."""
    import oracle as O
    O.build()

    def rate(st, threads, nsteps):
        p = O.default_params(k=64, knn_mode=O.KNN_TREE, gamma=5 / 3, mass=st["mass"],
                             hor=(0, 1), vert=(0, 1), use_cfl=1, cfl=0.3,
                             visc_alpha=1.0, visc_beta=2.0)
        s = O.OracleSim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"], threads=threads)
        s.step(1)  # step-0 pass + first step, untimed
        t0 = time.perf_counter()
        s.step(nsteps)
        dt = time.perf_counter() - t0
        return st["n"] * nsteps / dt, dt / nsteps

    big = ics.clustered(4194304, seed=3, k=64)
    v_all, s_all = rate(big, threads_all, c3_steps)
    print(f"# cpu baseline: 4M on {threads_all} threads, {s_all:.1f} s per step",
          file=sys.stderr, flush=True)
    small = ics.clustered(32768, seed=3, k=64)
    v_one, s_one = rate(small, 1, 4)
    # the reference's speed-test harness input, one thread (the reference's
    # path is serial): its first Step() calls from creation, as it times them
    stt = ics.speed_test()
    ps = O.default_params(k=32, knn_mode=O.KNN_TREE, dt_half=0.02, accel=(0.0, 0.2),
                          use_cfl=0)
    s = O.OracleSim(ps, stt["x"], stt["y"], stt["vx"], stt["vy"], stt["e"], threads=1)
    nst = 3
    t0 = time.perf_counter()
    s.step(nst)
    st_fps = nst / (time.perf_counter() - t0)
    return {
        "value": v_all, "unit": "particle-steps/s", "cores": threads_all, "kind": "port",
        "sample": f"clustered 4,194,304 particles, k=64 (the headline workload): {c3_steps} "
                  "step(s) timed after the untimed step-0 pass and first step, "
                  f"{s_all:.1f} s per step; oracle restatement of the reference's tree "
                  "path, OpenMP over particles in kNN/density/force, tree build serial",
        **cpu_info(),
        "one_thread": {"value": v_one, "cores": 1, "seconds_per_step": round(s_one, 3),
                       "sample": "clustered layout at N=32768, k=64, 4 steps after step 0"},
        "speed_test": {"fps": round(st_fps, 3), "cores": 1, "steps": nst,
                       "sample": "the reference's speed-test harness input (100,000 uniform "
                                 "particles, MakeConfig defaults, DeltaTHalf 0.02, g = (0, 0.2)): "
                                 f"its first {nst} Step() calls, 1 thread"},
    }


def run_serve(njobs, precision, capacity, max_jobs=256, load=0.7):
    """Poisson arrivals of jobs (N log-uniform in [1024, 2^20], k in {32, 64},
    steps uniform in [10, 200]) into the continuous-batching server at `load`
    times its capacity measured on a burst of 64 jobs of the same mix."""
    import queue
    import threading

    make = ics.serve_ics(seed=5)
    prec = L.PREC_FP32C if precision == "fp32c" else L.PREC_FP64

    def submit(srv, i, n, k, steps):
        st = make(i, n, k)
        return srv.submit(params_for(st, prec), st, steps)

    # capacity: a burst of 64 jobs, submit to last completion
    srv = L.Server(capacity, max_jobs=max_jobs, kmax=64, precision=prec)
    cal = ics.serve_jobs(64, seed=6)
    ids = [submit(srv, 10**5 + i, n, k, s) for i, (n, k, s) in enumerate(cal)]
    srv.drain()
    cs = srv.stats()
    for j in ids:
        srv.release(j)
    srv.close()
    cap_rate = cs["particle_steps"] / cs["wall_s"]

    jobs = ics.serve_jobs(njobs, seed=5)
    mean_ps = float(np.mean([n * s for n, _, s in jobs]))
    lam = load * cap_rate / mean_ps
    arrivals = np.cumsum(np.random.default_rng(5).exponential(1.0 / lam, njobs))
    srv = L.Server(capacity, max_jobs=max_jobs, kmax=64, precision=prec)
    done = queue.Queue()
    lat = []
    errs = []

    def reaper():
        while True:
            jid = done.get()
            if jid is None:
                return
            try:
                srv.wait(jid)
                lat.append(srv.info(jid)["latency_ms"])
                srv.release(jid)
            except Exception as exc:  # noqa: BLE001 - reported below
                errs.append(str(exc))

    th = threading.Thread(target=reaper)
    th.start()
    t0 = time.perf_counter()
    for i, (n, k, s) in enumerate(jobs):
        st = make(i, n, k)
        p = params_for(st, prec)
        wait = t0 + arrivals[i] - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        done.put(srv.submit(p, st, s))
        if i % 100 == 99:
            print(f"# serve: {i + 1} jobs submitted at {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
    srv.drain()
    done.put(None)
    th.join()
    rs = srv.stats()
    srv.close()
    if errs:
        raise RuntimeError(f"serve: {len(errs)} jobs failed: {errs[:3]}")
    return {
        "jobs": njobs, "precision": precision, "pool_capacity": capacity, "max_jobs": max_jobs,
        "capacity_particle_steps_per_s": round(cap_rate, 1),
        "offered_load": load, "lambda_jobs_per_s": round(lam, 3),
        "mean_particle_steps_per_job": round(mean_ps, 1),
        "throughput_particle_steps_per_s": round(rs["particle_steps_per_s"], 1),
        "p50_ms": round(rs["p50_ms"], 2), "p99_ms": round(rs["p99_ms"], 2),
        "queue_wait_p50_ms": round(rs["queue_p50_ms"], 2),
        "queue_wait_p99_ms": round(rs["queue_p99_ms"], 2),
        "run_p50_ms": round(rs["run_p50_ms"], 2), "run_p99_ms": round(rs["run_p99_ms"], 2),
        "mean_latency_ms": round(float(np.mean(lat)), 2),
        "mean_occupancy": round(rs["occupancy"], 4), "ticks": int(rs["ticks"]),
        "tick_ms": round(rs["tick_ms"], 3), "tick_gpu_ms": round(rs["tick_gpu_ms"], 3),
        "wall_s": round(rs["wall_s"], 2),
        "ics": "clustered recipe (16 Gaussians sigma 0.02 holding 80%, 20% uniform, "
               "v ~ N(0, 0.1^2), e = 1): per-job window of one of 16 shuffled 2^20 templates "
               "with a per-job periodic shift",
    }


def error_report():
    """FP32C against plain FP64 on the same inputs: density and acceleration
    after the first force pass (max and rms relative error; acceleration
    relative to max |a|), neighbour sets, and total-energy drift
    E = sum m (e + v^2/2) over a run.  The plain FP64 form is the oracle's
    within 1e-12 (tests/test_gpu_parity.py), so it stands in for the oracle
    at sizes the CPU restatement cannot run quickly."""
    cases = [("uniform-2200", ics.uniform(2200, seed=12345678, k=32), 10),
             ("blast-262144", ics.blast(262144, seed=2, k=32), 100),
             ("clustered-4194304", ics.clustered(4194304, seed=3, k=64), 50)]
    out = {}
    for name, st, steps in cases:
        res = {}
        sims = {}
        for prec in ("fp64", "fp32c"):
            p = params_for(st, L.PREC_FP32C if prec == "fp32c" else L.PREC_FP64)
            s = L.Sim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"])
            s.set("vpx", st["vx"])
            s.set("vpy", st["vy"])
            s.set("epred", st["e"])
            s.calculate_forces()
            sims[prec] = {f: s.get(f) for f in ("rho", "ax", "ay", "h")}
            sims[prec]["ids"] = s.neighbours(with_d2=False)[0]
            s.close()
            s = L.Sim(p, st["x"], st["y"], st["vx"], st["vy"], st["e"])
            s.step(1)
            e0 = s.reductions()[2]
            s.step(steps - 1)
            e1 = s.reductions()[2]
            sims[prec]["drift"] = (e1 - e0) / abs(e0)
            sims[prec]["E_end"] = e1
            s.close()
        a, b = sims["fp32c"], sims["fp64"]
        rr = np.abs(a["rho"] - b["rho"]) / b["rho"]
        amag = np.sqrt(b["ax"] ** 2 + b["ay"] ** 2)
        ar = np.sqrt((a["ax"] - b["ax"]) ** 2 + (a["ay"] - b["ay"]) ** 2) / amag.max()
        res["neighbour_sets_equal"] = bool((a["ids"] == b["ids"]).all())
        res["h_equal"] = bool((a["h"] == b["h"]).all())
        res["rho_rel_max"] = float(rr.max())
        res["rho_rel_rms"] = float(np.sqrt(np.mean(rr ** 2)))
        res["accel_rel_max"] = float(ar.max())
        res["accel_rel_rms"] = float(np.sqrt(np.mean(ar ** 2)))
        res["steps"] = steps
        res["energy_drift_fp64"] = float(b["drift"])
        res["energy_drift_fp32c"] = float(a["drift"])
        res["energy_end_rel_diff"] = float(abs(a["E_end"] - b["E_end"]) / abs(b["E_end"]))
        out[name] = res
        print(f"# errors {name}: {res}", file=sys.stderr, flush=True)
    return out


def profile_step(precision: str, steps: int) -> None:
    """A fixed workload to run under rocprofv3 counters: the clustered 4M
    case (k = 64) created, one untimed step, then `steps` steps, every pass
    launching each kernel once."""
    st = ics.clustered(4194304, seed=3, k=64)
    sim = L.Sim(params_for(st, L.PREC_FP32C if precision == "fp32c" else L.PREC_FP64),
                st["x"], st["y"], st["vx"], st["vy"], st["e"])
    sim.step(1)
    sim.step(steps)
    print(f"stepped {steps} steps at N={st['n']} ({precision}); E = {sim.total_energy():.12g}",
          flush=True)
    sim.close()


def profile_copy(mib: int) -> None:
    """One device-to-device copy of known size: the reference that scales
    the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB (read {mib} MiB, wrote {mib} MiB)", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="clustered")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--precision", default="both", choices=["both", "fp32c", "fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="timed steps of the all-cores CPU baseline at 4M")
    ap.add_argument("--graph", action="store_true", help="replay steps from a hipGraph")
    ap.add_argument("--tune", action="store_true", help="autotune the grid before stepping")
    ap.add_argument("--batch-count", type=int, default=0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-speed-test", action="store_true",
                    help="skip the reference's speed-test harness workload")
    ap.add_argument("--speed-test-only", action="store_true",
                    help="run only the reference's speed-test harness workload (GPU) and print it")
    ap.add_argument("--errors", action="store_true",
                    help="print the FP32C error report (FP32C vs FP64) and exit")
    ap.add_argument("--jobs", type=int, default=1000, help="served jobs")
    ap.add_argument("--serve-precision", default="fp32c", choices=["fp32c", "fp64"])
    ap.add_argument("--serve-capacity", type=int, default=1 << 23, help="pool particle slots")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving workload and print its object")
    ap.add_argument("--profile-step", choices=["fp32c", "fp64"],
                    help="run the clustered 4M case for counter runs (no timing)")
    ap.add_argument("--profile-copy", type=int, metavar="MIB",
                    help="one device copy of MIB MiB (counter scaling)")
    args = ap.parse_args()
    if args.profile_step:
        profile_step(args.profile_step, args.steps)
        return
    if args.profile_copy:
        profile_copy(args.profile_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design, not a run)"}))
        sys.exit(2)
    if args.errors:
        print(json.dumps({"fp32c_error_report": error_report()}), flush=True)
        return
    if args.speed_test_only:
        print(json.dumps({"speed_test": run_speed_test()}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"serve": run_serve(args.jobs, args.serve_precision,
                                             args.serve_capacity)}), flush=True)
        return
    st = make_state(args)
    k = st["k"]
    res = {}
    if args.precision in ("both", "fp32c"):
        res["fp32c"] = run_variant(st, L.PREC_FP32C, args.steps, args.warmup,
                                   graph=args.graph, tune=args.tune)
        print(f"# fp32c: {res['fp32c']}", file=sys.stderr, flush=True)
    if args.precision in ("both", "fp64"):
        res["fp64"] = run_variant(st, L.PREC_FP64, args.steps, args.warmup,
                                  graph=args.graph, tune=args.tune)
        print(f"# fp64: {res['fp64']}", file=sys.stderr, flush=True)
    head = res.get("fp32c") or res["fp64"]
    # dominant single kernel and its roofline
    km = head["kernel_ms"]
    prec_name = "fp32c" if "fp32c" in res else "fp64"
    cands = ("knn_walk", "knn_finish", "force", "permute", "keys")
    dom = max(cands, key=lambda n: km[n])
    bytes_launch = algo_bytes(dom, k, prec_name) * st["n"]
    ach = bytes_launch / (km[dom] * 1e-3) / 1e9
    traffic = profiled_traffic(prec_name, dom) if args.workload == "clustered" else None
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algo_bytes_per_particle": algo_bytes(dom, k, prec_name), "avg_ms": km[dom]}
    df = ("knn_walk", "knn_finish", "force")
    t_df = km["knn_density"] + km["force"]
    b_df = sum(algo_bytes(n, k, prec_name) for n in df) * st["n"]
    roof["knn_density_force_frac"] = round(b_df / (t_df * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    line = {
        "metric": METRIC,
        "value": round(head["value"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if "fp32c" in res else "f64",
        "data": "synthetic",
        "config": {"workload": f"{st['name']} (k={k}, CFL 0.3, alpha=1, beta=2, periodic unit box)",
                   "n_particles": st["n"], "k": k,
                   "precision": "fp32c" if "fp32c" in res else "fp64"},
        "roofline": roof,
    }
    if "fp64" in res:
        line["fp64"] = {"value": round(res["fp64"]["value"], 1),
                        "ms_per_step": round(res["fp64"]["ms_per_step"], 4),
                        "kernel_ms": res["fp64"]["kernel_ms"]}
    if "fp32c" in res:
        line["fp32c_kernel_ms"] = res["fp32c"]["kernel_ms"]
    if "fp32c" in res and "fp64" in res:
        line["speedup_fp32c_over_fp64"] = round(res["fp32c"]["value"] / res["fp64"]["value"], 3)
    if not args.no_batch and args.workload == "clustered":
        line["batch"] = run_batch()
    if not args.no_speed_test and args.workload == "clustered":
        line["speed_test"] = run_speed_test()
    if not args.no_cpu:
        threads = min(16, cpu_info()["nproc"])
        line["cpu_baseline"] = cpu_baseline(threads, args.cpu_steps)
        if "speed_test" in line:
            cpu_fps = line["cpu_baseline"]["speed_test"]["fps"]
            line["speed_test"]["cpu_1thread_fps"] = cpu_fps
            line["speed_test"]["fp64_fps_over_cpu"] = round(line["speed_test"]["fp64"]["fps"] / cpu_fps, 1)
    if not args.no_serve and args.workload == "clustered":
        line["serve"] = run_serve(args.jobs, args.serve_precision, args.serve_capacity)
        print(f"# serve: {line['serve']}", file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
