#!/usr/bin/env python3
"""Benchmark: modal TL solves/s of the MI355X effective-sound-speed normal-mode path.

One step = one batch of (profile, frequency, azimuth) solves through the HIP library:
grid setup -> k_min/k_max -> all eigenvalues in [k_min^2, k_max^2] with perturbative
attenuation and psi(zs) psi(zr) -> 1-D transmission loss at 10000 ranges.  Profiles are
resident on the device before the timed region; outputs stay on the device.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

Prints one JSON line (see DESIGN.md for the fields).  The CPU baseline is the FP64 oracle
(scipy LAPACK bisection + inverse iteration, numpy TL sums) timed on a bounded sample of the
same workload on this host, single-threaded.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ncpaprop-release_amd")
sys.path.insert(0, PKG)

FREQ = 1.0
AZIMUTHS = [0.0, 45.0, 90.0, 135.0, 180.0, 225.0, 270.0, 315.0]
SEEDS = list(range(8))
NZ = 20000
NR = 10000
MAXRANGE_M = 1000e3
FLOPS_PER_STURM_STEP = 3.0     # one row of one sweep: sub, mul/div, sub
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def write_profiles(d):
    from ncpamodes import synth
    return [synth.write_profile(os.path.join(d, "P%d.dat" % s), seed=s) for s in SEEDS]


def cpu_baseline(n_problems=2):
    """FP64 oracle on a bounded sample of the workload (runs in a 1-thread subprocess)."""
    sys.path.insert(0, ROOT)
    from oracle import ess
    with tempfile.TemporaryDirectory() as d:
        paths = write_profiles(d)
        t = time.perf_counter()
        for i in range(n_problems):
            sol = ess.solve(paths[i % len(paths)], freq=FREQ, azimuth=AZIMUTHS[i % len(AZIMUTHS)],
                            Nz=NZ, turnoff_wkb=True)
            ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], MAXRANGE_M / NR, NR)
        dt = time.perf_counter() - t
    return {"value": n_problems / dt, "seconds": dt, "problems": n_problems}


def run_cpu_leg():
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        env[k] = "1"
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg"], env=env,
                         capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        sys.stderr.write(out.stderr)
        return None
    return json.loads(out.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-leg", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg:
        print(json.dumps(cpu_baseline()))
        return 0
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); the path shards "
              "by independent (profile, frequency, azimuth) solves, see DESIGN.md" % args.gpus,
              file=sys.stderr)
        return 2
    import numpy as np
    import ncpamodes

    tmp = tempfile.mkdtemp(prefix="ncpa_bench_")
    paths = write_profiles(tmp)
    ctx = ncpamodes.Context(0)
    pids = [ctx.read_profile(p)[0] for p in paths]
    grid = ctx.grid(pids[0], Nz=NZ)
    problems = [ncpamodes.problem(pid, FREQ, az) for pid in pids for az in AZIMUTHS]
    P = len(problems)
    bufs = [ctx.empty(P * NR) for _ in range(3)]
    dr = MAXRANGE_M / NR

    def step():
        m = ctx.solve(grid, problems)
        m.tl1d_device(dr, NR, False, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr)
        return m

    for _ in range(args.warmup):
        step().free()
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.reset_timing()
    t0 = time.perf_counter()
    last = None
    for k in range(args.steps):
        m = step()
        if k == args.steps - 1:
            last = m
        else:
            m.free()
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    modes_ms, modes_n = ctx.kernel_time("modes")
    tl_ms, _ = ctx.kernel_time("tl1d")
    iters = np.abs(last.iterations()).astype(np.float64)
    total_modes = last.total
    # one final sensitivity sweep per mode in addition to the counted sweeps
    steps_per_launch = float((iters + 1.0).sum()) * NZ
    modes_avg_s = modes_ms / max(modes_n, 1) / 1e3
    achieved_tflops = FLOPS_PER_STURM_STEP * steps_per_launch / modes_avg_s / 1e12
    last.free()
    ctx.set_timing(False)

    cpu = None if args.no_cpu else run_cpu_leg()
    value = P * args.steps / elapsed
    line = {
        "metric": "modal TL solves/sec at <=0.1 dB vs FP64 plain on GPU; p50/p99 request latency",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "ESS normal modes FP64 at 1 Hz, Nz=20000 over 0-150 km, all modes in "
                        "[k_min,k_max] + perturbative attenuation + 1-D TL at 10000 ranges "
                        "(0.1-1000 km); batch of 64 solves = 8 seeded profiles x 8 azimuths",
            "global_batch": P,
            "freq_hz": FREQ,
            "Nz": NZ,
            "n_ranges": NR,
            "modes_per_step": int(total_modes),
            "parallelism": "single GPU",
        },
        "stage_ms_per_step": {
            "modes": modes_ms / args.steps,
            "tl1d": tl_ms / args.steps,
        },
        "roofline": {
            "kernel": "k_modes",
            "bound": "mfma",
            "note": "FP64 VALU recurrence; peak is MI355X FP64 (vector = matrix) 78.6 TFLOP/s",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": None,
        },
        "cpu_baseline": None if cpu is None else {
            "value": cpu["value"],
            "unit": "solves/s",
            "cores": 1,
            "kind": "port",
            "sample": "%d of the benchmark's 1 Hz solves (eigen + perturbation + 1-D TL at "
                      "10000 ranges), FP64 oracle: scipy LAPACK stebz/stein + numpy, 1 thread, "
                      "%.1f s" % (cpu["problems"], cpu["seconds"]),
        },
    }
    print(json.dumps(line), flush=True)
    for b in bufs:
        b.free()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
