#!/usr/bin/env python3
"""Benchmark of the MI355X effective-sound-speed normal-mode path (ncpaprop Modess / ModBB).

Default workload (--workload field, the configuration the headline metric is quoted on):
one step = a batch of 64 (profile, frequency, azimuth) solves at 1 Hz (8 seeded synthetic
profiles x 8 azimuths), each through the HIP library: grid -> k_min/k_max -> every mode in
[k_min^2, k_max^2] with perturbative attenuation -> 1-D TL at 10000 ranges (0.1-1000 km) ->
eigenvector rows at 1500 depths -> 2-D field 1500 depths x 10000 ranges.  Eigenvalues,
phases (reduced mod 2 pi) and the 1-D TL are FP64; the field contraction runs in the fastest
precision whose TL stays within 0.1 dB of the FP64 contraction wherever TL > -140 dB, which
the run checks itself on every point of all 64 solves before timing (the FP64-plain rate is
reported beside it).  Profiles are resident on the device before the timed region; outputs stay there.
The steps are timed twice: on one device stream (the stage split and the roofline line come
from this region, each kernel alone on the GPU) and as a two-stage pipeline over two device
streams (step i's eigenpairs, 1-D TL, rows and field staging beside step i-1's contraction), which gives
the headline value (--schedule streams: independent batches on --streams streams instead).

--workload precision: the 2-D field at 2 Hz, 1500 x 10000, in every contraction precision
against FP64 (time, speed-up, max/RMS dB error).

--workload broadband: one step = a batched ModBB job: dispersion solves at 1024 frequencies
(f_min = f_step = 2/1024 Hz) x 8 azimuths on the seed-0 profile, then transfer functions at
16 receiver ranges (200-1000 km) and pulse synthesis (NFFT = max(4096, 4 Nfreq) = 4100, the
reference's rule) with a Gaussian-windowed 0.5 Hz source given as a spectrum.

--workload wmod: the field workload's 64-problem batch through the wide-angle (WMod) solver
(eigenpairs + attenuation + 1-D TL, FP64), the effective-sound-speed batch timed beside it.
--workload broadband-wmod: the broadband job with wide-angle dispersion (ModBB --method wmod).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload field|broadband|serving|precision|wmod|broadband-wmod]
                    [--schedule pipeline|streams] [--streams S] [--no-cpu]

The default run (field workload) then adds the rest of the metric: "latency" (p50/p99 of the
continuous-batching server at the production precision under Poisson arrivals at 100 and 500
requests/s), "broadband_job" (the batched ModBB job and its CPU comparison) and
"precision_2hz" (the --workload precision comparison, 10 launches per precision);
--headline-only skips them.  Before timing, every context that contracts runs the contraction
autotuner (ncpa_field_tune) on one solve of the batch ("contraction_tuning").

Prints one JSON line.  cpu_baseline (field and broadband): the oracle's C restatement of the
reference loops (oracle/ess_port.c, gcc -O2, FP64) on the host cores that go with one GPU
(16), one solve per thread, on a bounded sample of the workload scaled to its unit; the
wide-angle workload times the Python oracle in worker processes instead.
"""
import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ncpaprop-release_amd")
sys.path.insert(0, PKG)

METRIC = "modal TL solves/sec at ≤0.1 dB vs FP64 plain on GPU; p50/p99 request latency"
FREQ = 1.0
AZIMUTHS = [0.0, 45.0, 90.0, 135.0, 180.0, 225.0, 270.0, 315.0]
SEEDS = list(range(8))
NZ = 20000
NR = 10000
MAXRANGE_M = 1000e3
FIELD_ROWS, FIELD_STRIDE = 1500, 13        # 1500 depths, 0 - 146 km AGL
FLOPS_PER_STURM_STEP = 3.0                 # one row of one sweep
WIDE_CPU_STRIDE = 8                        # wide-angle CPU baseline: every 8th eigenvalue branch
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}
PREC_NAMES = {0: "f64", 1: "f32", 2: "bf16x3", 3: "bf16x6", 4: "fp16x3", 5: "fp16x4", 6: "bf16x6f",
              7: "fp16x4f"}
PREC_PEAK = {0: "f64", 1: "f32", 2: "bf16", 3: "bf16", 4: "bf16", 5: "bf16", 6: "bf16", 7: "bf16"}   # f16 = bf16
# split-precision kernel (parts, products, FP16, FP64 fold) per precision
PREC_KERNEL = {2: "k_gemm_split<2, 3, false>", 3: "k_gemm_split<3, 6, false>",
               4: "k_gemm_split<2, 3, true>", 5: "k_gemm_split<2, 4, true>",
               6: "k_gemm_fold<3, 6, false>", 7: "k_gemm_fold<2, 4, true>"}
PRODUCTS = {0: 1, 1: 1, 2: 3, 3: 6, 4: 3, 5: 4, 6: 6, 7: 4}
# candidate field precisions, fastest first
CANDIDATES = [2, 4, 5, 7, 3, 6, 1]
TL_GATE_DB = 0.1

BB_NF = 1024
BB_FSTEP = 2.0 / 1024
BB_RANGES_KM = [200.0 + i * (800.0 / 15.0) for i in range(16)]
BB_MAXCEL = 340.0


def write_profiles(d, seeds=SEEDS):
    from ncpamodes import synth
    return [synth.write_profile(os.path.join(d, "P%d.dat" % s), seed=s) for s in seeds]


def bb_source_spectrum(f):
    """Gaussian-windowed 0.5 Hz source pulse s(t) = sin(2 pi 0.5 (t - 10)) exp(-((t - 10)/2)^2),
    its spectrum S(f) = int s(t) e^{+2 pi i f t} dt in closed form (the --source spectrum
    convention of BroadbandPropagator.cpp:313-325)."""
    import numpy as np
    f0, tc, sg = 0.5, 10.0, 2.0
    # sin(2 pi f0 (t - tc)) = (e^{i a} - e^{-i a}) / 2i; Gaussian FT: sg sqrt(pi) e^{-(pi sg nu)^2}
    g = lambda nu: sg * math.sqrt(math.pi) * np.exp(-(math.pi * sg * nu) ** 2)
    S = (g(f + f0) - g(f - f0)) / 2j * np.exp(2j * math.pi * f * tc)
    return S


# ------------------------------------------------------------------------------------------
# CPU baseline (oracle) -- worker processes, one thread each
# ------------------------------------------------------------------------------------------
def cpu_worker(spec):
    sys.path.insert(0, ROOT)
    import numpy as np
    from oracle import ess
    out = {}
    if spec["workload"] == "field":
        t = time.perf_counter()
        sol = ess.solve(spec["path"], freq=FREQ, azimuth=spec["az"], Nz=NZ, turnoff_wkb=True)
        t_eig = time.perf_counter() - t
        t = time.perf_counter()
        ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], MAXRANGE_M / NR, NR)
        t_tl = time.perf_counter() - t
        rows = np.arange(FIELD_ROWS) * FIELD_STRIDE
        nr_s = spec["field_ranges"]
        t = time.perf_counter()
        ess.tl2d(sol.k_pert, sol.psi[rows], sol.psi[sol.i_s], MAXRANGE_M / NR, nr_s)
        t_f = time.perf_counter() - t
        out["solve_s"] = t_eig + t_tl + t_f * (NR / nr_s)
        out["parts"] = [t_eig, t_tl, t_f]
    elif spec["workload"] == "wmod":
        # a wide-angle solve on a sample of its eigenvalue branches (every WIDE_CPU_STRIDE-th),
        # the eigen time scaled by the branch ratio; trace, perturbation and TL timed in full
        from oracle import atmos, wmod
        t = time.perf_counter()
        grid = wmod.WGrid(atmos.AtmosphereNodes(atmos.read_profile(spec["path"])), NZ, 150.0, 0.0)
        grid.set_attenuation(FREQ)
        kd, md, cd, _, _, kmin, kmax, _ = wmod.modal_trace(grid, FREQ, spec["az"], 0.0, 0.0, True)
        t_tr = time.perf_counter() - t
        t = time.perf_counter()
        k, v = wmod.eigen(kd, md, cd, grid.dz_m, kmin, kmax, stride=WIDE_CPU_STRIDE)
        t_eig = (time.perf_counter() - t) * WIDE_CPU_STRIDE
        t = time.perf_counter()
        kp = ess.perturb(k, v, wmod.perturb_weights(grid, FREQ))
        kp = np.repeat(kp, WIDE_CPU_STRIDE)
        ps = np.repeat(v[0], WIDE_CPU_STRIDE)
        ess.tl1d(kp, ps, ps, MAXRANGE_M / NR, NR)
        t_tl = time.perf_counter() - t
        out["solve_s"] = t_tr + t_eig + t_tl
        out["parts"] = [t_tr, t_eig, t_tl]
    else:
        from oracle import broadband as obb
        t = time.perf_counter()
        sol = ess.solve(spec["path"], freq=spec["freq"], azimuth=spec["az"], Nz=NZ, turnoff_wkb=True)
        kp, ps, pr = sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r]
        t_eig = time.perf_counter() - t
        # compute_modal_sum terms of this frequency at the 16 ranges
        t = time.perf_counter()
        for r in BB_RANGES_KM:
            np.sum(ps * pr * np.exp(1j * r * 1e3 * kp) / np.sqrt(kp))
        t_h = time.perf_counter() - t
        out["solve_s"] = t_eig + t_h
        out["parts"] = [t_eig, t_h]
        if spec.get("fft"):
            nf = BB_NF + 1
            f = np.arange(nf) * BB_FSTEP
            nfft = 4 * nf
            t = time.perf_counter()
            H = np.ones(nf, dtype=complex)
            for _ in range(len(BB_RANGES_KM) * len(AZIMUTHS)):
                obb.pulse_prop(H, H, f, 1.0, nfft)
            out["fft_s"] = time.perf_counter() - t
    print(json.dumps(out), flush=True)


def cpu_baseline(workload, n_workers):
    """Runs n_workers oracle processes at once (one thread each); returns the aggregate rate."""
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        env[k] = "1"
    tmp = tempfile.mkdtemp(prefix="ncpa_cpu_")
    specs = []
    if workload in ("field", "wmod"):
        paths = write_profiles(tmp)
        for i in range(n_workers):
            specs.append({"workload": workload, "path": paths[i % len(paths)],
                          "az": AZIMUTHS[(i * 3) % len(AZIMUTHS)], "field_ranges": 100})
    else:
        path = write_profiles(tmp, [0])[0]
        # frequencies spread uniformly over the band so the mean cost matches the band's
        idx = [int((i + 0.5) * BB_NF / n_workers) + 1 for i in range(n_workers)]
        for i, fi in enumerate(idx):
            specs.append({"workload": "broadband", "path": path, "freq": fi * BB_FSTEP,
                          "az": AZIMUTHS[i % len(AZIMUTHS)], "fft": i == 0})
    t = time.perf_counter()
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", json.dumps(s)],
                              env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
             for s in specs]
    res = []
    for p in procs:
        out, err = p.communicate(timeout=900)
        if p.returncode != 0:
            sys.stderr.write(err)
            return None
        res.append(json.loads(out.strip().splitlines()[-1]))
    wall = time.perf_counter() - t
    mean_solve = sum(r["solve_s"] for r in res) / len(res)
    out = {"workers": n_workers, "wall_s": wall, "mean_solve_s": mean_solve,
           "rate": n_workers / mean_solve, "parts": [sum(x) / len(res) for x in zip(*[r["parts"] for r in res])]}
    if workload == "broadband":
        out["fft_s"] = next(r["fft_s"] for r in res if "fft_s" in r)
    return out


HOST_CPU_SHARE = 16   # host cores that go with one GPU on the GPU boxes (their stated CPU share)


def cpu_cores():
    """Threads for the CPU baseline: every core this process may use, capped at the host CPU
    share that goes with one GPU on the GPU boxes (nproc there counts the whole host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(HOST_CPU_SHARE, n))


def host_info():
    """nproc, the affinity count and the CPU model of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"nproc": os.cpu_count(), "affinity": aff, "cpu_model": model}


# ------------------------------------------------------------------------------------------
# CPU baseline (the oracle's C restatement of the reference loops, one job per thread)
# ------------------------------------------------------------------------------------------
PORT_FIELD_RANGES = 4     # 2-D field ranges timed per solve (of NR; scaled by NR / this)


def port_baseline(workload, n_threads):
    """Times oracle/_build/ess_port (FP64 C, gcc -O2: Sturm bisection + inverse iteration,
    doPerturb, getTLoss1D / getTLoss2D with a complex exp and sqrt per term, as the reference
    loops) on n_threads threads, one job per thread; returns the aggregate rate and the sample."""
    port = os.path.join(ROOT, "oracle", "_build", "ess_port")
    if not os.path.exists(port):
        sys.stderr.write("bench.py: %s is not built (run __graft_entry__.build())\n" % port)
        return None
    tmp = tempfile.mkdtemp(prefix="ncpa_cpu_")
    if workload == "field":
        paths = write_profiles(tmp)
        jobs = ["%s:%g:%g" % (paths[i % len(paths)], AZIMUTHS[(i * 3) % len(AZIMUTHS)], FREQ)
                for i in range(n_threads)]
        opts = ["--Nz", str(NZ), "--nr", str(NR), "--dr", str(MAXRANGE_M / NR), "--rows", str(FIELD_ROWS),
                "--stride", str(FIELD_STRIDE), "--field_ranges", str(PORT_FIELD_RANGES)]
    else:
        path = write_profiles(tmp, [0])[0]
        # frequencies spread uniformly over the band so the mean cost matches the band's
        idx = [int((i + 0.5) * BB_NF / n_threads) + 1 for i in range(n_threads)]
        jobs = ["%s:%g:%.17g" % (path, AZIMUTHS[i % len(AZIMUTHS)], fi * BB_FSTEP) for i, fi in enumerate(idx)]
        # compute_modal_sum at the 16 receiver ranges: one complex exp and sqrt per (mode,
        # range), the per-term work of the 1-D TL loop at 16 ranges
        opts = ["--Nz", str(NZ), "--nr", str(len(BB_RANGES_KM)), "--dr", "50000"]
    cmd = [port, "bench", "--threads", str(n_threads)] + opts
    for j in jobs:
        cmd += ["--job", j]
    t = time.perf_counter()
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    wall = time.perf_counter() - t
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        return None
    out = json.loads(r.stdout)
    per = []
    for j in out["jobs"]:
        setup, eig, tl1d, tl2d = j["times"]
        per.append(setup + eig + tl1d + (tl2d * NR / PORT_FIELD_RANGES if workload == "field" else 0.0))
    mean = sum(per) / len(per)
    parts = [sum(j["times"][k] for j in out["jobs"]) / len(per) for k in range(4)]
    return {"workers": n_threads, "wall_s": wall, "mean_solve_s": mean, "rate": n_threads / mean,
            "parts": parts, "jobs": len(per), "host": host_info()}


# ------------------------------------------------------------------------------------------
# GPU legs
# ------------------------------------------------------------------------------------------
def field_gate(ctx, grid, problems):
    """Max / RMS |TL_prec - TL_fp64| [dB] over TL > -140 dB at every point of every problem of
    the step, for every candidate precision (compared on the device, untimed); returns the
    errors and the first candidate within TL_GATE_DB (0 = FP64 if none is)."""
    m = ctx.solve(grid, problems)
    rows = m.rows_device(0, FIELD_STRIDE, FIELD_ROWS)
    dr = MAXRANGE_M / NR
    o64 = m.field2d_device(rows, FIELD_ROWS, dr, NR, 0)
    n = len(problems) * FIELD_ROWS * NR
    errs, chosen = {}, 0
    for prec in CANDIDATES:
        oc = m.field2d_device(rows, FIELD_ROWS, dr, NR, prec)
        mx, rms, cnt = ctx.field_compare(o64, oc, prec, n)
        oc.free()
        errs[prec] = (mx, rms, cnt)
        if not chosen and mx <= TL_GATE_DB:
            chosen = prec
    for x in (o64, rows):
        x.free()
    m.free()
    return errs, chosen


def nonzero_chunks(psi, counts):
    """Chunks (problem, 128-row tile, 32-mode k-step) of the field's Psi operand [FIELD_ROWS x
    modes] that hold a nonzero (-0 counts as zero): k_gemm_fold skips the others."""
    import numpy as np
    a = psi.reshape(FIELD_ROWS, -1)
    mp = -(-FIELD_ROWS // 128) * 128
    n, off = 0, 0
    for k in counts:
        k = int(k)
        kp = -(-k // 32) * 32
        nz = np.zeros((mp, kp), dtype=bool)
        nz[:FIELD_ROWS, :k] = a[:, off:off + k] != 0.0
        off += k
        n += int(nz.reshape(mp // 128, 128, kp // 32, 32).any(axis=(1, 3)).sum())
    return n


def issued_mfma_flops(prec, counts, kernel="", psi=None):
    """MFMA flops one split contraction issues: 2 x (rows padded to 128) x (2 ranges padded to
    the block width: 64 for k_gemm_fold, else 128) x (sum of modes padded to 32) x products,
    and for k_gemm_fold only over the chunks of Psi that hold a nonzero (psi: its host copy);
    None for the FP64 / FP32 forms."""
    if prec not in PREC_KERNEL:
        return None
    mp = -(-FIELD_ROWS // 128) * 128
    bw = 64 if kernel.startswith("k_gemm_fold<") else 128
    np_ = -(-2 * NR // bw) * bw
    if bw == 64 and psi is not None:
        return 2.0 * 128 * np_ * 32 * nonzero_chunks(psi, counts) * PRODUCTS[prec]
    kp = sum(-(-int(k) // 32) * 32 for k in counts)
    return 2.0 * mp * np_ * kp * PRODUCTS[prec]


def run_field(lanes, args, prec):
    """Timed loop of the config-1 step at one field precision; returns a timing dict.

    lanes: list of (ctx, problems, grid), one per device stream.  Steps are dealt round-robin
    to the lanes, each driven by its own host thread, so independent batches overlap on the
    GPU (an eigenvalue stage that leaves SIMDs idle beside another batch's contraction).
    Kernel times are HIP events on each lane's stream over the same timed region."""
    import threading
    import numpy as np
    dr = MAXRANGE_M / NR
    lane_state = []
    for ctx, problems, grid in lanes:
        P = len(problems)
        lane_state.append({"tl": [ctx.empty(P * NR) for _ in range(3)]})

    def step(li):
        ctx, problems, grid = lanes[li]
        st = lane_state[li]
        P = len(problems)
        m = ctx.solve(grid, problems)
        m.tl1d_device(dr, NR, False, st["tl"][0].ptr, st["tl"][1].ptr, st["tl"][2].ptr)
        if "rows" not in st:
            st["rows"] = ctx.empty(FIELD_ROWS * max(m.total, 1))
            st["out"] = ctx.empty(P * FIELD_ROWS * 2 * NR, np.float64 if prec == 0 else np.float32)
        ctx.check(ctx.lib.ncpa_modes_rows(ctx.h, m.h, 0, FIELD_STRIDE, FIELD_ROWS,
                                          ctypes.c_void_p(st["rows"].ptr)))
        m.field2d_device(st["rows"], FIELD_ROWS, dr, NR, prec, out=st["out"])
        return m

    def run_lane(li, n_steps, keep):
        last = None
        for k in range(n_steps):
            m = step(li)
            if keep and k == n_steps - 1:
                last = m
            else:
                m.free()
        lanes[li][0].synchronize()
        lane_state[li]["last"] = last

    L = len(lanes)
    for li in range(L):
        for _ in range(args.warmup):
            step(li).free()
        lanes[li][0].synchronize()
    for ctx, _, _ in lanes:
        ctx.set_timing(True)
        ctx.reset_timing()
    counts = [args.steps // L + (1 if li < args.steps % L else 0) for li in range(L)]
    threads = [threading.Thread(target=run_lane, args=(li, counts[li], li == 0)) for li in range(L) if counts[li]]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    elapsed = time.perf_counter() - t0
    fams = {}
    for name in ("setup", "interval", "modes", "tl1d", "rows", "emat", "field2d"):
        tot, n = 0.0, 0
        for ctx, _, _ in lanes:
            ms, c = ctx.kernel_time(name)
            tot += ms
            n += c
        fams[name] = (tot, n)
    for ctx, _, _ in lanes:
        ctx.set_timing(False)
    last = lane_state[0]["last"]
    res = {"elapsed": elapsed, "fams": fams, "total_modes": int(last.total),
           # rows each mode's search swept (sweeps it used x rows below its block's evanescent cut)
           "sturm_steps": float(last.sweep_rows().astype(np.float64).sum()),
           "field_flops": 4.0 * FIELD_ROWS * last.total * NR, "streams": L,
           "field_mfma_flops": issued_mfma_flops(prec, last.counts,
                                                 lanes[0][0].field_variant_info(prec, lanes[0][0].field_variant(prec))[0],
                                                 lane_state[0]["rows"].numpy() if prec in PREC_KERNEL else None)}
    last.free()
    for st in lane_state:
        for b in st["tl"] + [st.get("rows"), st.get("out")]:
            if b is not None:
                b.free()
    return res


def tune_contraction(ctx, contractors, problems, grid, precs):
    """Contraction autotuner (ncpa_field_tune), untimed: on one solve of the step's batch,
    every launch variant of each precision is timed on each context that will contract, and
    that context keeps the fastest.  Returns {precision name: {"ms": [...], "best": i}} for the
    first context (the variants agree to the bit, so the choice changes speed only)."""
    dr = MAXRANGE_M / NR
    m = ctx.solve(grid, problems)
    rows = ctx.empty(FIELD_ROWS * max(m.total, 1))
    ctx.check(ctx.lib.ncpa_modes_rows(ctx.h, m.h, 0, FIELD_STRIDE, FIELD_ROWS, ctypes.c_void_p(rows.ptr)))
    report = {}
    for prec in precs:
        stage = m.field2d_stage(rows, FIELD_ROWS, dr, NR, prec)
        for i, c in enumerate(contractors):
            if c is not ctx:
                ready = ctx.record()
                c.wait(ready)
                ready.free()
            best, ms = stage.tune(c, reps=3)
            if i == 0:
                report[PREC_NAMES[prec]] = {"ms": [round(t, 3) for t in ms], "best": best}
        stage.free()
    ctx.synchronize()
    rows.free()
    m.free()
    return report


def run_field_pipelined(solver, contractor, problems, grid, args, prec, depth=2):
    """Timed loop of the config-1 step as a two-stage pipeline over two device streams, one
    host thread: the solver context runs step i's eigenpairs, 1-D TL, eigenvector rows and the
    field's operand staging while the contractor context contracts step i-1's field, so the
    contraction (MFMA) and the latency-bound eigenvalue sweeps (FP64 VALU) of consecutive
    batches share the GPU.  Stream ordering by events: a contraction waits for its staging;
    the solver reuses a row buffer and frees a step's modes and staged operands only after
    the contraction that read them (the solver runs at most `depth` steps ahead).  Returns a
    timing dict like run_field."""
    import numpy as np
    dr = MAXRANGE_M / NR
    P = len(problems)
    tl = [solver.empty(P * NR) for _ in range(3)]
    out = contractor.empty(P * FIELD_ROWS * 2 * NR, np.float64 if prec == 0 else np.float32)
    rows = [None] * (depth + 1)
    inflight = []            # (modes, staged operands, event after the contraction)

    def retire(keep):
        while len(inflight) > keep:
            m, stage, ev = inflight.pop(0)
            solver.wait(ev)          # the solver's later work (buffer reuse) follows the contraction
            stage.free()
            m.free()
            ev.free()

    def step(i):
        m = solver.solve(grid, problems)
        m.tl1d_device(dr, NR, False, tl[0].ptr, tl[1].ptr, tl[2].ptr)
        slot = i % (depth + 1)
        if rows[slot] is None:
            rows[slot] = solver.empty(FIELD_ROWS * max(m.total, 1))
        solver.check(solver.lib.ncpa_modes_rows(solver.h, m.h, 0, FIELD_STRIDE, FIELD_ROWS,
                                                ctypes.c_void_p(rows[slot].ptr)))
        stage = m.field2d_stage(rows[slot], FIELD_ROWS, dr, NR, prec)
        ready = solver.record()
        contractor.wait(ready)
        ready.free()
        stage.contract(contractor, out)
        inflight.append((m, stage, contractor.record()))
        retire(depth)

    # untimed warm-up: at least depth + 1 steps, so that every rotating row buffer and every
    # in-flight set of staged operands has been allocated before the timed region
    n_warm = max(args.warmup, depth + 1)
    for i in range(n_warm):
        step(i)
    retire(0)
    solver.synchronize()
    contractor.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(n_warm + i)
    retire(0)
    solver.synchronize()
    contractor.synchronize()
    elapsed = time.perf_counter() - t0
    for b in tl + [out] + [r for r in rows if r is not None]:
        b.free()
    return {"elapsed": elapsed, "streams": 2}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (copy-kernel calibrated, column "HBM bytes / launch"); None when that file has no
    line for it."""
    try:
        col = None
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if cells and cells[0] == "kernel":
                    col = cells.index("HBM bytes / launch") if "HBM bytes / launch" in cells else None
                elif col is not None and len(cells) > col and cells[0] == kernel:
                    return float(cells[col])
    except (OSError, ValueError):
        return None
    return None


def copy_calibration(ctx, mib):
    """One launch of the plain copy kernel over `mib` MiB (reference traffic for the counters)."""
    n = mib * (1 << 20)
    a, b = ctx.empty(n // 8), ctx.empty(n // 8)
    ctx.check(ctx.lib.ncpa_dev_copy(ctx.h, ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr), n))
    ctx.synchronize()
    a.free()
    b.free()


def field_main(args, cpu):
    import numpy as np
    import ncpamodes
    tmp = tempfile.mkdtemp(prefix="ncpa_bench_")
    paths = write_profiles(tmp)
    ctx = ncpamodes.Context(0)
    ctx.set_modes_form(args.modes_form)
    if args.copy_calibration:
        copy_calibration(ctx, args.copy_calibration)
    pids = [ctx.read_profile(p)[0] for p in paths]
    grid = ctx.grid(pids[0], Nz=NZ)
    problems = [ncpamodes.problem(pid, FREQ, az) for pid in pids for az in AZIMUTHS]
    P = len(problems)
    # precision gate over every point of all 64 solves, before any timing
    errs, chosen = field_gate(ctx, grid, problems)
    lanes = [(ctx, problems, grid)]
    n_lanes = args.streams if args.schedule == "streams" else 1
    for _ in range(n_lanes - 1):
        c2 = ncpamodes.Context(0)
        c2.set_modes_form(args.modes_form)
        p2 = [c2.read_profile(p)[0] for p in paths]
        g2 = c2.grid(p2[0], Nz=NZ)
        lanes.append((c2, [ncpamodes.problem(pid, FREQ, az) for pid in p2 for az in AZIMUTHS], g2))
    # one stream: the stage split and the roofline (each kernel alone on the GPU); then the
    # headline throughput: the two-stage pipeline (default) or independent batches in flight
    # on several streams
    contractor = ncpamodes.Context(0) if args.schedule == "pipeline" else None
    tuned = tune_contraction(ctx, [ctx] + [c for c, _, _ in lanes[1:]] + ([contractor] if contractor else []),
                             problems, grid, sorted({chosen, 0}))
    # the launch variant each precision's contraction runs with on the single-stream context
    kvariant = {prec: ctx.field_variant_info(prec, ctx.field_variant(prec)) for prec in {chosen, 0}}
    res = run_field(lanes[:1], args, chosen)
    plain_multi = None
    if args.schedule == "pipeline":
        multi = run_field_pipelined(ctx, contractor, problems, grid, args, chosen)
        # the FP64 contraction in the same two-stage schedule: the precision's own speed-up
        plain_multi = run_field_pipelined(ctx, contractor, problems, grid, args, 0) if chosen != 0 else multi
        contractor.close()
    else:
        multi = run_field(lanes, args, chosen) if len(lanes) > 1 else res
    plain = run_field(lanes[:1], args, 0) if chosen != 0 else res
    for c, _, _ in lanes[1:]:
        c.close()
    ctx.close()

    value = P * args.steps / multi["elapsed"]
    fams = res["fams"]
    per_step = {k: v[0] / args.steps for k, v in fams.items()}
    # dominant kernel family of the measured step
    dom = max(("modes", "field2d"), key=lambda k: fams[k][0])
    if dom == "field2d":
        avg_s = fams["field2d"][0] / max(fams["field2d"][1], 1) / 1e3
        achieved = res["field_flops"] / avg_s / 1e12
        peak = PEAK_TFLOPS[PREC_PEAK[chosen]]
        kname, krg = kvariant[chosen]
        roof = {"kernel": "%s (2-D field contraction, %s; row groups per XCD class %d)" % (kname, PREC_NAMES[chosen], krg),
                "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": measured_traffic("%s rg %d" % (kname, krg)),
                "algorithmic": "4 x depths x modes x ranges real flops per problem (real x complex)"}
        if res.get("field_mfma_flops"):
            issued = res["field_mfma_flops"] / avg_s / 1e12
            roof["issued_mfma"] = {
                "tflops": issued, "frac": issued / peak,
                "note": "MFMA work the kernel issues: %d part products per algorithmic product, "
                        "padded shapes, k-steps whose Psi chunk is all zero skipped by k_gemm_fold; "
                        "SQ_INSTS_MFMA x 16384 agrees (profiles/field_mfma_counters.md)" % PRODUCTS[chosen]}
    else:
        avg_s = fams["modes"][0] / max(fams["modes"][1], 1) / 1e3
        achieved = FLOPS_PER_STURM_STEP * res["sturm_steps"] / avg_s / 1e12
        roof = {"kernel": "k_modes<2>", "bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["f64"],
                "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["f64"], "traffic": measured_traffic("k_modes<2, false>"),
                "algorithmic": "3 FP64 flops per Sturm step (row x sweep); FP64 VALU, MI355X FP64 peak"}
    modes_avg_s = fams["modes"][0] / max(fams["modes"][1], 1) / 1e3
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": multi["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+%s" % PREC_NAMES[chosen] if chosen else "f64",
        "data": "synthetic",
        "config": {
            "workload": "ESS normal modes at 1 Hz, Nz=20000 over 0-150 km: all modes in [k_min,k_max] "
                        "(FP64 Sturm bisection/Newton) + perturbative attenuation + 1-D TL at 10000 "
                        "ranges (0.1-1000 km) + 2-D field 1500 depths x 10000 ranges; batch of 64 "
                        "solves = 8 seeded profiles x 8 azimuths",
            "global_batch": P,
            "freq_hz": FREQ,
            "Nz": NZ,
            "n_ranges": NR,
            "field_depths": FIELD_ROWS,
            "modes_per_step": res["total_modes"],
            "field_precision": PREC_NAMES[chosen],
            "eigen_sweeps": "diagonal via scalar loads" if args.modes_form else "diagonal via LDS ring",
            "device_streams": multi["streams"],
            "schedule": ("two-stage pipeline: eigenpairs, 1-D TL, rows and field operand staging of step i "
                         "on one stream while step i-1's field contraction runs on the other"
                         if args.schedule == "pipeline"
                         else "independent batches on %d streams" % multi["streams"]),
            "parallelism": "single GPU",
        },
        "tl_check_db": {PREC_NAMES[p]: {"max": e[0], "rms": e[1], "points": e[2]} for p, e in errs.items()},
        "tl_check_scope": "all %d solves of the step, every field point with TL_fp64 > -140 dB" % P,
        "fp64_plain": {"value": P * args.steps / plain["elapsed"], "unit": "solves/s",
                       "ms_per_step": plain["elapsed"] / args.steps * 1e3,
                       "note": "FP64 contraction, one stream (compare single_stream)"},
        "fp64_plain_pipelined": None if plain_multi is None else {
            "value": P * args.steps / plain_multi["elapsed"], "unit": "solves/s",
            "ms_per_step": plain_multi["elapsed"] / args.steps * 1e3,
            "note": "FP64 contraction in the headline's two-stage schedule (compare value)"},
        "single_stream": {"value": P * args.steps / res["elapsed"], "unit": "solves/s",
                          "ms_per_step": res["elapsed"] / args.steps * 1e3,
                          "note": "stage split and roofline are measured in this one-stream region"},
        "stage_ms_per_step": per_step,
        "roofline": roof,
        "contraction_tuning": {"note": "ncpa_field_tune on one solve of the batch before timing: ms per launch "
                                       "variant (block tile, LDS stages, row tiles per XCD); the variants agree "
                                       "to the bit", **tuned},
        "k_modes": {"avg_ms": modes_avg_s * 1e3,
                    "tflops": FLOPS_PER_STURM_STEP * res["sturm_steps"] / modes_avg_s / 1e12},
        "cpu_baseline": None if cpu is None else {
            "value": cpu["rate"], "unit": "solves/s", "cores": cpu["workers"], "kind": "port",
            "sample": "the oracle's C restatement of the reference loops (gcc -O2, FP64: Sturm bisection + "
                      "inverse iteration, doPerturb, getTLoss1D at all 10000 ranges, getTLoss2D with a complex "
                      "exp and sqrt per (range, depth, mode) term on 1500 depths x %d ranges, scaled x%d to "
                      "10000 ranges) on %d threads, one of the benchmark's 1 Hz solves per thread; mean "
                      "%.1f s per solve (eigen %.2f s, 1-D TL %.2f s, 2-D field %.0f s), %.1f s wall"
                      % (PORT_FIELD_RANGES, NR // PORT_FIELD_RANGES, cpu["workers"], cpu["mean_solve_s"],
                         cpu["parts"][1], cpu["parts"][2], cpu["parts"][3] * NR / PORT_FIELD_RANGES,
                         cpu["wall_s"]),
            "host": cpu["host"],
        },
    }
    return line


def broadband_main(args, cpu):
    import numpy as np
    import ncpamodes
    from ncpamodes import broadband as bb
    tmp = tempfile.mkdtemp(prefix="ncpa_bench_")
    path = write_profiles(tmp, [0])[0]
    ctx = ncpamodes.Context(0)
    if args.copy_calibration:
        copy_calibration(ctx, args.copy_calibration)
    pid = ctx.read_profile(path)[0]
    grid = ctx.grid(pid, Nz=NZ)
    f = np.arange(BB_NF + 1) * BB_FSTEP
    problems = [ncpamodes.problem(pid, f[i], az) for az in AZIMUTHS for i in range(1, BB_NF + 1)]
    ranges = np.array(BB_RANGES_KM) * 1e3
    nfft = bb.default_nfft(4096, f.size)
    S, _ = bb.source(ctx, "spectrum", f, nfft, -1.0, np.stack([f, bb_source_spectrum(f).real,
                                                                bb_source_spectrum(f).imag], axis=1))
    nset = len(AZIMUTHS)
    # the job as one graph: batched dispersion (eigenvalue launch form and multisection depth
    # autotuned), transfer functions, pulse synthesis
    t = time.perf_counter()
    plan = bb.BroadbandPlan(ctx, grid, problems, nset, f, S, ranges, np.ones(nset), BB_MAXCEL, nfft, tune=True)
    plan_s = time.perf_counter() - t
    tuning = plan.tuning()
    # one eager step (same tuned eigenvalue kernel) for the stage split and the Sturm-step count
    ctx.set_timing(True)
    ctx.reset_timing()
    m = ctx.solve(grid, problems)
    W, kb = bb.synthesize_from_modes(ctx, m, nset, f, ranges, np.ones(nset), S, BB_MAXCEL, nfft)
    ctx.synchronize()
    total_modes = m.total
    sturm = float(m.sweep_rows().astype(np.float64).sum())   # rows swept per mode, summed
    fams = {n: ctx.kernel_time(n) for n in ("setup", "interval", "modes", "bb_transfer", "fft")}
    ctx.set_timing(False)
    for b in (W,) + kb:
        b.free()
    m.free()
    for _ in range(args.warmup):
        plan.run()
    plan.result()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run()
    _, plan_modes = plan.result()
    elapsed = time.perf_counter() - t0
    plan.close()
    ctx.close()
    assert plan_modes == total_modes, (plan_modes, total_modes)
    P = len(problems)
    # the tuned k_modes kernel (the plan set the context's form and depth) in the eager solve of
    # the same batch, HIP events on the library stream around its one launch
    modes_s = fams["modes"][0] / max(fams["modes"][1], 1) / 1e3
    achieved = FLOPS_PER_STURM_STEP * sturm / modes_s / 1e12
    value = P * args.steps / elapsed
    line = {
        "metric": METRIC + " (broadband job)",
        "value": value,
        "unit": "dispersion solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "ModBB job: dispersion at 1024 frequencies (2/1024 - 2 Hz) x 8 azimuths, "
                               "Nz=20000, + transfer functions at 16 ranges (200-1000 km) + pulse "
                               "synthesis NFFT %d, Gaussian-windowed 0.5 Hz source spectrum; one step = one "
                               "replay of the job's HIP graph (eigenvalue launch form picked by the plan's autotuner)" % nfft,
                   "global_batch": P, "modes_per_step": int(total_modes), "nfft": nfft,
                   "parallelism": "single GPU"},
        "eager_stage_ms": {k: v[0] for k, v in fams.items()},
        "plan": {"create_s": plan_s, "ms_passes": tuning["ms_passes"], "grid_blocks": tuning["grid_blocks"],
                 "k_modes_ms": tuning["modes_ms"],
                 "tuner_ms": {"%d/%d" % (int(a), int(b)): round(c, 3) for a, b, c in tuning["log"]}},
        "roofline": {"kernel": "k_modes<%d%s> (%s)" % (tuning["ms_passes"], ", true" if tuning["grid_blocks"] < 0 else "",
                                                       "one block per work-list slot, diagonal through LDS"
                                                       if tuning["grid_blocks"] == 0 else
                                                       "one block per work-list slot, diagonal by scalar loads"
                                                       if tuning["grid_blocks"] < 0 else
                                                       "persistent, %d blocks" % tuning["grid_blocks"]),
                     "bound": "valu",
                     "achieved": achieved, "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS["f64"],
                     "traffic": measured_traffic("k_modes<%d%s> (8192-problem batch)"
                                                 % (tuning["ms_passes"], ", true" if tuning["grid_blocks"] < 0 else "")),
                     "algorithmic": "3 FP64 flops per Sturm step (row x sweep), counting the rows each "
                                    "mode's sweeps covered (the evanescent top left out, ncpa_modes_sweep_rows); "
                                    "FP64 VALU (peak: the FP64 vector rate, equal to the FP64 MFMA rate on MI355X)"},
        "cpu_baseline": None,
    }
    if cpu is not None:
        line["cpu_baseline"] = {
            "value": cpu["rate"], "unit": "dispersion solves/s", "cores": cpu["workers"], "kind": "port",
            "sample": "the oracle's C restatement of the reference loops (gcc -O2, FP64) on %d threads, one "
                      "dispersion solve per thread at frequencies spread uniformly over the band (Sturm "
                      "bisection + inverse iteration, doPerturb, a complex exp and sqrt per (mode, range) "
                      "term at 16 ranges); the 128 pulse FFTs of a job (~1 ms on the GPU) are not timed; "
                      "mean %.2f s per solve, %.1f s wall" % (cpu["workers"], cpu["mean_solve_s"], cpu["wall_s"]),
            "host": cpu["host"]}
        line["speedup_vs_cpu_all_cores"] = value / line["cpu_baseline"]["value"]
    return line


FLOPS_PER_WIDE_STEP = 6.0   # one row of one wide-angle count sweep: a_i(k) (2 FMA) + continuant (1 FMA)


def wmod_main(args, cpu):
    """Wide-angle (WMod) solves: one step = the 64-problem batch of the field workload (1 Hz,
    8 profiles x 8 azimuths, Nz = 20000) through ncpa_wmod_solve (every real eigenvalue of
    the quadratic problem in [k_min, k_max], perturbative attenuation) + 1-D TL at 10000
    ranges, all FP64.  The same batch through the effective-sound-speed solver is timed
    beside it."""
    import numpy as np
    import ncpamodes
    tmp = tempfile.mkdtemp(prefix="ncpa_bench_")
    paths = write_profiles(tmp)
    ctx = ncpamodes.Context(0)
    pids = [ctx.read_profile(p)[0] for p in paths]
    grid = ctx.grid(pids[0], Nz=NZ)
    problems = [ncpamodes.problem(pid, FREQ, az) for pid in pids for az in AZIMUTHS]
    P = len(problems)
    dr = MAXRANGE_M / NR
    bufs = [ctx.empty(P * NR) for _ in range(3)]

    def step(wide):
        m = ctx.solve_wide(grid, problems) if wide else ctx.solve(grid, problems)
        m.tl1d_device(dr, NR, False, *[b.ptr for b in bufs])
        return m

    # one step with kernel timing (after an untimed one that loads the kernels): the stage
    # split and the Sturm-step count
    step(True).free()
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.reset_timing()
    m = step(True)
    ctx.synchronize()
    total_modes = m.total
    # rows swept per eigenvalue (the search's count/Newton sweeps and the final sweep, each
    # below its block's evanescent cut), summed
    rows_swept = float(m.sweep_rows().astype(np.float64).sum())
    fams = {n: ctx.kernel_time(n) for n in ("setup", "interval", "modes", "tl1d")}
    ctx.set_timing(False)
    m.free()

    def timed(wide):
        for _ in range(args.warmup):
            step(wide).free()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(wide).free()
        ctx.synchronize()
        return time.perf_counter() - t0

    elapsed = timed(True)
    ess_elapsed = timed(False)
    for b in bufs:
        b.free()
    ctx.close()
    value = P * args.steps / elapsed
    modes_s = fams["modes"][0] / max(fams["modes"][1], 1) / 1e3
    achieved = FLOPS_PER_WIDE_STEP * rows_swept / modes_s / 1e12
    line = {
        "metric": "wide-angle (WMod) modal TL solves/sec, FP64",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "WMod batch: 64 solves at 1 Hz (8 seeded jet profiles x 8 azimuths), "
                               "Nz=20000, every real eigenvalue of the quadratic problem in [k_min, k_max] "
                               "with perturbative attenuation, 1-D TL at 10000 ranges (0.1-1000 km)",
                   "global_batch": P, "modes_per_step": int(total_modes), "parallelism": "single GPU"},
        "stage_ms": {k: v[0] for k, v in fams.items()},
        "ess_same_batch": {"solves_per_s": P * args.steps / ess_elapsed,
                           "wide_over_ess_time": elapsed / ess_elapsed},
        "roofline": {"kernel": "k_wmodes<2, 256, true>", "bound": "valu", "achieved": achieved,
                     "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["f64"],
                     "traffic": measured_traffic("k_wmodes<2, 256, true>"),
                     "algorithmic": "6 FP64 flops per row per sweep (the row value a_i(k) = (P - kQ)^2 "
                                    "- 2 - dz^2 k^2 and the continuant; Newton sweeps do 12, counted as 6); "
                                    "rows = those each eigenvalue's search and final sweeps covered "
                                    "(ncpa_modes_sweep_rows, the evanescent top left out); FP64 VALU, "
                                    "MI355X FP64 peak"},
        "cpu_baseline": None,
    }
    if cpu is not None:
        line["cpu_baseline"] = {
            "value": cpu["rate"], "unit": "solves/s", "cores": cpu["workers"], "kind": "port",
            "sample": "%d concurrent 1-thread processes, each one 1 Hz wide-angle solve (FP64 oracle: "
                      "branch crossings by Brent on LAPACK bisection, inverse-iteration vectors) on every "
                      "%dth eigenvalue branch, eigen time scaled by %d; trace, perturbation and 1-D TL at "
                      "10000 ranges in full; mean %.1f s per solve"
                      % (cpu["workers"], WIDE_CPU_STRIDE, WIDE_CPU_STRIDE, cpu["mean_solve_s"])}
        line["speedup_vs_cpu_all_cores"] = value / cpu["rate"]
    return line


def broadband_wmod_main(args):
    """The broadband job's dispersion set (1024 frequencies x 8 azimuths, Nz = 20000) through the
    wide-angle solver (ModBB --dispersion --method wmod), eagerly (the graph plan is built for
    the effective-sound-speed kernels), followed by the same transfer functions and pulse
    synthesis.  One step = one job."""
    import numpy as np
    import ncpamodes
    from ncpamodes import broadband as bb
    tmp = tempfile.mkdtemp(prefix="ncpa_bench_")
    path = write_profiles(tmp, [0])[0]
    ctx = ncpamodes.Context(0)
    pid = ctx.read_profile(path)[0]
    grid = ctx.grid(pid, Nz=NZ)
    f = np.arange(BB_NF + 1) * BB_FSTEP
    problems = [ncpamodes.problem(pid, f[i], az) for az in AZIMUTHS for i in range(1, BB_NF + 1)]
    ranges = np.array(BB_RANGES_KM) * 1e3
    nfft = bb.default_nfft(4096, f.size)
    S, _ = bb.source(ctx, "spectrum", f, nfft, -1.0, np.stack([f, bb_source_spectrum(f).real,
                                                                bb_source_spectrum(f).imag], axis=1))
    nset = len(AZIMUTHS)

    def job(timing=False):
        m = ctx.solve_wide(grid, problems)
        W, kb = bb.synthesize_from_modes(ctx, m, nset, f, ranges, np.ones(nset), S, BB_MAXCEL, nfft)
        return m, W, kb

    def release(res):
        m, W, kb = res
        for b in (W,) + kb:
            b.free()
        m.free()

    release(job())
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.reset_timing()
    res = job()
    ctx.synchronize()
    total_modes = res[0].total
    rows_swept = float(res[0].sweep_rows().astype(np.float64).sum())
    fams = {n: ctx.kernel_time(n) for n in ("setup", "interval", "modes", "bb_transfer", "fft")}
    ctx.set_timing(False)
    release(res)
    for _ in range(args.warmup):
        release(job())
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        release(job())
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.close()
    P = len(problems)
    modes_s = fams["modes"][0] / max(fams["modes"][1], 1) / 1e3
    achieved = FLOPS_PER_WIDE_STEP * rows_swept / modes_s / 1e12
    return {
        "metric": "wide-angle (WMod) dispersion solves/sec, broadband job, FP64",
        "value": P * args.steps / elapsed, "unit": "dispersion solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "ModBB --method wmod job: wide-angle dispersion at 1024 frequencies (2/1024 - 2 Hz) "
                               "x 8 azimuths, Nz=20000, + transfer functions at 16 ranges + pulse synthesis NFFT "
                               "%d; eager (one batched solve per job)" % nfft,
                   "global_batch": P, "modes_per_step": int(total_modes), "parallelism": "single GPU"},
        "stage_ms": {k: v[0] for k, v in fams.items()},
        "roofline": {"kernel": "k_wmodes<2, 128, true>", "bound": "valu", "achieved": achieved,
                     "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["f64"],
                     "traffic": None, "algorithmic": "6 FP64 flops per swept row, as the wmod workload"},
        "cpu_baseline": None,
    }


def precision_main(args):
    """Config 3: the 2-D field at 2 Hz on 1500 depths x 10000 ranges, every contraction precision
    against the FP64 contraction of the same modes: time per launch (HIP events), speed-up over
    FP64, max / RMS |TL - TL_fp64| over TL > -140 dB, and the algorithmic and MFMA-work rates."""
    import numpy as np
    import ncpamodes
    tmp = tempfile.mkdtemp(prefix="ncpa_bench_")
    paths = write_profiles(tmp, [0, 1])
    ctx = ncpamodes.Context(0)
    pids = [ctx.read_profile(p)[0] for p in paths]
    grid = ctx.grid(pids[0], Nz=NZ)
    problems = [ncpamodes.problem(pids[0], 2.0, 90.0), ncpamodes.problem(pids[1], 2.0, 225.0)]
    m = ctx.solve(grid, problems)
    rows = m.rows_device(0, FIELD_STRIDE, FIELD_ROWS)
    dr = MAXRANGE_M / NR
    P = len(problems)
    times, res = {}, {}
    flops = 4.0 * FIELD_ROWS * m.total * NR
    n = P * FIELD_ROWS * NR
    o64 = m.field2d_device(rows, FIELD_ROWS, dr, NR, 0)
    for prec in (0, 1, 6, 3, 7, 5, 4, 2):
        out = ctx.empty(P * FIELD_ROWS * 2 * NR, np.float64 if prec == 0 else np.float32)
        for _ in range(max(args.warmup, 1)):
            m.field2d_device(rows, FIELD_ROWS, dr, NR, prec, out=out)
        ctx.synchronize()
        ctx.set_timing(True)
        ctx.reset_timing()
        for _ in range(args.steps):
            m.field2d_device(rows, FIELD_ROWS, dr, NR, prec, out=out)
        ctx.synchronize()
        ms, n_l = ctx.kernel_time("field2d")
        ctx.set_timing(False)
        t = ms / max(n_l, 1)
        times[prec] = t
        mx, rms, _ = ctx.field_compare(o64, out, prec, n)
        out.free()
        res[PREC_NAMES[prec]] = {
            "ms": t, "speedup_vs_fp64": times[0] / t, "max_db": mx, "rms_db": rms,
            "algorithmic_tflops": flops / (t / 1e3) / 1e12,
            "mfma_work_tflops": flops * PRODUCTS[prec] / (t / 1e3) / 1e12,
            "peak_tflops": PEAK_TFLOPS[PREC_PEAK[prec]]}
    total = m.total
    o64.free()
    rows.free()
    m.free()
    ctx.close()
    best = min((p for p in times if p != 0 and res[PREC_NAMES[p]]["max_db"] <= TL_GATE_DB), key=lambda p: times[p])
    t = times[best]
    return {
        "metric": METRIC + " (mixed-precision 2-D field)",
        "value": P / (t / 1e3), "unit": "2-D fields/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+%s" % PREC_NAMES[best], "data": "synthetic",
        "config": {"workload": "2-D field at 2 Hz, 1500 depths x 10000 ranges (150 km x 1000 km), Nz=20000, "
                               "2 problems (seeds 0, 1; azimuths 90, 225), %d modes; one step = one contraction "
                               "launch of both fields in the fastest precision within 0.1 dB of FP64" % total,
                   "modes": int(total), "parallelism": "single GPU"},
        "precisions": res,
        "roofline": {"kernel": "field contraction (%s)" % PREC_NAMES[best], "bound": "mfma",
                     "achieved": res[PREC_NAMES[best]]["algorithmic_tflops"], "peak": PEAK_TFLOPS[PREC_PEAK[best]],
                     "unit": "TFLOP/s", "frac": res[PREC_NAMES[best]]["algorithmic_tflops"] / PEAK_TFLOPS[PREC_PEAK[best]],
                     "traffic": None,
                     "algorithmic": "4 x depths x modes x ranges real flops per problem (real x complex)"},
        "cpu_baseline": None,
    }


def serving_main(args):
    """Config 4: Poisson request streams through the continuous-batching server."""
    import numpy as np
    from ncpamodes import serve
    n_profiles = 64
    streams = [(10.0, 150), (100.0, 1000), (500.0, 2000)]
    srv = serve.Server(n_workers=3, precision=6, max_batch=64)
    srv.warm(range(n_profiles))
    warm = serve.make_requests(64, 1000.0, seed=99, n_profiles=n_profiles)
    srv.serve(warm, realtime=False)
    # capacity: the whole stream queued at t = 0
    cap_reqs = serve.make_requests(2000, 500.0, seed=1, n_profiles=n_profiles)
    for r in cap_reqs:
        r.t_arrival = 0.0
    wall = srv.serve(cap_reqs, realtime=False)
    capacity = len(cap_reqs) / wall
    lat = {}
    for rate, n in streams:
        reqs = serve.make_requests(n, rate, seed=int(rate), n_profiles=n_profiles)
        w = srv.serve(reqs, realtime=True)
        st = serve.latency_stats(reqs)
        st["offered_per_s"] = rate
        st["achieved_per_s"] = n / w
        st["requests"] = n
        lat["%g" % rate] = st
    srv.close()
    # plain per-request FP64 path: one request at a time, FP64 contraction, no batching
    plain = serve.Server(n_workers=1, precision=0, max_batch=1)
    plain.warm(range(n_profiles))
    preqs = serve.make_requests(300, 500.0, seed=1, n_profiles=n_profiles)
    for r in preqs:
        r.t_arrival = 0.0
    pw = plain.serve(preqs, realtime=False)
    p10 = serve.make_requests(150, 10.0, seed=10, n_profiles=n_profiles)
    plain.serve(p10, realtime=True)
    plain_lat = serve.latency_stats(p10)
    plain.close()
    return {
        "metric": METRIC,
        "value": capacity,
        "unit": "solves/s",
        "n_gpus": 1, "steps": len(cap_reqs), "warmup": len(warm),
        "ms_per_step": 1e3 / capacity,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+bf16x6f",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson request streams; per request a seeded profile (jet amplitudes "
                               "+-40%, heights +-5 km, 64-profile pool resident on the device), f log-uniform "
                               "0.05-2 Hz, azimuth U(0,360), max range U(200,1500) km, 80% 1-D TL (100 m steps) "
                               "/ 20% 2-D field (500 depths x 1 km steps); 3 device contexts, batches <= 64",
                   "parallelism": "single GPU", "Nz": NZ},
        "latency": lat,
        "fp64_plain_per_request": {"capacity_per_s": len(preqs) / pw, "latency_at_10_per_s": plain_lat},
        "roofline": None,
        "cpu_baseline": None,
    }


def serving_short():
    """p50/p99 request latency of the continuous-batching server (production precision BF16x6F,
    3 device contexts) on two short Poisson streams, and its capacity with the stream queued."""
    from ncpamodes import serve
    n_profiles = 64
    srv = serve.Server(n_workers=3, precision=6, max_batch=64)
    srv.warm(range(n_profiles))
    warm = serve.make_requests(64, 1000.0, seed=99, n_profiles=n_profiles)
    srv.serve(warm, realtime=False)
    cap = serve.make_requests(1000, 500.0, seed=1, n_profiles=n_profiles)
    for r in cap:
        r.t_arrival = 0.0
    capacity = len(cap) / srv.serve(cap, realtime=False)
    out = {"unit": "ms", "capacity_per_s": capacity, "precision": "bf16x6f",
           "workload": "Poisson arrivals; per request a seeded profile (jet amplitudes +-40%, heights +-5 km, "
                       "64-profile pool resident), f log-uniform 0.05-2 Hz, azimuth U(0,360), max range "
                       "U(200,1500) km, 80% 1-D TL / 20% 2-D field; 3 device contexts, batches <= 64"}
    for rate, n in ((100.0, 400), (500.0, 1000)):
        reqs = serve.make_requests(n, rate, seed=int(rate), n_profiles=n_profiles)
        w = srv.serve(reqs, realtime=True)
        st = serve.latency_stats(reqs)
        st["offered_per_s"], st["achieved_per_s"], st["requests"] = rate, n / w, n
        out["%g_per_s" % rate] = st
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["field", "broadband", "serving", "precision", "wmod", "broadband-wmod"],
                    default="field")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--schedule", choices=["pipeline", "streams"], default="pipeline",
                    help="field workload headline: two-stage pipeline over two streams, or "
                         "independent batches on --streams streams")
    ap.add_argument("--streams", type=int, default=3,
                    help="field workload with --schedule streams: batches in flight on this many streams")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run the plain copy kernel over MIB MiB first (byte-counter calibration)")
    ap.add_argument("--modes-form", type=int, choices=[0, 1], default=1,
                    help="field workload: eigenvalue sweeps read the diagonal through LDS (0) or scalar loads (1)")
    ap.add_argument("--headline-only", action="store_true",
                    help="field workload: skip the latency and broadband legs that follow the headline")
    ap.add_argument("--cpu-worker", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(json.loads(args.cpu_worker))
        return 0
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); independent "
              "(profile, frequency, azimuth) solves would shard across GPUs with no communication"
              % args.gpus, file=sys.stderr)
        return 2
    if args.workload == "precision":
        print(json.dumps(precision_main(args)), flush=True)
        return 0
    if args.workload == "serving":
        print(json.dumps(serving_main(args)), flush=True)
        return 0
    if args.workload == "broadband-wmod":
        print(json.dumps(broadband_wmod_main(args)), flush=True)
        return 0
    cores = cpu_cores()
    if args.no_cpu:
        cpu = None
    elif args.workload in ("field", "broadband"):
        cpu = port_baseline(args.workload, cores)
    else:
        cpu = cpu_baseline(args.workload, cores)
    if args.workload == "field":
        line = field_main(args, cpu)
        if not args.headline_only:
            # the rest of the metric in the same run: request latency under Poisson arrivals
            # and the batched broadband job with its CPU comparison
            line["latency"] = serving_short()
            bb_args = argparse.Namespace(**vars(args))
            bb_args.steps, bb_args.warmup = 5, 1
            bb_cpu = None if args.no_cpu else port_baseline("broadband", cores)
            bbl = broadband_main(bb_args, bb_cpu)
            line["broadband_job"] = {k: bbl[k] for k in ("value", "unit", "ms_per_step", "config", "roofline",
                                                         "cpu_baseline", "plan") if k in bbl}
            if "speedup_vs_cpu_all_cores" in bbl:
                line["broadband_job"]["speedup_vs_cpu_all_cores"] = bbl["speedup_vs_cpu_all_cores"]
            # the mixed-precision mode sum against the FP64 contraction at 2 Hz (config 3)
            pr_args = argparse.Namespace(**vars(args))
            pr_args.steps, pr_args.warmup = 10, 2
            prl = precision_main(pr_args)
            line["precision_2hz"] = {"config": prl["config"]["workload"], "chosen": prl["dtype"],
                                     "speedup_vs_fp64": prl["precisions"][prl["dtype"].split("+")[1]]["speedup_vs_fp64"],
                                     "precisions": {k: {kk: v[kk] for kk in ("ms", "speedup_vs_fp64", "max_db", "rms_db")}
                                                    for k, v in prl["precisions"].items()}}
    elif args.workload == "wmod":
        line = wmod_main(args, cpu)
    else:
        line = broadband_main(args, cpu)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
