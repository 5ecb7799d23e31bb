"""Request serving over the ncpamodes C ABI: continuous batching of propagation solves.

Requests (each with its own atmosphere, frequency, azimuth, maximum range and output kind)
arrive over time.  A dispatcher admits queued requests into batches while their estimated
device-memory footprint fits the budget, and hands each batch to the first idle worker.
Every worker owns a device context (its own HIP stream, default priority) and a cache of
device-resident profiles, so batches on different workers run concurrently on the GPU and
requests that arrive while others are in flight are admitted as soon as a worker and memory
free up.  One batch is one ncpa_ess_solve launch over all of its problems (uneven profiles,
frequencies and mode counts), followed by one 1-D TL launch for the 1-D requests, and
eigenvector rows + one 2-D field contraction for the 2-D requests.

Per-request results do not depend on the batch a request lands in: every kernel computes
each problem from its own inputs only (checked by tests/test_gpu_serve.py).
"""
import collections
import ctypes as C
import math
import threading
import time

import numpy as np

from . import synth

# output grids of a request (fixed steps; the number of ranges follows its maximum range)
TL1D_DR_M = 100.0
TL2D_DR_M = 1000.0
TL2D_ROWS = 500
ROOM_AGL_KM = 150.0


class Request:
    __slots__ = ("rid", "seed", "freq", "azimuth", "rmax_km", "kind", "t_arrival", "t_done", "result", "batch")

    def __init__(self, rid, seed, freq, azimuth, rmax_km, kind, t_arrival=0.0):
        self.rid, self.seed, self.freq, self.azimuth = rid, seed, freq, azimuth
        self.rmax_km, self.kind, self.t_arrival = rmax_km, kind, t_arrival
        self.t_done = None
        self.result = None
        self.batch = None

    def n_ranges(self):
        dr = TL1D_DR_M if self.kind == "1d" else TL2D_DR_M
        return max(1, int(round(self.rmax_km * 1000.0 / dr)))

    def footprint_bytes(self, Nz):
        """Device-memory estimate: grid arrays, modes (<= 4000), outputs and contraction staging."""
        modes_est = min(4000, int(900 * self.freq) + 50)
        b = 3 * Nz * 8 + modes_est * 12 * 8
        if self.kind == "1d":
            b += 3 * self.n_ranges() * 8
        else:
            nr = self.n_ranges()
            b += TL2D_ROWS * modes_est * 8 + TL2D_ROWS * 2 * nr * 8 + modes_est * 2 * nr * 3 * 2 + 3 * TL2D_ROWS * modes_est * 2
        return b


def make_requests(n, rate_hz, seed=0, n_profiles=None, frac_2d=0.2):
    """Seeded request stream: Poisson arrivals at rate_hz; per request a profile seed (jet
    amplitudes +-40 %, jet heights +-5 km), f log-uniform in 0.05-2 Hz, azimuth U(0, 360),
    maximum range U(200, 1500) km, output 2-D with probability frac_2d else 1-D."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, size=n))
    out = []
    for i in range(n):
        pseed = int(rng.integers(0, n_profiles if n_profiles else 1 << 30))
        f = float(math.exp(rng.uniform(math.log(0.05), math.log(2.0))))
        az = float(rng.uniform(0.0, 360.0))
        rmax = float(rng.uniform(200.0, 1500.0))
        kind = "2d" if rng.uniform() < frac_2d else "1d"
        out.append(Request(i, pseed, f, az, rmax, kind, float(t[i])))
    return out


def profile_arrays(seed):
    """SI node arrays of the serving profile family (jet amplitudes +-40 %, heights +-5 km)."""
    z, u, v, w, T, rho_gcc, P_mbar = synth.profile_columns(seed, amp=0.4, height_shift_km=5.0)
    P = P_mbar * 100.0
    rho = rho_gcc * 1000.0
    c0 = np.sqrt(1.4 * P / rho)          # C0 from P and RHO (ESSModeSolver.cpp:159-183)
    return z * 1000.0, c0, u, v, T, P, rho


class Worker:
    """A device context with its own stream and a cache of device-resident profiles."""

    def __init__(self, wid, Nz, precision, ctx_factory):
        self.wid = wid
        self.ctx = ctx_factory()
        self.Nz = Nz
        self.precision = precision
        self.pids = {}
        self.grid = None

    def profile(self, seed):
        pid = self.pids.get(seed)
        if pid is None:
            z, c0, u, v, T, P, rho = (np.ascontiguousarray(a, dtype=np.float64) for a in profile_arrays(seed))
            dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
            h = C.c_int()
            self.ctx.check(self.ctx.lib.ncpa_profile_upload(self.ctx.h, z.size, dp(z), dp(c0), dp(u), dp(v), dp(T),
                                                            dp(P), dp(rho), C.byref(h)))
            pid = h.value
            self.pids[seed] = pid
            if self.grid is None:
                self.grid = self.ctx.grid(pid, Nz=self.Nz, maxheight_km=ROOM_AGL_KM)
        return pid

    def run_batch(self, reqs, precision=None):
        """Solves a batch and stores each request's result (host arrays)."""
        import ncpamodes
        prec = self.precision if precision is None else precision
        ctx = self.ctx
        groups = {"1d": [r for r in reqs if r.kind == "1d"], "2d": [r for r in reqs if r.kind == "2d"]}
        for kind, rs in groups.items():
            if not rs:
                continue
            probs = [ncpamodes.problem(self.profile(r.seed), r.freq, r.azimuth) for r in rs]
            m = ctx.solve(self.grid, probs)
            nr = max(r.n_ranges() for r in rs)
            if kind == "1d":
                p, inc = m.tl1d(TL1D_DR_M, nr)
                for i, r in enumerate(rs):
                    n = r.n_ranges()
                    r.result = (p[i, :n], inc[i, :n])
            else:
                stride = max(1, self.Nz // TL2D_ROWS)
                rows = m.rows_device(0, stride, TL2D_ROWS)
                out = m.field2d_device(rows, TL2D_ROWS, TL2D_DR_M, nr, prec)
                a = out.numpy().astype(np.float64).reshape(len(rs), TL2D_ROWS, nr, 2)
                out.free()
                rows.free()
                for i, r in enumerate(rs):
                    n = r.n_ranges()
                    r.result = a[i, :, :n, 0] + 1j * a[i, :, :n, 1]
            m.free()
        ctx.synchronize()

    def close(self):
        self.ctx.close()


class Server:
    """Continuous-batching dispatcher over n_workers device contexts."""

    def __init__(self, n_workers=2, Nz=20000, precision=6, hbm_budget_gb=32.0, max_batch=64, ctx_factory=None):
        import ncpamodes
        factory = ctx_factory or (lambda: ncpamodes.Context(0))
        self.workers = [Worker(i, Nz, precision, factory) for i in range(n_workers)]
        self.Nz = Nz
        self.budget = hbm_budget_gb * 1e9
        self.max_batch = max_batch
        self.in_flight = 0.0
        self.lock = threading.Lock()
        self.cv = threading.Condition(self.lock)
        self.idle = collections.deque(self.workers)
        self.errors = []
        self.batches = 0
        self.max_in_flight = 0.0

    def warm(self, seeds):
        for w in self.workers:
            for s in seeds:
                w.profile(s)

    def _run(self, w, batch, fp, t0):
        try:
            w.run_batch(batch)
            t = time.perf_counter() - t0
            for r in batch:
                r.t_done = t
        except Exception as e:   # recorded and re-raised by serve()
            self.errors.append(e)
        with self.cv:
            self.in_flight -= fp
            self.idle.append(w)
            self.cv.notify_all()

    def serve(self, requests, realtime=True):
        """Runs the stream (arrival times in seconds from the start); returns wall seconds."""
        pending = collections.deque(sorted(requests, key=lambda r: r.t_arrival))
        queue = collections.deque()
        threads = []
        t0 = time.perf_counter()
        with self.cv:
            while pending or queue or len(self.idle) < len(self.workers):
                if self.errors:
                    break
                now = time.perf_counter() - t0
                while pending and (pending[0].t_arrival <= now or not realtime):
                    queue.append(pending.popleft())
                launched = False
                while queue and self.idle:
                    batch, fp = [], 0.0
                    while queue and len(batch) < self.max_batch:
                        need = queue[0].footprint_bytes(self.Nz)
                        if batch and self.in_flight + fp + need > self.budget:
                            break
                        if not batch and self.in_flight + need > self.budget and self.in_flight > 0:
                            break   # wait until memory frees up
                        batch.append(queue.popleft())
                        fp += need
                    if not batch:
                        break
                    w = self.idle.popleft()
                    self.in_flight += fp
                    self.max_in_flight = max(self.max_in_flight, self.in_flight)
                    self.batches += 1
                    for r in batch:
                        r.batch = self.batches
                    th = threading.Thread(target=self._run, args=(w, batch, fp, t0), daemon=True)
                    threads.append(th)
                    th.start()
                    launched = True
                if launched:
                    continue
                timeout = None
                if pending and realtime:
                    timeout = max(0.0, pending[0].t_arrival - (time.perf_counter() - t0))
                self.cv.wait(timeout=timeout if timeout is not None else 0.05)
        for th in threads:
            th.join()
        if self.errors:
            raise self.errors[0]
        return time.perf_counter() - t0

    def close(self):
        for w in self.workers:
            w.close()


def latency_stats(requests):
    lat = np.array([r.t_done - r.t_arrival for r in requests])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}
