"""Request serving over the ncpamodes C ABI: continuous batching of propagation solves.

Requests (each with its own atmosphere, frequency, azimuth, maximum range and output kind)
arrive over time.  As they arrive, the dispatcher counts their modes exactly on the device
(ncpa_ess_count: the grid, the interval and its Sturm counts, no eigenvalues), then admits
queued requests into batches while the batches' device-memory footprint fits the budget, and
hands each batch to the first idle worker.  The footprint of a batch is the peak of the
library's own allocations for it (batch_footprint mirrors the allocation sites of
ncpa_ess_solve, ncpa_tl1d, ncpa_modes_rows and ncpa_field2d, and tests compare it with the
allocator's measured peak, ncpa_ctx_memory).  Every worker owns a device context (its own HIP
stream, default priority) and a cache of device-resident profiles with the grid each one
implies (setParams: the maxheight clamp to the profile top and the ground at Z0), so batches
on different workers run concurrently on the GPU and requests that arrive while others are in
flight are admitted as soon as a worker and memory free up.  One batch is one ncpa_ess_solve
launch per grid over its problems (uneven profiles, frequencies and mode counts), followed by
one 1-D TL launch for the 1-D requests, and eigenvector rows + one 2-D field contraction for
the 2-D requests.

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
    __slots__ = ("rid", "seed", "freq", "azimuth", "rmax_km", "kind", "t_arrival", "t_done", "result", "batch",
                 "nm")

    def __init__(self, rid, seed, freq, azimuth, rmax_km, kind, t_arrival=0.0):
        self.rid, self.seed, self.freq, self.azimuth = rid, seed, freq, azimuth
        self.rmax_km, self.kind, self.t_arrival = rmax_km, kind, t_arrival
        self.t_done = None
        self.result = None
        self.batch = None
        self.nm = None   # mode count (ncpa_ess_count), set when the request is queued

    def n_ranges(self):
        dr = TL1D_DR_M if self.kind == "1d" else TL2D_DR_M
        return max(1, int(round(self.rmax_km * 1000.0 / dr)))


# ---------------------------------------------------------------------------------------
# device-memory footprint of a batch, from the library's allocation sites
# ---------------------------------------------------------------------------------------
def _r256(b):
    """DevicePool::get rounds every request up to 256 bytes (ctx.cpp)."""
    b = max(int(b), 1)
    return (b + 255) // 256 * 256


SIZEOF = {"ProbDev": 96, "IntervalOut": 48, "ModeProb": 88, "ModeWork": 8, "TlProb": 8, "SplitProb": 32,
          "ETile": 8, "GemmProb": 32}
SPLIT_PARTS = {2: 2, 3: 3, 4: 2, 5: 2, 6: 3, 7: 2}   # 16-bit parts per operand, precisions 2..7
SPLIT_F16 = (4, 5, 7)                                # FP16 forms (row / range scales)


def _solve_bytes(nms, Nz, max_nodes):
    """(persistent, transient) bytes of ncpa_ess_solve over problems with mode counts nms
    (ess.hip: the grid samples and mode arrays stay with the modes; the problem records, node
    tables and intervals are returned when the call ends)."""
    P, tot = len(nms), max(int(sum(nms)), 1)
    groups = [(n + 63) // 64 for n in nms]
    # XCD-balanced work list: at most ceil(groups / 8) + the largest problem's groups per XCD
    work = 8 * ((sum(groups) + 7) // 8 + max(groups + [0]))
    persistent = (3 * _r256(8 * P * Nz) + 5 * _r256(8 * tot) + 3 * _r256(4 * tot) + _r256(4 * (P + 1))
                  + _r256(SIZEOF["ModeProb"] * P) + _r256(SIZEOF["ModeWork"] * max(work, 1)))
    transient = (_r256(SIZEOF["ProbDev"] * P) + _r256(16 * P * max_nodes) + _r256(SIZEOF["IntervalOut"] * P))
    return persistent, transient


def _field_bytes(nms, nrows, nr, precision):
    """(held, staging-only) bytes of ncpa_field2d's operand staging (field2d.hip)."""
    P, tot = len(nms), max(int(sum(nms)), 1)
    if precision in SPLIT_PARTS:
        S, f16 = SPLIT_PARTS[precision], precision in SPLIT_F16
        Mp = -(-nrows // 128) * 128
        Np = -(-2 * nr // 128) * 128
        Kp = [-(-int(n) // 32) * 32 for n in nms]
        sa, sb = sum(Mp * k for k in Kp), sum(Np * k for k in Kp)
        held = _r256(SIZEOF["SplitProb"] * P) + _r256(2 * max(S * sa, 8)) + _r256(2 * max(S * sb, 8))
        if max(Kp) <= 128 * 32 and sa > 0:
            held += _r256(4 * (sa // (128 * 32)))            # per-chunk flags of Ahat (k_asplit)
        temps = _r256(SIZEOF["ETile"] * max(sum(Kp) // 32, 1)) + _r256(6 * 8 * tot)   # + per-mode factors
        if f16:
            held += _r256(4 * P * Mp) + _r256(4 * P * (Np // 2))
            temps += 2 * _r256(8 * tot)
        return held, temps
    esz = 8 if precision == 0 else 4
    held = _r256(SIZEOF["GemmProb"] * P) + _r256(tot * 2 * nr * esz)
    if precision != 0:
        held += _r256(4 * nrows * tot)
    return held, 0


def batch_footprint(reqs, Nz, precision, max_nodes=1501, nrows=TL2D_ROWS):
    """Peak device bytes Worker.run_batch allocates for `reqs` (mode counts in r.nm): per output
    kind, the solve's held arrays plus the largest of its transient tables, the 1-D outputs, the
    eigenvector-rows pass and the field staging + output; the kinds run one after the other."""
    peak = 0
    for kind in ("1d", "2d"):
        rs = [r for r in reqs if r.kind == kind]
        if not rs:
            continue
        nms = [int(r.nm) for r in rs]
        P, tot = len(rs), max(sum(nms), 1)
        nr = max(r.n_ranges() for r in rs)
        held, transient = _solve_bytes(nms, Nz, max_nodes)
        if kind == "1d":
            # three outputs of P * nr doubles (ncpa_dev_alloc) and the problem table of ncpa_tl1d
            peak = max(peak, held + max(transient, 3 * 8 * P * nr + _r256(SIZEOF["TlProb"] * P)))
            continue
        slots = nrows + 1
        rows_out = 8 * nrows * tot                                        # ncpa_dev_alloc
        rows_scratch = (2 * _r256(8 * slots * tot) + 2 * _r256(4 * slots * tot) + 3 * _r256(8 * tot)
                        + 2 * _r256(4 * tot))
        psis_rows = _r256(8 * tot)                                        # kept with the modes
        f_held, f_temps = _field_bytes(nms, nrows, nr, precision)
        out = P * nrows * 2 * nr * (8 if precision == 0 else 4)           # ncpa_dev_alloc
        peak = max(peak, held + transient, held + rows_out + rows_scratch + psis_rows,
                   held + rows_out + psis_rows + out + f_held + f_temps)
    return peak


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


class ProfileCache:
    """Device-resident profiles of one context and the grid each one implies."""

    def __init__(self, ctx, Nz):
        self.ctx, self.Nz = ctx, Nz
        self.pids, self.grids = {}, {}

    def get(self, seed):
        """(profile id, grid, grid key) of the serving profile `seed`, uploaded on first use.  The
        grid is made for this profile (ncpa_grid_make: maxheight clamped to its top, ground at
        its Z0), so profiles with different tops or grounds are never solved on one grid."""
        pid = self.pids.get(seed)
        if pid is None:
            z, c0, u, v, T, P, rho = (np.ascontiguousarray(a, dtype=np.float64) for a in profile_arrays(seed))
            dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
            h = C.c_int()
            self.ctx.check(self.ctx.lib.ncpa_profile_upload(self.ctx.h, z.size, dp(z), dp(c0), dp(u), dp(v), dp(T),
                                                            dp(P), dp(rho), C.byref(h)))
            pid = h.value
            self.pids[seed] = pid
            g = self.ctx.grid(pid, Nz=self.Nz, maxheight_km=ROOM_AGL_KM)
            self.grids[seed] = (g, (g.Nz, g.z_min, g.dz, g.maxheight))
        g, key = self.grids[seed]
        return pid, g, key


def _by_grid(cache, reqs):
    """Requests grouped by the grid of their profile: {key: (grid, [(request, pid)])}."""
    out = {}
    for r in reqs:
        pid, g, key = cache.get(r.seed)
        out.setdefault(key, (g, []))[1].append((r, pid))
    return out


class Worker:
    """A device context with its own stream and a cache of device-resident profiles."""

    def __init__(self, wid, Nz, precision, ctx_factory):
        self.wid = wid
        self.ctx = ctx_factory()
        self.Nz = Nz
        self.precision = precision
        self.cache = ProfileCache(self.ctx, Nz)

    def profile(self, seed):
        return self.cache.get(seed)[0]

    def run_batch(self, reqs, precision=None):
        """Solves a batch and stores each request's result (host arrays)."""
        import ncpamodes
        prec = self.precision if precision is None else precision
        ctx = self.ctx
        for kind in ("1d", "2d"):
            for grid, items in _by_grid(self.cache, [r for r in reqs if r.kind == kind]).values():
                rs = [r for r, _ in items]
                m = ctx.solve(grid, [ncpamodes.problem(pid, r.freq, r.azimuth) for r, pid in items])
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
        self.counter = ProfileCache(factory(), Nz)   # the dispatcher's context: mode counts
        self.Nz = Nz
        self.precision = precision
        self.budget = hbm_budget_gb * 1e9
        self.max_batch = max_batch
        self.in_flight = 0.0
        self.lock = threading.Lock()
        self.cv = threading.Condition(self.lock)
        self.idle = collections.deque(self.workers)
        self.errors = []
        self.batches = 0
        self.max_in_flight = 0.0
        self.spans = []          # (batch, worker, t_start, t_end, footprint) per batch

    def warm(self, seeds):
        for s in seeds:
            self.counter.get(s)
        for w in self.workers:
            for s in seeds:
                w.profile(s)

    def count(self, reqs):
        """Sets r.nm for every request: one ncpa_ess_count launch per grid."""
        import ncpamodes
        todo = [r for r in reqs if r.nm is None]
        for grid, items in _by_grid(self.counter, todo).values():
            counts = self.counter.ctx.count(grid, [ncpamodes.problem(pid, r.freq, r.azimuth) for r, pid in items])
            for (r, _), n in zip(items, counts):
                r.nm = int(n)

    def footprint(self, reqs):
        return batch_footprint(reqs, self.Nz, self.precision)

    def _run(self, w, batch, fp, t0):
        t_start = time.perf_counter() - t0
        try:
            w.run_batch(batch)
            t = time.perf_counter() - t0
            for r in batch:
                r.t_done = t
        except Exception as e:   # recorded and re-raised by serve()
            self.errors.append(e)
            t = time.perf_counter() - t0
        with self.cv:
            self.spans.append((batch[0].batch, w.wid, t_start, t, fp))
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
                arrived = []
                while pending and (pending[0].t_arrival <= now or not realtime):
                    arrived.append(pending.popleft())
                if arrived:
                    self.count(arrived)   # exact mode counts before admission
                    queue.extend(arrived)
                launched = False
                while queue and self.idle:
                    batch, fp = [], 0.0
                    while queue and len(batch) < self.max_batch:
                        need = self.footprint(batch + [queue[0]])
                        if batch and self.in_flight + need > self.budget:
                            break
                        if not batch and self.in_flight + need > self.budget and self.in_flight > 0:
                            break   # wait until memory frees up
                        batch.append(queue.popleft())
                        fp = need
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
        self.counter.ctx.close()


def latency_stats(requests):
    lat = np.array([r.t_done - r.t_arrival for r in requests])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}
