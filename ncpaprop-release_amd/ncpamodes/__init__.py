"""Python host mirror of the ESSModeSolver path over the ncpamodes C ABI.

Mirrors the reference's solver interface for tests and benchmarks: a Context owns the HIP
stream and profiles; EssBatch solves (profile, frequency, azimuth) problems on the device
and exposes modes and 1-D transmission loss.  All numerics run in libncpamodes.so.
"""
import ctypes as C
import math
import numpy as np

from . import _lib
from ._lib import Grid, Problem, ProfileInfo, NcpaError, NCPA_F_WKB, NCPA_F_WVNUM_FILTER

PI = 3.141592653589793


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """A device context (one HIP stream).  stream: an int hipStream_t handle or None."""

    def __init__(self, device=0, stream=None):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.ncpa_ctx_create(device, C.c_void_p(stream) if stream else None, C.byref(h))
        if rc != 0:
            raise NcpaError(rc, "ncpa_ctx_create failed (no HIP device?)")
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise NcpaError(rc, self.lib.ncpa_ctx_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.ncpa_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self.check(self.lib.ncpa_ctx_synchronize(self.h))

    def record(self):
        """An Event at the current end of this context's stream."""
        h = C.c_void_p()
        self.check(self.lib.ncpa_event_record(self.h, C.byref(h)))
        return Event(self.lib, h)

    def wait(self, event):
        """Work queued on this context after the call waits for `event` (no host wait)."""
        self.check(self.lib.ncpa_event_wait(self.h, event.h))

    def field_compare(self, ref, test, test_precision, n_points, floor_db=-140.0):
        """(max, rms) |TL_test - TL_ref| [dB] of two device fields of n_points complex values
        over TL_ref > floor_db, and the number of such points (computed on the device)."""
        mx, rms, cnt = C.c_double(), C.c_double(), C.c_size_t()
        self.check(self.lib.ncpa_field_compare(self.h, C.c_void_p(ref.ptr), C.c_void_p(test.ptr), test_precision,
                                               n_points, floor_db, C.byref(mx), C.byref(rms), C.byref(cnt)))
        return mx.value, rms.value, cnt.value

    @property
    def stream(self):
        return self.lib.ncpa_ctx_stream(self.h)

    def set_timing(self, on=True):
        self.check(self.lib.ncpa_ctx_set_timing(self.h, 1 if on else 0))

    def kernel_time(self, name):
        ms = C.c_double()
        n = C.c_int()
        self.check(self.lib.ncpa_ctx_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_modes_form(self, form):
        """Eigenvalue sweeps of eager solves and eigenvector-row sweeps (rows_device): 1 (default)
        diagonal through scalar loads, 0 through LDS; the same results to the bit."""
        self.check(self.lib.ncpa_ctx_set_modes_form(self.h, int(form)))

    def reset_timing(self):
        self.check(self.lib.ncpa_ctx_reset_timing(self.h))

    def memory(self, reset_peak=False):
        """(live, peak, cached) device bytes of this context's allocations (ncpa_ctx_memory)."""
        live, peak, cached = C.c_size_t(), C.c_size_t(), C.c_size_t()
        self.check(self.lib.ncpa_ctx_memory(self.h, 1 if reset_peak else 0, C.byref(live), C.byref(peak),
                                            C.byref(cached)))
        return live.value, peak.value, cached.value

    # contraction variants ----------------------------------------------------------------
    def field_variants(self, precision):
        """Number of launch variants of a precision's contraction (ncpa_field_variants)."""
        n = C.c_int()
        self.check(self.lib.ncpa_field_variants(precision, C.byref(n)))
        return n.value

    def field_variant_info(self, precision, variant):
        """(kernel name, row groups) of a contraction variant (ncpa_field_variant_info)."""
        name, rg = C.c_char_p(), C.c_int()
        self.check(self.lib.ncpa_field_variant_info(precision, variant, C.byref(name), C.byref(rg)))
        return name.value.decode(), rg.value

    def set_field_variant(self, precision, variant):
        self.check(self.lib.ncpa_field_set_variant(self.h, precision, variant))

    def field_variant(self, precision):
        v = C.c_int()
        self.check(self.lib.ncpa_field_get_variant(self.h, precision, C.byref(v)))
        return v.value

    # device memory -------------------------------------------------------------------
    def empty(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    # profiles ------------------------------------------------------------------------
    def read_profile(self, path, header=""):
        pid = C.c_int()
        info = ProfileInfo()
        self.check(self.lib.ncpa_profile_read(self.h, path.encode(), header.encode(), C.byref(pid),
                                              C.byref(info)))
        return pid.value, info

    def grid(self, profile, Nz=20000, maxheight_km=150.0, zground_km=0.0):
        g = Grid()
        self.check(self.lib.ncpa_grid_make(self.h, profile, Nz, maxheight_km, zground_km, C.byref(g)))
        return g

    def admittance(self, profile, z_min):
        a = C.c_double()
        self.check(self.lib.ncpa_profile_admittance(self.h, profile, z_min, C.byref(a)))
        return a.value

    def sample(self, profile, grid):
        rho = np.empty(grid.Nz)
        P = np.empty(grid.Nz)
        T = np.empty(grid.Nz)
        self.check(self.lib.ncpa_profile_sample(self.h, profile, C.byref(grid), _dp(rho), _dp(P),
                                                _dp(T), None))
        return rho, P, T

    # solves --------------------------------------------------------------------------
    def solve(self, grid, problems):
        arr = (Problem * len(problems))(*problems)
        h = C.c_void_p()
        self.check(self.lib.ncpa_ess_solve(self.h, C.byref(grid), arr, len(problems), C.byref(h)))
        return Modes(self, h, len(problems), grid)

    def count(self, grid, problems):
        """Mode counts ncpa_ess_solve would return for each problem (no modes computed)."""
        arr = (Problem * len(problems))(*problems)
        out = (C.c_int * len(problems))()
        self.check(self.lib.ncpa_ess_count(self.h, C.byref(grid), arr, len(problems), out))
        return np.array(out[:], dtype=np.int64)

    def solve_wide(self, grid, problems):
        """Wide-angle (WMod) solve of a batch; the modes' grid spacing is the matrix
        dz = (maxheight - z_min)/Nz of WModeSolver::solve."""
        arr = (Problem * len(problems))(*problems)
        h = C.c_void_p()
        self.check(self.lib.ncpa_wmod_solve(self.h, C.byref(grid), arr, len(problems), C.byref(h)))
        gm = Grid(grid.Nz, grid.z_min, (grid.maxheight - grid.z_min) / grid.Nz, grid.maxheight)
        return Modes(self, h, len(problems), gm)


def ess_eigen(ctx, diag, dz, k_min, k_max, max_modes=4000):
    """Drop-in of EigenEngine::doESSCalculation: (k2 [nconv], psi [Nz, nconv])."""
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    Nz = diag.size
    kmin = C.c_double(k_min)
    kmax = C.c_double(k_max)
    nconv = C.c_int()
    k2 = np.zeros(max_modes)
    v = np.zeros((Nz, max_modes))
    ctx.check(ctx.lib.ncpa_ess_eigen(ctx.h, _dp(diag), Nz, dz, 1e-8, C.byref(kmin), C.byref(kmax),
                                     max_modes, C.byref(nconv), _dp(k2), _dp(v), max_modes))
    n = nconv.value
    return k2[:n], v[:, :n]


def wmod_eigen(ctx, kd, md, cd, dz, k_min, k_max, max_modes=4000):
    """Drop-in of EigenEngine::doWideAngleCalculation + doSelect: (kH [nconv], v [Nz, nconv])."""
    kd = np.ascontiguousarray(kd, dtype=np.float64)
    md = np.ascontiguousarray(md, dtype=np.float64)
    cd = np.ascontiguousarray(cd, dtype=np.float64)
    Nz = kd.size
    nconv = C.c_int()
    kh = np.zeros(max_modes)
    v = np.zeros((Nz, max_modes))
    ctx.check(ctx.lib.ncpa_wmod_eigen(ctx.h, _dp(kd), _dp(md), _dp(cd), Nz, dz, 1e-8, k_min, k_max,
                                      max_modes, C.byref(nconv), _dp(kh), _dp(v), max_modes))
    n = nconv.value
    return kh[:n], v[:, :n]


class Event:
    """A point in one context's stream (ncpa_event_*): other contexts wait for it with
    Context.wait, the host with synchronize()."""

    def __init__(self, lib, h):
        self.lib, self.h = lib, h

    def synchronize(self):
        if self.lib.ncpa_event_synchronize(self.h) != 0:
            raise NcpaError(_lib.NCPA_ERR_HIP, "ncpa_event_synchronize failed")

    def free(self):
        if self.h:
            self.lib.ncpa_event_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FieldStage:
    """Staged operands of one 2-D field (ncpa_field2d_stage)."""

    def __init__(self, lib, h, n_out, precision):
        self.lib, self.h, self.n_out, self.precision = lib, h, n_out, precision

    def contract(self, ctx, out=None):
        """Queue the contraction on `ctx` into `out` (allocated on ctx when None)."""
        if out is None:
            out = ctx.empty(self.n_out, np.float64 if self.precision == 0 else np.float32)
        ctx.check(self.lib.ncpa_field2d_contract(ctx.h, self.h, C.c_void_p(out.ptr)))
        return out

    def tune(self, ctx, out=None, reps=3):
        """Time every contraction variant of this precision on the staged batch
        (ncpa_field_tune); ctx keeps the fastest.  Returns (best, [ms per variant])."""
        own = out is None
        if own:
            out = ctx.empty(self.n_out, np.float64 if self.precision == 0 else np.float32)
        n = ctx.field_variants(self.precision)
        best, ms = C.c_int(), (C.c_double * n)()
        try:
            ctx.check(self.lib.ncpa_field_tune(ctx.h, self.h, C.c_void_p(out.ptr), reps, C.byref(best), ms, n))
        finally:
            if own:   # the tuner has waited for its launches
                out.free()
        return best.value, list(ms)

    def free(self):
        if self.h:
            self.lib.ncpa_field_stage_free(self.h)
            self.h = None


class DeviceArray:
    """A flat device buffer allocated through the C ABI.  Not a torch tensor: torch bundles its
    own HIP runtime, and two HIP runtimes cannot share a process."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(ctx.lib.ncpa_dev_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        self.ctx.check(self.ctx.lib.ncpa_dev_download(self.ctx.h, out.ctypes.data, C.c_void_p(self.ptr),
                                                      self.nbytes))
        return out

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.size == self.n
        self.ctx.check(self.ctx.lib.ncpa_dev_upload(self.ctx.h, C.c_void_p(self.ptr), arr.ctypes.data,
                                                    self.nbytes))

    def free(self):
        if self.ptr:
            self.ctx.lib.ncpa_dev_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def problem(profile, freq, azimuth, zs_m=0.0, zr_m=0.0, wkb=False, alpha_freq=None,
            admittance=0.0, wvnum_filter=None):
    flags = NCPA_F_WKB if wkb else 0
    c_min = c_max = 0.0
    if wvnum_filter is not None:
        flags |= NCPA_F_WVNUM_FILTER
        c_min, c_max = wvnum_filter
    return Problem(profile, flags, freq, azimuth, freq if alpha_freq is None else alpha_freq,
                   zs_m, zr_m, admittance, c_min, c_max)


class Modes:
    """Device-resident modes of a solved batch."""

    def __init__(self, ctx, h, n_problems, grid):
        self.ctx, self.h, self.P, self.grid = ctx, h, n_problems, grid
        counts = (C.c_int * n_problems)()
        tot = C.c_int()
        ctx.check(ctx.lib.ncpa_modes_counts(h, counts, C.byref(tot)))
        self.counts = np.array(counts[:], dtype=np.int64)
        self.total = tot.value
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)])

    def free(self):
        if self.h:
            self.ctx.lib.ncpa_modes_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def interval(self):
        kmin = np.empty(self.P)
        kmax = np.empty(self.P)
        self.ctx.check(self.ctx.lib.ncpa_modes_interval(self.h, _dp(kmin), _dp(kmax)))
        return kmin, kmax

    def download(self):
        n = max(self.total, 1)
        kr, ki, psr, ps, pr = (np.zeros(n) for _ in range(5))
        self.ctx.check(self.ctx.lib.ncpa_modes_download(self.h, _dp(kr), _dp(ki), _dp(psr),
                                                        _dp(ps), _dp(pr)))
        n = self.total
        return kr[:n] + 1j * ki[:n], psr[:n], ps[:n], pr[:n]

    def iterations(self):
        it = np.zeros(max(self.total, 1), dtype=np.int32)
        self.ctx.check(self.ctx.lib.ncpa_modes_iterations(self.h, it.ctypes.data_as(C.POINTER(C.c_int))))
        return it[:self.total]

    def sweep_rows(self):
        """Rows each mode's eigenvalue search swept in total (ncpa_modes_sweep_rows)."""
        r = np.zeros(max(self.total, 1), dtype=np.int32)
        self.ctx.check(self.ctx.lib.ncpa_modes_sweep_rows(self.h, r.ctypes.data_as(C.POINTER(C.c_int))))
        return r[:self.total]

    def ceff(self, p):
        """c_eff [m/s] on the grid rows of problem p (ncpa_modes_ceff)."""
        out = np.empty(self.grid.Nz)
        self.ctx.check(self.ctx.lib.ncpa_modes_ceff(self.h, int(p), _dp(out)))
        return out

    def problem_modes(self, p):
        k, psr, ps, pr = self.download()
        s = slice(self.offsets[p], self.offsets[p + 1])
        return k[s], psr[s], ps[s], pr[s]

    def rows(self, row0, stride, nrows):
        """Eigenvectors at rows row0 + k*stride: array [nrows, total] (psi, sum psi^2 dz = 1)."""
        buf = self.ctx.empty(nrows * max(self.total, 1))
        self.ctx.check(self.ctx.lib.ncpa_modes_rows(self.ctx.h, self.h, row0, stride, nrows,
                                                    C.c_void_p(buf.ptr)))
        out = buf.numpy()[:nrows * self.total].reshape(nrows, self.total)
        buf.free()
        return out

    def rows_device(self, row0, stride, nrows):
        """Eigenvector rows left on the device (DeviceArray of nrows * total doubles)."""
        buf = self.ctx.empty(nrows * max(self.total, 1))
        self.ctx.check(self.ctx.lib.ncpa_modes_rows(self.ctx.h, self.h, row0, stride, nrows,
                                                    C.c_void_p(buf.ptr)))
        return buf

    def field2d_device(self, psi_rows, nrows, dr, n_r, precision=0, lossless=False, out=None, ctx=None):
        """2-D field into a device buffer [P, nrows, 2 n_r] (float64 for FP64 else float32),
        launched on `ctx` (default: the context that solved; another one must first wait for
        an event recorded after ncpa_modes_rows)."""
        c = self.ctx if ctx is None else ctx
        dt = np.float64 if precision == 0 else np.float32
        if out is None:
            out = c.empty(self.P * nrows * 2 * n_r, dt)
        c.check(c.lib.ncpa_field2d(c.h, self.h, C.c_void_p(psi_rows.ptr), nrows, dr, n_r,
                                   1 if lossless else 0, precision, C.c_void_p(out.ptr)))
        return out

    def field2d_stage(self, psi_rows, nrows, dr, n_r, precision=0, lossless=False):
        """Queue the operand staging of the 2-D field on the solving context; the returned
        FieldStage is contracted with FieldStage.contract (on any context that waited for an
        event recorded after this call) and freed once that contraction has completed."""
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.ncpa_field2d_stage(self.ctx.h, self.h, C.c_void_p(psi_rows.ptr), nrows, dr,
                                                       n_r, 1 if lossless else 0, precision, C.byref(h)))
        return FieldStage(self.ctx.lib, h, self.P * nrows * 2 * n_r, precision)

    def field2d(self, row0, stride, nrows, dr, n_r, precision=0, lossless=False):
        """Host copy of the complex field [P, nrows, n_r]."""
        rows = self.rows_device(row0, stride, nrows)
        out = self.field2d_device(rows, nrows, dr, n_r, precision, lossless)
        a = out.numpy().reshape(self.P, nrows, n_r, 2)
        out.free()
        rows.free()
        return a[..., 0] + 1j * a[..., 1]

    def tl1d(self, dr, n_r, lossless=False):
        """1-D TL of every problem: (p complex [P, n_r], incoherent [P, n_r]) on the host."""
        n = self.P * n_r
        bufs = [self.ctx.empty(n) for _ in range(3)]
        self.tl1d_device(dr, n_r, lossless, *(b.ptr for b in bufs))
        re, im, inc = (b.numpy().reshape(self.P, n_r) for b in bufs)
        for b in bufs:
            b.free()
        return re + 1j * im, inc

    def tl1d_device(self, dr, n_r, lossless, p_re, p_im, incoh):
        """Device pointers (ints) for the three outputs, each n_problems * n_r doubles."""
        self.ctx.check(self.ctx.lib.ncpa_tl1d(self.ctx.h, self.h, dr, n_r, 1 if lossless else 0,
                                              C.c_void_p(p_re), C.c_void_p(p_im), C.c_void_p(incoh)))
