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

    def reset_timing(self):
        self.check(self.lib.ncpa_ctx_reset_timing(self.h))

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


class DeviceArray:
    """A flat device buffer allocated through the C ABI (no torch: see DESIGN.md)."""

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

    def problem_modes(self, p):
        k, psr, ps, pr = self.download()
        s = slice(self.offsets[p], self.offsets[p + 1])
        return k[s], psr[s], ps[s], pr[s]

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
