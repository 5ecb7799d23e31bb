"""Broadband pulse synthesis over the ncpamodes C ABI -- host mirror of ModBB --propagation.

``ModalBroadbandPropagator`` follows the reference class of the same name
(src/libbroadband/ModalBroadbandPropagator.cpp:25-277, BroadbandPropagator.cpp:244-662):
read a dispersion file (DC bin prepended), build the source spectrum, form the transfer
function at each receiver range and FFT the propagated pulse.  Transfer functions, FFTs and
the time-phase product run in libncpamodes.so; this module only parses text and moves
arrays.  ``synthesize_from_modes`` is the fused path: device-resident modes of a batched
dispersion solve go straight into the transfer-function kernel without a text round trip.
"""
import ctypes as C
import math

import numpy as np

from ._lib import NCPA_SRC, BbJob, Problem


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def read_dispersion(path):
    """ModalBroadbandPropagator::read_dispersion_file (ModalBroadbandPropagator.cpp:98-138):
    f (DC first), counts, flat mode arrays k_re, k_im, psi_S, psi_R and rho_s/rho_r of the
    last line read."""
    f, counts, rows = [0.0], [0], []
    rho_s = rho_r = None
    with open(path) as fh:
        tok = fh.read().split()
    pos = 0
    while pos < len(tok):
        f.append(float(tok[pos]))
        n = int(tok[pos + 1])
        rho_s, rho_r = float(tok[pos + 2]), float(tok[pos + 3])
        pos += 4
        if n > 4000:
            raise ValueError("%s: more than MAX_MODES = 4000 modes on one line" % path)
        v = np.array(tok[pos:pos + 4 * n], dtype=float)
        if v.size != 4 * n:
            raise ValueError("%s: read_dispersion_data format error" % path)
        rows.append(v.reshape(n, 4))
        counts.append(n)
        pos += 4 * n
    q = np.concatenate(rows) if rows else np.zeros((0, 4))
    return {"f": np.array(f), "counts": np.array(counts, dtype=np.int32), "k_re": q[:, 0].copy(),
            "k_im": q[:, 1].copy(), "psi_s": q[:, 2].copy(), "psi_r": q[:, 3].copy(),
            "rho_s": rho_s, "rho_r": rho_r}


def fft(ctx, x, sign=-1):
    """Unnormalised batched complex FFT of the rows of x on the device (sign -1 forward)."""
    x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex128)
    batch, n = x.shape
    a = ctx.empty(2 * x.size)
    b = ctx.empty(2 * x.size)
    a.upload(x.view(np.float64).ravel())
    ctx.check(ctx.lib.ncpa_fft_batch(ctx.h, C.c_void_p(a.ptr), C.c_void_p(b.ptr), n, batch, sign))
    out = b.numpy().view(np.complex128).reshape(batch, n)
    a.free()
    b.free()
    return out


def source(ctx, kind, f, nfft, f_center=-1.0, rows=None):
    """Source spectrum S[Nfreq] and time-domain source pulse[NFFT]
    (BroadbandPropagator::get_source_spectrum, BroadbandPropagator.cpp:244-417).
    rows: (n, 3) spectrum rows (f, Re, Im) or (n, 2) waveform rows (t, amplitude)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    S = np.zeros(2 * f.size)
    pulse = np.zeros(2 * nfft)
    if rows is not None:
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        n_in, inp = rows.shape[0], _dp(rows)
    else:
        n_in, inp = 0, None
    ctx.check(ctx.lib.ncpa_bb_source(ctx.h, NCPA_SRC[kind], _dp(f), f.size, nfft, f_center, inp, n_in,
                                     _dp(S), _dp(pulse)))
    return S.view(np.complex128), pulse.view(np.complex128)


def transfer_device(ctx, n_sets, n_freq, begin, count, kre, kim, psr, ranges_m, rho_ratio, zero_attn=False):
    """Transfer functions [n_sets, n_ranges, n_freq] into a device buffer (complex interleaved)."""
    begin = np.ascontiguousarray(begin, dtype=np.int32)
    count = np.ascontiguousarray(count, dtype=np.int32)
    ranges_m = np.ascontiguousarray(ranges_m, dtype=np.float64)
    rho_ratio = np.ascontiguousarray(np.broadcast_to(rho_ratio, (n_sets,)), dtype=np.float64)
    H = ctx.empty(2 * n_sets * ranges_m.size * n_freq)
    ctx.check(ctx.lib.ncpa_bb_transfer(ctx.h, n_sets, n_freq, _ip(begin), _ip(count), C.c_void_p(kre),
                                       C.c_void_p(kim), C.c_void_p(psr), 1 if zero_attn else 0,
                                       _dp(ranges_m), ranges_m.size, _dp(rho_ratio), C.c_void_p(H.ptr)))
    return H


def pulse_device(ctx, n_sets, H, S, f, ranges_m, max_cel, nfft):
    """Waveforms 2 Re(df FFT(e^{-2 pi i f t0} S H)) [n_sets, n_ranges, NFFT] on the device."""
    ranges_m = np.ascontiguousarray(ranges_m, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.complex128)
    f = np.ascontiguousarray(f, dtype=np.float64)
    W = ctx.empty(n_sets * ranges_m.size * nfft)
    ctx.check(ctx.lib.ncpa_bb_pulse(ctx.h, n_sets, ranges_m.size, f.size, C.c_void_p(H.ptr),
                                    _dp(S.view(np.float64)), _dp(f), _dp(ranges_m), max_cel, nfft,
                                    C.c_void_p(W.ptr)))
    return W


def default_nfft(nfft, nfreq):
    """NFFT = max(nfft, 4 Nfreq), Nfreq counting DC (ModalBroadbandPropagator.cpp:71-74)."""
    return 4 * nfreq if nfft < 4 * nfreq else nfft


def receiver_ranges(receiver="single", range_km=0.0, start_range_km=0.0, end_range_km=0.0,
                    range_step_km=0.0):
    """Receiver ranges in m (ModalBroadbandPropagator.cpp:37-54)."""
    if receiver == "single":
        return np.array([range_km * 1000.0])
    if end_range_km <= start_range_km:
        raise ValueError("end_range_km must be greater than start_range_km")
    n = int(math.floor((end_range_km - start_range_km) / range_step_km)) + 1
    return (start_range_km + np.arange(n) * range_step_km) * 1000.0


class ModalBroadbandPropagator:
    """Host mirror of NCPA::ModalBroadbandPropagator on the GPU path."""

    def __init__(self, ctx, dispersion, ranges_m, source="impulse", nfft=0, max_celerity=340.0,
                 f_center=-1.0, source_rows=None, zero_attn=False):
        self.ctx = ctx
        self.disp = read_dispersion(dispersion) if isinstance(dispersion, str) else dispersion
        self.ranges = np.asarray(ranges_m, dtype=np.float64)
        f = self.disp["f"]
        self.f_step = f[2] - f[1]
        self.nfft = default_nfft(nfft, f.size)
        if source == "waveform":
            while source_rows.shape[0] > self.nfft:
                self.nfft *= 2
        self.source, self.f_center, self.rows = source, f_center, source_rows
        self.max_cel = max_celerity
        self.zero_attn = zero_attn

    def calculate_waveform(self):
        """Returns (times [nr, NFFT], waveform [nr, NFFT]) as ModBB writes them."""
        ctx, d = self.ctx, self.disp
        f = d["f"]
        S, self.source_pulse = source(ctx, self.source, f, self.nfft, self.f_center, self.rows)
        self.spectrum = S
        counts = d["counts"].astype(np.int32)
        begin = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int32)
        n = max(int(counts.sum()), 1)
        bufs = [ctx.empty(n) for _ in range(3)]
        for b, a in zip(bufs, (d["k_re"], d["k_im"], d["psi_s"] * d["psi_r"])):
            b.upload(np.resize(a, n) if a.size else np.zeros(n))
        H = transfer_device(ctx, 1, f.size, begin, counts, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                            self.ranges, d["rho_r"] / d["rho_s"], self.zero_attn)
        W = pulse_device(ctx, 1, H, S, f, self.ranges, self.max_cel, self.nfft)
        wave = W.numpy().reshape(self.ranges.size, self.nfft)
        for b in bufs + [H, W]:
            b.free()
        fmx = self.nfft * self.f_step
        t0 = self.ranges / self.max_cel
        times = np.arange(self.nfft)[None, :] / fmx + t0[:, None]
        return times, wave


def synthesize_from_modes(ctx, modes, n_sets, f, ranges_m, rho_ratio, S, max_cel, nfft):
    """Fused broadband path: problem p = s * (Nfreq - 1) + (i - 1) of ``modes`` holds the modes of
    set s at positive frequency f[i] (f[0] = 0 is the DC bin with no modes).  Returns the
    device waveform buffer [n_sets, n_ranges, NFFT] and the device k_pert buffers."""
    nf = f.size
    assert modes.P == n_sets * (nf - 1)
    n = max(modes.total, 1)
    kre, kim = ctx.empty(n), ctx.empty(n)
    offs, k2, ab, psr = (C.c_void_p() for _ in range(4))
    ctx.check(ctx.lib.ncpa_modes_device(modes.h, C.byref(offs), C.byref(k2), C.byref(ab), C.byref(psr)))
    ctx.check(ctx.lib.ncpa_kpert(ctx.h, k2, ab, modes.total, C.c_void_p(kre.ptr), C.c_void_p(kim.ptr)))
    begin = np.zeros((n_sets, nf), dtype=np.int32)
    count = np.zeros((n_sets, nf), dtype=np.int32)
    begin[:, 1:] = modes.offsets[:-1].reshape(n_sets, nf - 1)
    count[:, 1:] = modes.counts.reshape(n_sets, nf - 1)
    H = transfer_device(ctx, n_sets, nf, begin.ravel(), count.ravel(), kre.ptr, kim.ptr, psr.value,
                        ranges_m, rho_ratio)
    W = pulse_device(ctx, n_sets, H, S, f, ranges_m, max_cel, nfft)
    H.free()
    return W, (kre, kim)


class BroadbandPlan:
    """A whole broadband job on the device as one captured HIP graph (ncpa_bb_plan_*).

    problems[s * (Nfreq - 1) + i - 1] is set s (e.g. an azimuth) at f[i] (f[0] = 0 is DC).
    run() replays the graph (dispersion solve of every problem, transfer functions at the
    ranges, pulse synthesis); wave() waits and returns the waveforms [n_sets, n_ranges, NFFT].
    """

    def __init__(self, ctx, grid, problems, n_sets, f, S, ranges_m, rho_ratio, max_cel, nfft,
                 zero_attn=False, max_modes_total=0, tune=True):
        self.ctx = ctx
        f = np.ascontiguousarray(f, dtype=np.float64)
        S = np.ascontiguousarray(S, dtype=np.complex128).view(np.float64)
        ranges_m = np.ascontiguousarray(ranges_m, dtype=np.float64)
        rho = np.ascontiguousarray(np.broadcast_to(rho_ratio, (n_sets,)), dtype=np.float64)
        self._keep = (f, S, ranges_m, rho)
        job = BbJob(n_sets, f.size, _dp(f), _dp(S), _dp(ranges_m), ranges_m.size, _dp(rho), max_cel, nfft,
                    1 if zero_attn else 0, max_modes_total, 1 if tune else 0)
        self.shape = (n_sets, ranges_m.size, nfft)
        self.P = n_sets * (f.size - 1)
        if len(problems) != self.P:
            raise ValueError("a plan needs n_sets * (len(f) - 1) problems")
        arr = (Problem * self.P)(*problems)
        h = C.c_void_p()
        ctx.check(ctx.lib.ncpa_bb_plan_create(ctx.h, C.byref(grid), arr, C.byref(job), C.byref(h)))
        self.h = h

    def set_problems(self, problems):
        arr = (Problem * self.P)(*problems)
        self.ctx.check(self.ctx.lib.ncpa_bb_plan_set_problems(self.h, arr))

    def run(self):
        self.ctx.check(self.ctx.lib.ncpa_bb_plan_run(self.h))

    def result(self):
        """(device pointers dict, total modes) of the last run; waits for it."""
        ptr = [C.c_void_p() for _ in range(5)]
        total = C.c_int()
        self.ctx.check(self.ctx.lib.ncpa_bb_plan_result(self.h, *[C.byref(p) for p in ptr], C.byref(total)))
        names = ("wave", "kre", "kim", "psi_sr", "offsets")
        return {n: p.value for n, p in zip(names, ptr)}, total.value

    def wave(self):
        ptrs, _ = self.result()
        out = np.empty(self.shape)
        self.ctx.check(self.ctx.lib.ncpa_dev_download(self.ctx.h, out.ctypes.data_as(C.c_void_p),
                                                      C.c_void_p(ptrs["wave"]), out.nbytes))
        return out

    def modes(self):
        """Host copies (offsets, k_re, k_im, psi_sr) of the last run."""
        ptrs, total = self.result()
        offs = np.empty(self.P + 1, dtype=np.int32)
        arrs = [np.empty(total) for _ in range(3)]
        lib, h = self.ctx.lib, self.ctx.h
        self.ctx.check(lib.ncpa_dev_download(h, offs.ctypes.data_as(C.c_void_p), C.c_void_p(ptrs["offsets"]),
                                             offs.nbytes))
        for a, n in zip(arrs, ("kre", "kim", "psi_sr")):
            if total:
                self.ctx.check(lib.ncpa_dev_download(h, a.ctypes.data_as(C.c_void_p), C.c_void_p(ptrs[n]), a.nbytes))
        return offs, arrs[0], arrs[1], arrs[2]

    def tuning(self):
        ms, blocks, tms, n = C.c_int(), C.c_int(), C.c_double(), C.c_int(64)
        log = np.zeros(3 * 64)
        self.ctx.check(self.ctx.lib.ncpa_bb_plan_tuning(self.h, C.byref(ms), C.byref(blocks), C.byref(tms),
                                                         _dp(log), C.byref(n)))
        return {"ms_passes": ms.value, "grid_blocks": blocks.value, "modes_ms": tms.value,
                "log": log[:3 * n.value].reshape(-1, 3).tolist()}

    def close(self):
        if self.h:
            self.ctx.lib.ncpa_bb_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
