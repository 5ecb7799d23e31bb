"""ncpa_bb_pulse_noalloc (every intermediate in the library's fixed device array, nothing
allocated) against ncpa_bb_pulse on the same device inputs: two calls in a row, the second
smaller; the FFT's global work buffer (NFFT above the LDS) inside the fixed array; jobs larger
than the array rejected with the context left usable."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _inputs(ctx, seed, n_sets, n_ranges, n_freq):
    """Device transfer functions H[n_sets][n_ranges][n_freq] (complex), host spectrum, f_vec
    with the DC bin first, ranges 200.. km."""
    rng = np.random.default_rng(seed)
    H = (rng.standard_normal((n_sets, n_ranges, n_freq)) + 1j * rng.standard_normal((n_sets, n_ranges, n_freq)))
    H *= 1e-3
    dH = ctx.empty(2 * H.size)
    dH.upload(np.ascontiguousarray(H).view(np.float64).ravel())
    S = np.ascontiguousarray(rng.standard_normal(n_freq) + 1j * rng.standard_normal(n_freq))
    f = np.arange(n_freq, dtype=np.float64) * (2.0 / 1024)
    ranges = np.linspace(200e3, 1000e3, n_ranges)
    return dH, S.view(np.float64).copy(), f, ranges


def _run(ctx, name, dH, S, f, ranges, n_sets, nfft, max_cel=300.0):
    W = ctx.empty(n_sets * ranges.size * nfft)
    rc = getattr(ctx.lib, name)(ctx.h, n_sets, ranges.size, f.size, C.c_void_p(dH.ptr), _dp(S), _dp(f), _dp(ranges),
                                max_cel, nfft, C.c_void_p(W.ptr))
    ctx.check(rc)
    out = W.numpy()
    W.free()
    return out


# (n_sets, n_ranges, n_freq, nfft): the broadband job's shape (8 azimuths, 16 ranges, 1024
# frequencies + DC, NFFT 4100), a smaller one, and NFFT 16384 (ping-pong buffers beyond LDS)
CASES = [(8, 16, 1025, 4100), (2, 3, 257, 1028), (2, 2, 1025, 16384)]


def test_noalloc_equals_pulse(gpu_ctx):
    for seed, (ns, nr, nf, nfft) in enumerate(CASES):
        print("case pulse_noalloc sets=%d ranges=%d freq=%d nfft=%d" % (ns, nr, nf, nfft), flush=True)
        dH, S, f, ranges = _inputs(gpu_ctx, seed, ns, nr, nf)
        ref = _run(gpu_ctx, "ncpa_bb_pulse", dH, S, f, ranges, ns, nfft)
        got = _run(gpu_ctx, "ncpa_bb_pulse_noalloc", dH, S, f, ranges, ns, nfft)
        assert np.all(np.isfinite(ref)) and np.max(np.abs(ref)) > 0
        # the same kernels on the same inputs: equal to the bit
        assert np.array_equal(got, ref)
        dH.free()


def test_noalloc_consecutive_calls(gpu_ctx):
    """Two calls in a row without a host wait between them, the second smaller: each equals
    ncpa_bb_pulse of its own job (the second reads nothing the first left behind)."""
    big, small = CASES[0], CASES[1]
    print("case pulse_noalloc consecutive", flush=True)
    jobs = [(_inputs(gpu_ctx, 10 + i, c[0], c[1], c[2]), c) for i, c in enumerate((big, small))]
    outs = []
    for (dH, S, f, ranges), c in jobs:
        W = gpu_ctx.empty(c[0] * c[1] * c[3])
        gpu_ctx.check(gpu_ctx.lib.ncpa_bb_pulse_noalloc(gpu_ctx.h, c[0], c[1], c[2], C.c_void_p(dH.ptr), _dp(S),
                                                        _dp(f), _dp(ranges), 300.0, c[3], C.c_void_p(W.ptr)))
        outs.append(W)
    for ((dH, S, f, ranges), c), W in zip(jobs, outs):
        got = W.numpy()
        ref = _run(gpu_ctx, "ncpa_bb_pulse", dH, S, f, ranges, c[0], c[3])
        assert np.array_equal(got, ref)
        W.free()
        dH.free()


def test_noalloc_rejects_oversize(gpu_ctx):
    """64 sets x 64 ranges x NFFT 4100 needs ~537 MB of intermediates: refused before any
    launch (the output is left untouched), after which the context works as before."""
    from ncpamodes import _lib
    ns, nr, nf, nfft = 64, 64, 1025, 4100
    print("case pulse_noalloc oversize", flush=True)
    dH, S, f, ranges = _inputs(gpu_ctx, 99, ns, nr, nf)
    W = gpu_ctx.empty(ns * nr * nfft)
    W.upload(np.full(ns * nr * nfft, 7.0))
    rc = gpu_ctx.lib.ncpa_bb_pulse_noalloc(gpu_ctx.h, ns, nr, nf, C.c_void_p(dH.ptr), _dp(S), _dp(f), _dp(ranges),
                                           300.0, nfft, C.c_void_p(W.ptr))
    assert rc == _lib.NCPA_ERR_ARG
    msg = gpu_ctx.lib.ncpa_ctx_error(gpu_ctx.h).decode()
    assert "maximum of %d" % _lib.NCPA_PULSE_SCRATCH_BYTES in msg, msg
    assert np.all(W.numpy() == 7.0)
    W.free()
    dH.free()
    ns, nr, nf, nfft = CASES[1]
    dH, S, f, ranges = _inputs(gpu_ctx, 100, ns, nr, nf)
    assert np.array_equal(_run(gpu_ctx, "ncpa_bb_pulse_noalloc", dH, S, f, ranges, ns, nfft),
                          _run(gpu_ctx, "ncpa_bb_pulse", dH, S, f, ranges, ns, nfft))
    dH.free()
