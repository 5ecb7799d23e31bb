"""Broadband pulse synthesis from dispersion data -- CPU oracle, TEST INFRASTRUCTURE ONLY.

FP64 numpy restatement of ncpaprop's ModBB ``--propagation`` path
(src/libbroadband/ModalBroadbandPropagator.cpp, src/libbroadband/BroadbandPropagator.cpp).
numpy.fft.fft is FFTW_FORWARD (e^{-2 pi i jk/N}, unnormalised) and N * numpy.fft.ifft is
FFTW_BACKWARD (BroadbandPropagator.cpp:234-239).

Pinned by (tests/test_oracle_broadband.py):
  * source_spectrum.dat (reference golden, pulse1 at f_center = f_max/5): frequency shape
    and phase to 1e-6 relative; the absolute scale of that older build differs by one
    constant factor (1.619126), recorded in the test.
  * mywavf.dat (reference golden, pulse1, 270 km, max celerity 320 m/s, NFFT 16384, from
    myDispersionFile.dat): the propagated waveform.
  * closed forms: the half-Hann taper end points, a single mode's transfer function,
    pulse2's spectrum against direct quadrature, FFT against a brute-force DFT.
"""
import math

import numpy as np

PI = 3.141592653589793


def read_dispersion(path):
    """ModalBroadbandPropagator::read_dispersion_file (ModalBroadbandPropagator.cpp:98-138).

    Returns a dict: f (with the DC bin prepended, f[0] = 0), counts (counts[0] = 0),
    modes = list of (n, 4) arrays [Re k, Im k, psi_S, psi_R] per frequency (empty at DC),
    rho_s, rho_r taken from the LAST line read (as the reference does)."""
    f = [0.0]
    counts = [0]
    modes = [np.zeros((0, 4))]
    rho_s = rho_r = None
    with open(path) as fh:
        tokens = fh.read().split()
    pos = 0
    while pos < len(tokens):
        f.append(float(tokens[pos]))
        n = int(tokens[pos + 1])
        rho_s = float(tokens[pos + 2])
        rho_r = float(tokens[pos + 3])
        pos += 4
        vals = np.array(tokens[pos:pos + 4 * n], dtype=float)
        if vals.size != 4 * n:
            raise ValueError("%s: read_dispersion_data format error" % path)
        modes.append(vals.reshape(n, 4))
        counts.append(n)
        pos += 4 * n
    return {"f": np.array(f), "counts": np.array(counts, dtype=int), "modes": modes,
            "rho_s": rho_s, "rho_r": rho_r}


def default_nfft(nfft, nfreq):
    """NFFT = max(nfft, 4 Nfreq) with Nfreq counting DC (ModalBroadbandPropagator.cpp:71-74)."""
    return 4 * nfreq if nfft < 4 * nfreq else nfft


def half_hann(begin, end, i):
    """BroadbandPropagator::half_hann (BroadbandPropagator.cpp:654-662)."""
    if i < begin:
        return 1.0
    if begin <= i <= end:
        if end == begin:
            # the reference evaluates 0.0/0 here (NaN); the build tapers that bin to 0
            return 0.0
        return 0.5 * (math.cos(PI * (i - begin) / (end - begin)) + 1.0)
    return 0.0


def transfer_function(disp, r):
    """ModalBroadbandPropagator::compute_modal_sum (ModalBroadbandPropagator.cpp:245-277):
    H_i = i e^{-i pi/4}/sqrt(8 pi r) sqrt(rho_r/rho_s) sum_j psi_S psi_R e^{i k r}/sqrt(k),
    then the half-Hann taper on the top floor(0.1 Nfreq) bins."""
    f = disp["f"]
    nf = f.size
    H = np.zeros(nf, dtype=complex)
    pref = 1j * np.exp(-1j * PI * 0.25) / math.sqrt(8.0 * PI * r) * math.sqrt(disp["rho_r"] / disp["rho_s"])
    for i in range(nf):
        m = disp["modes"][i]
        if m.shape[0] == 0:
            continue
        k = m[:, 0] + 1j * m[:, 1]
        H[i] = pref * np.sum(m[:, 2] * m[:, 3] * np.exp(1j * r * k) / np.sqrt(k))
    smooth = int(math.floor(0.1 * nf))
    for i in range(nf - smooth, nf):
        H[i] *= half_hann(nf - smooth, nf - 1, i)
    return H


def pulse_spec_fit(scale, x):
    """BroadbandPropagator::pulse_spec_fit (BroadbandPropagator.cpp:420-438)."""
    fnorm = 0.780776406404415
    x = scale * fnorm * np.asarray(x, dtype=float)
    ans = x * np.exp((-x - x * x) / 2) * np.exp(1j * (-0.5 * PI)) * scale / 33.4336
    return 4.0 * PI * 5.0 * ans * np.exp(1j * 2.0 * PI * x * 0.01)


def model_pulse_shape(power, scale, t):
    """BroadbandPropagator::model_pulse_shape (BroadbandPropagator.cpp:469-476)."""
    x = 2.094 * (1.0 + power) * np.asarray(t, dtype=float) / scale
    return np.power(x, power) * (1 - x / (1.0 + power)) * np.exp(-x) / 1.387566e+03


def pulse_scale(f, f_center):
    """f_center default f_max/5 and its check (BroadbandPropagator.cpp:270-283)."""
    if f_center < 0.0:
        f_center = f[-1] / 5.0
    if f_center > 0 and f_center <= f[-1] / 5.0 + 1.0e-16:
        return 1.0 / f_center
    raise ValueError("f_center = %g Hz is too large" % f_center)


def source_spectrum(kind, f, nfft, f_step, f_center=-1.0, spectrum=None, waveform=None):
    """BroadbandPropagator::get_source_spectrum (BroadbandPropagator.cpp:244-417).

    Returns (S[Nfreq], pulse_vec[NFFT]) -- the source spectrum at the dispersion
    frequencies and the time-domain source pulse the reference writes to
    source_waveform_input.dat (before its factor 2 / 1)."""
    nf = f.size
    fmx = nfft * f_step
    dt = 1.0 / fmx
    if kind == "impulse":
        S = np.exp(1j * 2.0 * PI * f * (nf / 2.0))
    elif kind == "pulse1":
        scale = pulse_scale(f, f_center)
        S = pulse_spec_fit(scale, f) * np.exp(1j * 2.0 * PI * f * scale) / 2.0
    elif kind == "pulse2":
        scale = pulse_scale(f, f_center)
        # model_pulse_fft (BroadbandPropagator.cpp:440-466): FFTW_BACKWARD * dt of the
        # pulse sampled at i dt, i < NFFT (the reference's one-element allocation there
        # is a heap overflow; the intended NFFT-long buffer is used)
        p = model_pulse_shape(8.0, scale, np.arange(nfft) * dt)
        S = (nfft * np.fft.ifft(p))[:nf] * dt
    elif kind == "spectrum":
        fs, S = spectrum
        if abs(fs[0] - f[0]) > 1.0e-8:
            raise ValueError("source spectrum frequencies do not match the dispersion file")
        S0 = np.asarray(S, dtype=complex)[:nf]
        S = np.zeros(nf, dtype=complex)
        S[:S0.size] = S0
    elif kind == "waveform":
        t, amp = waveform
        dtw = t[1] - t[0]
        pv = np.zeros(nfft, dtype=complex)
        pv[:len(amp)] = amp
        S = (nfft * np.fft.ifft(pv))[:nf] * dtw
        return S, pv
    else:
        raise ValueError("unknown source type " + kind)
    arg = np.zeros(nfft, dtype=complex)
    arg[:nf] = S
    pulse = f_step * np.fft.fft(arg)
    return S, pulse


def pulse_prop(H, S, f, t0, nfft):
    """BroadbandPropagator::fft_pulse_prop (BroadbandPropagator.cpp:573-651):
    df * FFTW_FORWARD of e^{-2 pi i f t0} S H zero padded to NFFT, df = f[1] - f[0]."""
    nf = f.size
    if nfft < nf:
        raise ValueError("fft too short (i.e. NFFT < Nfreq)")
    df = f[1] - f[0]
    arg = np.zeros(nfft, dtype=complex)
    arg[:nf] = np.exp(-1j * 2.0 * PI * f * t0) * S * H
    return df * np.fft.fft(arg)


def waveforms(disp, ranges_m, kind="impulse", nfft=0, max_cel=340.0, f_center=-1.0, spectrum=None,
              waveform=None):
    """ModalBroadbandPropagator::calculate_waveform (ModalBroadbandPropagator.cpp:141-242).

    Returns (nfft, times[nr][NFFT], wave[nr][NFFT]) with wave = 2 Re(pulse) as written."""
    f = disp["f"]
    f_step = f[2] - f[1]
    nfft = default_nfft(nfft, f.size)
    if kind == "waveform":
        while len(waveform[0]) > nfft:
            nfft *= 2
    fmx = nfft * f_step
    S, _ = source_spectrum(kind, f, nfft, f_step, f_center, spectrum, waveform)
    times, waves = [], []
    for r in ranges_m:
        t0 = r / max_cel
        H = transfer_function(disp, r)
        p = pulse_prop(H, S, f, t0, nfft)
        times.append(np.arange(nfft) / fmx + t0)
        waves.append(2.0 * p.real)
    return nfft, np.array(times), np.array(waves)


def disp_from_arrays(f_pos, counts, q, rho_s, rho_r):
    """Dispersion dict (as read_dispersion returns) from parsed arrays of the positive
    frequencies: q[total, 4] = Re k, Im k, psi_S, psi_R in file order."""
    modes = [np.zeros((0, 4))]
    o = 0
    for n in counts:
        modes.append(np.asarray(q[o:o + n]))
        o += n
    return {"f": np.concatenate([[0.0], f_pos]), "counts": np.concatenate([[0], counts]).astype(int),
            "modes": modes, "rho_s": float(rho_s), "rho_r": float(rho_r)}
