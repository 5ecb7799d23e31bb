"""Broadband oracle (oracle/broadband.py) pinned against the reference's ModBB goldens and
closed forms.  CPU only."""
import math
import os

import numpy as np
import pytest

from oracle import broadband as bb

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# The manual's goldens come from an older ModBB whose built-in pulse was 1.619126 times the
# current pulse_spec_fit (same shape and phase) and whose waveform writer did not apply the
# factor 2 of ModalBroadbandPropagator.cpp:223-226.
OLD_PULSE_SCALE = 1.619126


def golden_disp():
    z = np.load(os.path.join(GOLD, "disp250.npz"))
    return bb.disp_from_arrays(z["f"], z["counts"], z["q"], z["rho_s"], z["rho_r"])


def test_source_spectrum_golden_shape():
    """pulse1 at f_center = f_max/5 = 0.1 Hz (scale 10 s): ratio to the golden is one real
    constant at every frequency (shape and phase pinned to the golden's 7 digits)."""
    g = np.loadtxt(os.path.join(GOLD, "source_spectrum.dat"))
    disp = golden_disp()
    f = disp["f"]
    assert np.allclose(f[1:], g[:, 0], rtol=0, atol=1e-9)
    S, _ = bb.source_spectrum("pulse1", f, 16384, f[2] - f[1])
    ratio = (g[:, 1] + 1j * g[:, 2]) / S[1:]
    assert np.max(np.abs(ratio - OLD_PULSE_SCALE)) < 5e-6 * OLD_PULSE_SCALE
    # a wrong delay (scale) or sign in the exponent would rotate the phase with f
    assert np.max(np.abs(np.angle(ratio))) < 1e-5


def test_waveform_golden():
    """Propagated pulse1 waveform at 270 km (max celerity 320 m/s, NFFT 16384) against the
    reference's mywavf.dat: same samples and times, amplitude ratio OLD_PULSE_SCALE / 2,
    residual below 5e-5 of the peak (the golden prints 7 digits)."""
    g = np.loadtxt(os.path.join(GOLD, "mywavf_880_960.dat"))
    disp = golden_disp()
    nfft, t, w = bb.waveforms(disp, [270e3], "pulse1", nfft=16384, max_cel=320.0)
    assert nfft == 16384
    i0 = int(round((g[0, 0] - t[0, 0]) * nfft * (disp["f"][2] - disp["f"][1])))
    sel = slice(i0, i0 + g.shape[0])
    assert np.max(np.abs(t[0, sel] - g[:, 0])) < 1e-6
    ours = w[0, sel] * (OLD_PULSE_SCALE / 2.0)
    peak = np.max(np.abs(g[:, 1]))
    assert np.max(np.abs(ours - g[:, 1])) < 5e-5 * peak


def test_default_nfft_and_df_rule():
    disp = golden_disp()
    assert disp["f"].size == 251 and disp["f"][0] == 0.0
    assert bb.default_nfft(0, 251) == 1004
    assert bb.default_nfft(16384, 251) == 16384
    assert bb.default_nfft(4096, 1025) == 4100


def test_half_hann():
    assert bb.half_hann(10, 20, 9) == 1.0
    assert bb.half_hann(10, 20, 10) == 1.0
    assert abs(bb.half_hann(10, 20, 15) - 0.5) < 1e-15
    assert abs(bb.half_hann(10, 20, 20)) < 1e-15
    assert bb.half_hann(10, 20, 21) == 0.0
    w = [bb.half_hann(10, 20, i) for i in range(10, 21)]
    assert all(a >= b for a, b in zip(w, w[1:]))


def test_single_mode_transfer_closed_form():
    k = 0.02 + 3e-7j
    disp = bb.disp_from_arrays(np.array([0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1]),
                               np.ones(11, dtype=int),
                               np.tile([k.real, k.imag, 0.01, -0.02], (11, 1)), 1.2, 0.9)
    r = 123456.0
    H = bb.transfer_function(disp, r)
    expect = (np.exp(1j * math.pi / 4) / math.sqrt(8 * math.pi * r) * math.sqrt(0.9 / 1.2)
              * (0.01 * -0.02) * np.exp(1j * k * r) / np.sqrt(k))
    assert H[0] == 0
    assert np.allclose(H[1:11], expect, rtol=1e-13, atol=0)
    assert H[11] == 0    # top bin tapered to zero (floor(0.1 * 12) = 1 bin: end == begin)


def test_fft_is_fftw_convention():
    rng = np.random.default_rng(1)
    x = rng.normal(size=41) + 1j * rng.normal(size=41)
    n = np.arange(41)
    dft = np.exp(-2j * np.pi * np.outer(n, n) / 41) @ x
    assert np.allclose(np.fft.fft(x), dft, rtol=1e-12, atol=1e-12)
    assert np.allclose(41 * np.fft.ifft(x), np.conj(np.exp(-2j * np.pi * np.outer(n, n) / 41)) @ x,
                       rtol=1e-12, atol=1e-12)


def test_pulse2_spectrum_is_fourier_integral():
    """pulse2's S(f) = int p(t) e^{+2 pi i f t} dt (FFTW_BACKWARD times dt)."""
    from scipy.integrate import quad
    f = np.arange(0, 257) * (2.0 / 256)
    nfft = 4 * f.size
    S, _ = bb.source_spectrum("pulse2", f, nfft, f[2] - f[1])
    scale = 1.0 / (f[-1] / 5)
    T = 1.0 / (f[2] - f[1])
    for i in (3, 17, 40):
        re = quad(lambda t: bb.model_pulse_shape(8.0, scale, t) * math.cos(2 * math.pi * f[i] * t), 0, T,
                  limit=400)[0]
        im = quad(lambda t: bb.model_pulse_shape(8.0, scale, t) * math.sin(2 * math.pi * f[i] * t), 0, T,
                  limit=400)[0]
        assert abs(S[i] - (re + 1j * im)) < 2e-3 * abs(S).max()


def test_impulse_and_spectrum_sources():
    f = np.arange(0, 11) * 0.01
    S, p = bb.source_spectrum("impulse", f, 44, 0.01)
    assert np.allclose(np.abs(S), 1.0)
    assert np.allclose(S, np.exp(2j * np.pi * f * 5.5))
    rows_f = f.copy()
    Sx = np.arange(11) * (1 + 2j)
    S2, _ = bb.source_spectrum("spectrum", f, 44, 0.01, spectrum=(rows_f, Sx))
    assert np.array_equal(S2, Sx)
    with pytest.raises(ValueError):
        bb.source_spectrum("spectrum", f, 44, 0.01, spectrum=(rows_f + 0.5, Sx))
