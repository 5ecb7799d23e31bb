"""GPU parity: broadband synthesis (FFT, source spectra, transfer functions, waveforms) through
the C ABI vs the FP64 oracle and the reference's ModBB golden waveform."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OLD_PULSE_SCALE = 1.619126   # see tests/test_oracle_broadband.py


def _golden_disp():
    from oracle import broadband as obb
    z = np.load(os.path.join(GOLD, "disp250.npz"))
    return obb.disp_from_arrays(z["f"], z["counts"], z["q"], z["rho_s"], z["rho_r"])


def _as_mirror(d):
    """oracle dispersion dict -> the flat arrays ncpamodes.broadband uses."""
    q = np.concatenate(d["modes"])
    return {"f": d["f"], "counts": d["counts"].astype(np.int32), "k_re": q[:, 0].copy(),
            "k_im": q[:, 1].copy(), "psi_s": q[:, 2].copy(), "psi_r": q[:, 3].copy(),
            "rho_s": d["rho_s"], "rho_r": d["rho_r"]}


@pytest.mark.parametrize("n", [4096, 4100, 16384, 1004, 41, 1155, 7 * 11 * 13, 2 * 3 * 5 * 7 * 9, 1])
def test_fft_vs_numpy(gpu_ctx, n):
    from ncpamodes import broadband as bb
    print("case fft n=%d" % n, flush=True)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    fwd = bb.fft(gpu_ctx, x, -1)
    ref = np.fft.fft(x, axis=1)
    assert np.max(np.abs(fwd - ref)) <= 1e-13 * np.sqrt(n) * np.max(np.abs(ref)) + 1e-300
    bwd = bb.fft(gpu_ctx, x, +1)
    refb = n * np.fft.ifft(x, axis=1)
    assert np.max(np.abs(bwd - refb)) <= 1e-13 * np.sqrt(n) * np.max(np.abs(refb)) + 1e-300


@pytest.mark.parametrize("kind", ["impulse", "pulse1", "pulse2", "spectrum", "waveform"])
def test_source_vs_oracle(gpu_ctx, kind):
    from ncpamodes import broadband as bb
    from oracle import broadband as obb
    print("case source %s" % kind, flush=True)
    f = np.arange(0, 1025) * (2.0 / 1024)
    nfft = 4100
    rows, spec, wave = None, None, None
    if kind == "spectrum":
        rng = np.random.default_rng(3)
        S0 = rng.normal(size=f.size) + 1j * rng.normal(size=f.size)
        rows = np.stack([f, S0.real, S0.imag], axis=1)
        spec = (f, S0)
    elif kind == "waveform":
        t = np.arange(600) * 0.05
        a = np.sin(2 * np.pi * 0.5 * t) * np.exp(-((t - 10.0) / 3.0) ** 2)
        rows = np.stack([t, a], axis=1)
        wave = (t, a)
    S, pulse = bb.source(gpu_ctx, kind, f, nfft, -1.0, rows)
    So, po = obb.source_spectrum(kind, f, nfft, f[2] - f[1], -1.0, spec, wave)
    # impulse: the oracle (like the reference) forms the delay phase 2 pi f Nfreq/2 (up to
    # 6.4e3 rad) unreduced, so its own rounding is ~1e-12; the library reduces mod 1 cycle first
    tol = 1e-11 if kind == "impulse" else 1e-12
    assert np.max(np.abs(S - So)) <= tol * np.max(np.abs(So))
    assert np.max(np.abs(pulse - po)) <= 1e-12 * np.max(np.abs(po))


def test_waveform_golden_dispersion(gpu_ctx):
    """ModBB --propagation on the reference's myDispersionFile.dat (pulse1, 270 km, max celerity
    320 m/s, NFFT 16384) vs the oracle (1e-10 of peak) and the reference's mywavf.dat."""
    from ncpamodes import broadband as bb
    from oracle import broadband as obb
    print("case golden waveform", flush=True)
    d = _golden_disp()
    prop = bb.ModalBroadbandPropagator(gpu_ctx, _as_mirror(d), [270e3], "pulse1", nfft=16384,
                                       max_celerity=320.0)
    t, w = prop.calculate_waveform()
    nfft, to, wo = obb.waveforms(d, [270e3], "pulse1", nfft=16384, max_cel=320.0)
    peak = np.max(np.abs(wo))
    assert np.max(np.abs(t - to)) < 1e-9
    assert np.max(np.abs(w - wo)) <= 1e-10 * peak
    g = np.loadtxt(os.path.join(GOLD, "mywavf_880_960.dat"))
    i0 = int(round((g[0, 0] - t[0, 0]) * 16384 * 0.002))
    ours = w[0, i0:i0 + g.shape[0]] * (OLD_PULSE_SCALE / 2.0)
    assert np.max(np.abs(ours - g[:, 1])) < 5e-5 * np.max(np.abs(g[:, 1]))


def test_multi_receiver_default_nfft(gpu_ctx):
    """--receiver multiple, impulse source, default NFFT (4 Nfreq = 1004) vs the oracle."""
    from ncpamodes import broadband as bb
    from oracle import broadband as obb
    print("case multi receiver", flush=True)
    d = _golden_disp()
    r = bb.receiver_ranges("multiple", start_range_km=220, end_range_km=300, range_step_km=20)
    assert r.size == 5
    prop = bb.ModalBroadbandPropagator(gpu_ctx, _as_mirror(d), r, "impulse", max_celerity=320.0)
    t, w = prop.calculate_waveform()
    nfft, to, wo = obb.waveforms(d, r, "impulse", max_cel=320.0)
    assert nfft == 1004 and w.shape == (5, 1004)
    assert np.max(np.abs(w - wo)) <= 1e-10 * np.max(np.abs(wo))


def test_fused_dispersion_to_waveform(gpu_ctx, synth_profile):
    """Batched dispersion solve (2 azimuths x 24 frequencies, Nz 4000) feeding the transfer
    functions on the device, vs the oracle's per-frequency solves -> compute_modal_sum -> FFT."""
    import ncpamodes
    from ncpamodes import broadband as bb
    from oracle import broadband as obb
    from oracle import ess
    print("case fused dispersion->waveform", flush=True)
    nfp, df, azs, Nz = 24, 0.02, [90.0, 250.0], 4000
    f = np.arange(nfp + 1) * df
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=Nz)
    probs = [ncpamodes.problem(pid, f[i], az) for az in azs for i in range(1, nfp + 1)]
    m = gpu_ctx.solve(g, probs)
    ranges = np.array([150e3, 230e3, 310e3])
    nfft = 4 * f.size
    S, _ = bb.source(gpu_ctx, "pulse1", f, nfft)
    W, kbufs = bb.synthesize_from_modes(gpu_ctx, m, len(azs), f, ranges, np.ones(len(azs)), S, 320.0, nfft)
    w = W.numpy().reshape(len(azs), ranges.size, nfft)
    for b in (W,) + kbufs:
        b.free()
    for s, az in enumerate(azs):
        modes = [np.zeros((0, 4))]
        for i in range(1, nfp + 1):
            sol = ess.solve(synth_profile, freq=f[i], azimuth=az, Nz=Nz, turnoff_wkb=True)
            kp = sol.k_pert
            ps, pr = sol.psi[sol.i_s], sol.psi[sol.i_r]
            modes.append(np.stack([kp.real, kp.imag, ps, pr], axis=1))
        d = {"f": f, "counts": np.array([q.shape[0] for q in modes]), "modes": modes,
             "rho_s": 1.0, "rho_r": 1.0}
        assert np.array_equal(d["counts"][1:], m.counts[s * nfp:(s + 1) * nfp])
        _, _, wo = obb.waveforms(d, ranges, "pulse1", nfft=nfft, max_cel=320.0)
        err = np.max(np.abs(w[s] - wo)) / np.max(np.abs(wo))
        print("az %g: max waveform error %.2e of peak" % (az, err), flush=True)
        assert err <= 1e-5
    m.free()
