"""GPU parity of the WKB ground-to-ground cut-off (ESSModeSolver.cpp:561-621) against the oracle.

The Modess run on the synthetic profile family at 0.1 Hz, azimuth 90, source and receiver at
the ground, Nz = 20000 over 0-150 km, 1-D TL at 0.1..1000 km every 0.1 km: WKB on.  Further
cases take the two branches that bypass the stepping: c_eff smallest at the ground (dkk = 0)
and a cut-off below k_min (WKB turned off again).
"""
import math
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ncpaprop-release_amd", "bin")
NR, DR = 10000, 100.0


def _profiles(tmp):
    """name -> path: the family's seeds 0 and 3, a temperature inversion from the ground
    (c_eff smallest at the ground) and a fast ground jet (the cut-off falls below k_min)."""
    from ncpamodes import synth
    z = np.arange(1501) * 0.1
    T0, _ = synth.temperature_pressure(z)
    zero = np.zeros_like(z)
    return {
        "P0": synth.write_profile(os.path.join(tmp, "P0.dat"), seed=0),
        "P3": synth.write_profile(os.path.join(tmp, "P3.dat"), seed=3),
        "inversion": synth.write_columns(os.path.join(tmp, "inv.dat"), z, zero, zero, 250.0 + 1.5 * z),
        "ground_jet": synth.write_columns(os.path.join(tmp, "jet.dat"), z, 100.0 * np.exp(-0.5 * (z / 8.0) ** 2),
                                          zero, T0),
    }


CASES = [("P0", 0.1, 90.0), ("P3", 0.1, 90.0), ("P0", 0.1, 270.0), ("P3", 0.5, 90.0), ("P0", 1.0, 0.0),
         ("inversion", 0.1, 90.0), ("inversion", 0.5, 30.0), ("ground_jet", 0.5, 90.0)]


def _wkb_margin(sol, freq):
    """Distance of the oracle's deciding integrals from 10 (how far from the knife edge)."""
    from oracle import ess
    kk, steps = ess.wkb_cutoff(sol.ceffz, 2 * math.pi * freq, sol.grid.dz)
    return min((abs(s[2] - 10.0) for s in steps), default=float("inf")), len(steps)


def test_wkb_batch_vs_oracle(gpu_ctx, tmp_path):
    """One batch of every case: k_min, k_max (the same WKB step), mode counts, k^2 to 1e-9,
    Im k, psi(0)^2 and |TL| over 10000 ranges to 0.1 dB where TL > -140 dB."""
    import ncpamodes
    from oracle import ess
    paths = _profiles(str(tmp_path))
    pids = {name: gpu_ctx.read_profile(p)[0] for name, p in paths.items()}
    g = gpu_ctx.grid(pids["P0"])
    probs = [ncpamodes.problem(pids[n], f, az, 0.0, 0.0, wkb=True) for n, f, az in CASES]
    print("case WKB batch of %d problems, Nz %d" % (len(probs), g.Nz), flush=True)
    m = gpu_ctx.solve(g, probs)
    kmin, kmax = m.interval()
    p_all, inc_all = m.tl1d(DR, NR)
    normT = 4.0 / g.dz ** 2
    for q, (name, f, az) in enumerate(CASES):
        sol = ess.solve(paths[name], freq=f, azimuth=az, Nz=g.Nz, turnoff_wkb=False)
        margin, nsteps = _wkb_margin(sol, f)
        k, psr, ps, pr = m.problem_modes(q)
        print("%s f=%g az=%g: used_wkb %s, %d WKB steps, margin %.3e, k_max %.15e / %.15e, modes %d / %d"
              % (name, f, az, sol.used_wkb, nsteps, margin, kmax[q], sol.k_max, k.size, sol.k.size), flush=True)
        assert sol.used_wkb
        assert abs(kmin[q] - sol.k_min) <= 1e-13 * sol.k_min
        assert abs(kmax[q] - sol.k_max) <= 1e-13 * sol.k_max
        assert k.size == sol.k.size
        k2, k2o = (k ** 2).real, sol.k ** 2
        assert np.all(np.abs(k2 - k2o) <= 1e-9 * k2o + 64 * 2.2e-16 * normT)
        assert np.allclose(k.imag, sol.k_pert.imag, rtol=1e-4, atol=1e-14)
        # psi(0)^2 to 1e-6 of the largest, or of the mean psi^2 = 1/H when every mode is
        # evanescent at the ground (the fast ground jet: psi(0)^2 ~ 1e-20, at the FP64 noise
        # of the adjugate products, ~1e-10 of 1/H)
        ref = sol.psi[0] ** 2
        scale = max(np.max(ref), 1.0 / (g.Nz * g.dz))
        assert np.max(np.abs(psr - ref)) <= 1e-6 * scale
        r, po, inco = ess.tl1d(sol.k_pert, sol.psi[0], sol.psi[0], DR, NR)
        tl, tlo = ess.tl_db(p_all[q]), ess.tl_db(po)
        mask = tlo > -140.0
        if not mask.any():
            # every mode of the fallback case is slower than c_eff at the ground (its interval
            # lies above k_gnd), so the ground-level field is evanescent: ~ -300 dB
            print("   TL below -140 dB at every range (max %.1f dB; ours %.1f dB)" % (tlo.max(), tl.max()), flush=True)
            assert np.all(tl < -140.0)
            continue
        err = float(np.max(np.abs(tl - tlo)[mask]))
        print("   max |dTL| %.2e dB over %d ranges" % (err, mask.sum()), flush=True)
        assert err <= 0.1
        imask = ess.tl_db(inco) > -140.0
        assert np.max(np.abs(inc_all[q][imask] / inco[imask] - 1)) <= 1e-6
    # the branches: dkk = 0 gives k_max = omega / c_eff(ground); the fallback omega / min c_eff
    from oracle import ess as _e
    for q, (name, f, az) in enumerate(CASES):
        if name == "inversion":
            sol = _e.solve(paths[name], freq=f, azimuth=az, Nz=g.Nz)
            assert abs(kmax[q] - 2 * math.pi * f / sol.ceffz[0]) <= 1e-14 * kmax[q]
        if name == "ground_jet":
            sol = _e.solve(paths[name], freq=f, azimuth=az, Nz=g.Nz)
            kk, _ = _e.wkb_cutoff(sol.ceffz, 2 * math.pi * f, g.dz)
            assert math.sqrt(kk) < sol.k_min   # the cut-off fell below k_min ...
            assert abs(kmax[q] - 2 * math.pi * f / sol.ceffz.min()) <= 1e-14 * kmax[q]   # ... so it was dropped
    m.free()


def test_wkb_off_when_source_above_ground(gpu_ctx, tmp_path):
    """zs = 1 m: the same flags leave the interval at omega / min c_eff."""
    import ncpamodes
    from oracle import ess
    paths = _profiles(str(tmp_path))
    pid = gpu_ctx.read_profile(paths["P0"])[0]
    g = gpu_ctx.grid(pid)
    print("case WKB flag with zs = 1 m", flush=True)
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, 0.1, 90.0, 1.0, 0.0, wkb=True)])
    sol = ess.solve(paths["P0"], freq=0.1, azimuth=90.0, zs_km=1e-3, turnoff_wkb=False)
    assert not sol.used_wkb
    _, kmax = m.interval()
    assert abs(kmax[0] - sol.k_max) <= 1e-14 * sol.k_max
    assert m.counts[0] == sol.k.size
    m.free()


def test_modess_ground_to_ground_cli(tmp_path):
    """The drop-in command line at 0.1 Hz, azimuth 90, zs = zr = 0, Nz 20000, 1-D output only:
    WKB stays on (ESSModeSolver.cpp:78-81), tloss_1d.nm matches the oracle with WKB."""
    from oracle import ess
    from ncpamodes import synth
    path = synth.write_profile(str(tmp_path / "P0.dat"), seed=0)
    args = ["--atmosfile", path, "--freq", "0.1", "--singleprop", "--azimuth", "90", "--Nz_grid", "20000",
            "--maxheight_km", "150", "--sourceheight_km", "0", "--receiverheight_km", "0",
            "--maxrange_km", "1000", "--Nrng_steps", "10000", "--write_lossless"]
    print("case Modess ground-to-ground 0.1 Hz", flush=True)
    out = subprocess.run([os.path.join(BIN, "Modess")] + args, cwd=tmp_path, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    sol = ess.solve(path, freq=0.1, azimuth=90.0, Nz=20000, turnoff_wkb=False)
    assert sol.used_wkb
    a = np.loadtxt(tmp_path / "tloss_1d.nm")
    assert a.shape == (NR, 4)
    r, po, inco = ess.tl1d(sol.k_pert, sol.psi[0], sol.psi[0], DR, NR)
    assert np.allclose(a[:, 0], r / 1000.0, atol=1e-6)
    tl, tlo = ess.tl_db(a[:, 1] + 1j * a[:, 2]), ess.tl_db(po)
    mask = tlo > -140
    assert np.max(np.abs(tl - tlo)[mask]) <= 0.1
    assert np.allclose(a[:, 3], inco, rtol=1e-6, atol=0)
    b = np.loadtxt(tmp_path / "tloss_1d.lossless.nm")
    _, pl, _ = ess.tl1d(sol.k_pert, sol.psi[0], sol.psi[0], DR, NR, lossless=True)
    tll, tllo = ess.tl_db(b[:, 1] + 1j * b[:, 2]), ess.tl_db(pl)
    assert np.max(np.abs(tll - tllo)[tllo > -140]) <= 0.1
