"""Modess / WMod / ModBB command lines: option handling (CPU) and end-to-end runs against the
FP64 oracle (GPU).  The programs are the C++ drop-ins built into ncpaprop-release_amd/bin."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ncpaprop-release_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(prog, args, cwd, timeout=300):
    return subprocess.run([os.path.join(BIN, prog)] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


def test_programs_built():
    for prog in ("Modess", "WMod", "ModBB"):
        assert os.access(os.path.join(BIN, prog), os.X_OK), prog


def test_modess_validation_messages(tmp_path):
    out = run("Modess", ["--freq", "0.1"], tmp_path)
    assert "Parameter validation failed:" in out.stdout
    assert "atmosfile is not present." in out.stdout
    assert "operation_mode: 0 of { singleprop, multiprop } are present" in out.stdout
    out = run("Modess", ["--atmosfile", "x", "--freq", "0.1", "--singleprop", "--multiprop", "--azimuth", "90",
                         "--azimuth_start", "0", "--azimuth_end", "10", "--azimuth_step", "5"], tmp_path)
    assert "operation_mode: 2 of" in out.stdout
    out = run("Modess", ["--atmosfile", "x", "--freq", "-1", "--singleprop", "--azimuth", "400"], tmp_path)
    assert "freq (-1.000000) must be greater than 0.000000." in out.stdout
    assert "azimuth (400.000000) must be less than 360.000000." in out.stdout
    out = run("Modess", ["--atmosfile", "x", "--freq", "0.1", "--singleprop", "--bogus", "1"], tmp_path)
    assert "Unknown option '--bogus'." in out.stdout


def test_modess_help():
    out = run("Modess", ["--help"], ROOT)
    assert out.returncode == 1 and "--atmosfile" in out.stdout


def test_param_file_and_override(tmp_path):
    """modess.param supplies options; the command line overrides them (Modess_main.cpp:58-63)."""
    (tmp_path / "modess.param").write_text("# comment\natmosfile: nothere.dat\nfreq = 0.1\nsingleprop\n"
                                           "azimuth 90\n")
    out = run("Modess", [], tmp_path)
    assert "Parameter validation failed" not in out.stdout
    out = run("Modess", ["--freq", "-2"], tmp_path)
    assert "freq (-2.000000) must be greater than 0.000000." in out.stdout


def test_wmod_options(tmp_path):
    """WMod takes the Modess options except --modal_starter_file and --write_speeds
    (wmod_parameters.cpp:6-214); validation messages as Modess."""
    out = run("WMod", ["--help"], ROOT)
    assert out.returncode == 1 and "tloss_1d.wnm" in out.stdout and "write_speeds" not in out.stdout
    out = run("WMod", ["--freq", "0.1"], tmp_path)
    assert "atmosfile is not present." in out.stdout
    out = run("WMod", ["--atmosfile", "x", "--freq", "0.1", "--singleprop", "--azimuth", "90", "--write_speeds"],
              tmp_path)
    assert "Unknown option '--write_speeds'." in out.stdout
    out = run("WMod", ["--atmosfile", "x", "--freq", "0.1", "--singleprop", "--azimuth", "90",
                       "--modal_starter_file", "s.pe"], tmp_path)
    assert "Unknown option '--modal_starter_file'." in out.stdout


def test_modbb_validation(tmp_path):
    # --method has a default but is still required with --dispersion (modbb_parameters.cpp:64-69)
    out = run("ModBB", ["--dispersion", "--dispersion_file", "d.dat", "--atmosfile", "a", "--azimuth", "90",
                        "--f_min", "0.01", "--f_step", "0.01", "--f_max", "0.1"], tmp_path)
    assert "One of { dispersion } is set, but method is not set." in out.stdout
    out = run("ModBB", ["--propagation"], tmp_path)
    assert "input_dispersion_file is not set." in out.stdout
    out = run("ModBB", ["--dispersion", "--propagation"], tmp_path)
    assert "operation_mode: 2 of { dispersion, propagation }" in out.stdout
    out = run("ModBB", ["--dispersion", "--dispersion_file", "d.dat", "--atmosfile", "a", "--azimuth", "90",
                        "--f_min", "0.01", "--f_step", "0.01", "--f_max", "0.1", "--method", "other"], tmp_path)
    assert 'method: "other" is not in { modess, wmod }.' in out.stdout


# ------------------------------------------------------------------------------------------
# GPU: end-to-end
# ------------------------------------------------------------------------------------------
def _tl_db(p):
    return 20 * np.log10(np.maximum(np.abs(p), 1e-300))


@pytest.mark.gpu
def test_modess_singleprop_outputs(tmp_path, synth_profile):
    from oracle import ess
    f, az, Nz = 0.3, 60.0, 6000
    print("case Modess singleprop f=%g az=%g Nz=%d" % (f, az, Nz), flush=True)
    args = ["--atmosfile", synth_profile, "--freq", str(f), "--singleprop", "--azimuth", str(az),
            "--Nz_grid", str(Nz), "--maxrange_km", "500", "--Nrng_steps", "500", "--write_2d_tloss",
            "--write_lossless", "--write_phase_speeds", "--write_speeds", "--write_modes", "--write_atm_profile",
            "--dispersion_file", "disp.dat", "--filetag", "run1"]
    out = run("Modess", args, tmp_path)
    assert out.returncode == 0, out.stdout + out.stderr
    sol = ess.solve(synth_profile, freq=f, azimuth=az, Nz=Nz, turnoff_wkb=True)
    nm = sol.k.size
    # 1-D TL, format "%f %20.12e %20.12e %20.12e"
    lines = (tmp_path / "run1.tloss_1d.nm").read_text().splitlines()
    assert len(lines) == 500
    first = lines[0].split()[0]
    assert re.fullmatch(r"\d+\.\d{6}", first) and len(lines[0]) == len(first) + 3 * 21, lines[0]
    assert all(re.fullmatch(r" *-?\d\.\d{12}e[+-]\d\d", lines[0][len(first) + 1 + 21 * c:len(first) + 21 * (c + 1)])
               for c in range(3)), lines[0]
    a = np.loadtxt(tmp_path / "run1.tloss_1d.nm")
    r, po, inco = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], 1000.0, 500)
    assert np.allclose(a[:, 0], r / 1000.0, atol=1e-6)
    tl, tlo = _tl_db(a[:, 1] + 1j * a[:, 2]), _tl_db(po)
    mask = tlo > -140
    assert np.max(np.abs(tl - tlo)[mask]) <= 0.1
    b = np.loadtxt(tmp_path / "run1.tloss_1d.lossless.nm")
    _, pl, _ = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], 1000.0, 500, lossless=True)
    tll, tllo = _tl_db(b[:, 1] + 1j * b[:, 2]), _tl_db(pl)
    assert np.max(np.abs(tll - tllo)[tllo > -140]) <= 0.1
    # 2-D TL: rows j = 0, Nz/500, ... ; blank line per range
    rows = ess.tl2d_rows(Nz)
    t2 = np.loadtxt(tmp_path / "run1.tloss_2d.nm")
    assert t2.shape == (500 * rows.size, 4)
    _, F = ess.tl2d(sol.k_pert, sol.psi[rows], sol.psi[sol.i_s], 1000.0, 500)
    ours = (t2[:, 2] + 1j * t2[:, 3]).reshape(500, rows.size)
    tl2, tl2o = _tl_db(ours), _tl_db(F)
    assert np.max(np.abs(tl2 - tl2o)[tl2o > -140]) <= 0.1
    assert np.allclose(t2[:rows.size, 1], rows * sol.grid.dz / 1000.0, atol=1e-6)
    assert (tmp_path / "run1.tloss_2d.lossless.nm").exists()
    # phase speeds and group speeds
    ph = np.loadtxt(tmp_path / "run1.phasespeeds.nm")
    assert ph.shape == (nm, 3)
    assert np.allclose(ph[:, 1], 2 * np.pi * f / sol.k_pert.real, rtol=1e-8)
    sp = np.loadtxt(tmp_path / "run1.speeds.nm")
    vph, vg = ess.group_speeds(sol.psi, sol.ceffz, sol.grid.dz, f, sol.k_pert)
    assert np.allclose(sp[:, 1], vph, atol=2e-3) and np.allclose(sp[:, 2], vg, atol=2e-3)
    # mode files: |psi| sign-free, normalised
    for j in (0, nm // 2, nm - 1):
        md = np.loadtxt(tmp_path / ("run1.mode_%d.nm" % j))
        assert md.shape == (Nz, 2)
        assert np.max(np.abs(np.abs(md[:, 1]) - np.abs(sol.psi[:, j]))) <= 1e-6 * np.max(np.abs(sol.psi[:, j])) + 1e-7
    # dispersion line (not tagged)
    d = (tmp_path / "disp.dat").read_text().split()
    assert float(d[0]) == pytest.approx(f) and int(d[1]) == nm
    q = np.array(d[4:], dtype=float).reshape(nm, 4)
    assert np.allclose(q[:, 0], sol.k_pert.real, rtol=1e-9)
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(q[:, 2] * q[:, 3] - ref)) <= 1e-6 * np.max(np.abs(ref))
    # atmosphere dump: header and 8 columns
    atm = (tmp_path / "run1.atm_profile.nm").read_text().splitlines()
    assert atm[0].startswith("#% 0, Z0, m") and "#% 8, CEFF, m/s" in atm
    vals = np.loadtxt(tmp_path / "run1.atm_profile.nm", comments="#")
    assert vals.shape[1] == 8


@pytest.mark.gpu
def test_modess_multiprop(tmp_path, synth_profile):
    from oracle import ess
    f, Nz = 0.2, 4000
    print("case Modess multiprop f=%g Nz=%d" % (f, Nz), flush=True)
    out = run("Modess", ["--atmosfile", synth_profile, "--freq", str(f), "--multiprop", "--azimuth_start", "0",
                         "--azimuth_end", "90", "--azimuth_step", "45", "--Nz_grid", str(Nz), "--maxrange_km", "300",
                         "--Nrng_steps", "300", "--turnoff_WKB"], tmp_path)
    assert out.returncode == 0, out.stdout + out.stderr
    text = (tmp_path / "Nby2D_tloss_1d.nm").read_text()
    blocks = [b for b in text.split("\n\n") if b.strip()]
    assert len(blocks) == 3
    for az, blk in zip((0.0, 45.0, 90.0), blocks):
        a = np.array([l.split() for l in blk.strip().splitlines()], dtype=float)
        assert a.shape == (300, 5) and np.all(a[:, 1] == az)
        sol = ess.solve(synth_profile, freq=f, azimuth=az, Nz=Nz, turnoff_wkb=True)
        _, po, _ = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], 1000.0, 300)
        tl, tlo = _tl_db(a[:, 2] + 1j * a[:, 3]), _tl_db(po)
        assert np.max(np.abs(tl - tlo)[tlo > -140]) <= 0.1


@pytest.mark.gpu
def test_modbb_dispersion_then_propagation(tmp_path, synth_profile):
    from oracle import broadband as obb
    from oracle import ess
    Nz = 4000
    print("case ModBB dispersion+propagation Nz=%d" % Nz, flush=True)
    out = run("ModBB", ["--dispersion", "--dispersion_file", "disp.dat", "--atmosfile", synth_profile,
                        "--azimuth", "90", "--f_min", "0.01", "--f_step", "0.01", "--f_max", "0.12",
                        "--Nz_grid", str(Nz), "--method", "modess"], tmp_path)
    assert out.returncode == 0 and "validation failed" not in out.stdout, out.stdout + out.stderr
    lines = (tmp_path / "disp.dat").read_text().splitlines()
    assert len(lines) == 12
    assert re.match(r"\d\.\d{7}e[+-]\d\d   \d+    \d\.\d{7}e[+-]\d\d   \d\.\d{7}e[+-]\d\d   ", lines[0]), lines[0]
    for i in (0, 5, 11):
        v = lines[i].split()
        f = float(v[0])
        sol = ess.solve(synth_profile, freq=f, azimuth=90.0, Nz=Nz, turnoff_wkb=True)
        q = np.array(v[4:], dtype=float).reshape(int(v[1]), 4)
        assert q.shape[0] == sol.k.size
        assert np.allclose(q[:, 0], sol.k_pert.real, rtol=1e-9)
        ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
        assert np.max(np.abs(q[:, 2] * q[:, 3] - ref)) <= 1e-6 * np.max(np.abs(ref))
    # tloss_1d.nm of the last frequency is a side product of the dispersion run
    assert (tmp_path / "tloss_1d.nm").exists()
    out = run("ModBB", ["--propagation", "--input_dispersion_file", "disp.dat", "--output_waveform_file",
                        "wave.dat", "--receiver", "multiple", "--start_range_km", "100", "--end_range_km", "200",
                        "--range_step_km", "50", "--source", "pulse1", "--max_celerity", "320"], tmp_path)
    assert out.returncode == 0, out.stdout + out.stderr
    disp = obb.read_dispersion(str(tmp_path / "disp.dat"))
    nfft, to, wo = obb.waveforms(disp, [100e3, 150e3, 200e3], "pulse1", max_cel=320.0)
    text = (tmp_path / "wave.dat").read_text()
    blocks = [b for b in text.split("\n\n") if b.strip()]
    assert len(blocks) == 3
    for k, blk in enumerate(blocks):
        a = np.array([l.split() for l in blk.strip().splitlines()], dtype=float)
        assert a.shape == (nfft, 3)
        assert np.allclose(a[:, 1], to[k], atol=2e-6)
        assert np.max(np.abs(a[:, 2] - wo[k])) <= 1e-5 * np.max(np.abs(wo[k])) + 1e-300
    src = np.loadtxt(tmp_path / "source_spectrum_input.dat")
    S, _ = obb.source_spectrum("pulse1", disp["f"], nfft, disp["f"][2] - disp["f"][1])
    assert np.allclose(src[:, 1] + 1j * src[:, 2], S, rtol=1e-5, atol=1e-6 * np.abs(S).max())


@pytest.mark.gpu
def test_wmod_singleprop_outputs(tmp_path, synth_profile):
    """WMod end to end: .wnm files (WModeSolver.h:4-34), 1-D / 2-D TL, phase speeds, mode files
    and the dispersion line against the wide-angle oracle; matrix rows every H/Nz."""
    from oracle import ess, wmod
    f, az, Nz = 0.3, 60.0, 6000
    print("case WMod singleprop f=%g az=%g Nz=%d" % (f, az, Nz), flush=True)
    args = ["--atmosfile", synth_profile, "--freq", str(f), "--singleprop", "--azimuth", str(az),
            "--Nz_grid", str(Nz), "--maxrange_km", "500", "--Nrng_steps", "500", "--write_2d_tloss",
            "--write_lossless", "--write_phase_speeds", "--write_modes", "--write_atm_profile",
            "--dispersion_file", "disp.dat"]
    out = run("WMod", args, tmp_path)
    assert out.returncode == 0, out.stdout + out.stderr
    sol = wmod.solve(synth_profile, freq=f, azimuth=az, Nz=Nz, turnoff_wkb=True)
    nm = sol.k.size
    a = np.loadtxt(tmp_path / "tloss_1d.wnm")
    r, po, inco = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], 1000.0, 500)
    assert a.shape == (500, 4) and np.allclose(a[:, 0], r / 1000.0, atol=1e-6)
    tl, tlo = _tl_db(a[:, 1] + 1j * a[:, 2]), _tl_db(po)
    assert np.max(np.abs(tl - tlo)[tlo > -140]) <= 0.1
    assert np.max(np.abs(a[:, 3] / inco - 1)) <= 1e-6
    assert (tmp_path / "tloss_1d.lossless.wnm").exists() and not (tmp_path / "tloss_1d.nm").exists()
    rows = ess.tl2d_rows(Nz)
    t2 = np.loadtxt(tmp_path / "tloss_2d.wnm")
    assert t2.shape == (500 * rows.size, 4)
    _, F = ess.tl2d(sol.k_pert, sol.psi[rows], sol.psi[sol.i_s], 1000.0, 500)
    ours = (t2[:, 2] + 1j * t2[:, 3]).reshape(500, rows.size)
    tl2, tl2o = _tl_db(ours), _tl_db(F)
    assert np.max(np.abs(tl2 - tl2o)[tl2o > -140]) <= 0.1
    assert np.allclose(t2[:rows.size, 1], rows * sol.grid.dz_m / 1000.0, atol=1e-6)
    ph = np.loadtxt(tmp_path / "phasespeeds.wnm")
    assert ph.shape == (nm, 3) and np.allclose(ph[:, 1], 2 * np.pi * f / sol.k_pert.real, rtol=1e-8)
    for j in (0, nm // 2, nm - 1):
        md = np.loadtxt(tmp_path / ("mode_%d.wnm" % j))
        assert md.shape == (Nz, 2)
        assert np.allclose(md[1, 0], sol.grid.dz_m / 1000.0, rtol=1e-6)
        assert np.max(np.abs(np.abs(md[:, 1]) - np.abs(sol.psi[:, j]))) <= 1e-6 * np.max(np.abs(sol.psi[:, j])) + 1e-7
    assert not (tmp_path / "speeds.nm").exists()
    d = (tmp_path / "disp.dat").read_text().split()
    assert float(d[0]) == pytest.approx(f) and int(d[1]) == nm
    q = np.array(d[4:], dtype=float).reshape(nm, 4)
    assert np.allclose(q[:, 0], sol.k_pert.real, rtol=1e-9)
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(q[:, 2] * q[:, 3] - ref)) <= 1e-6 * np.max(np.abs(ref))
    assert (tmp_path / "atm_profile.wnm").exists()


@pytest.mark.gpu
def test_modbb_wmod_dispersion(tmp_path, synth_profile):
    """ModBB --dispersion --method wmod writes the wide-angle dispersion lines (ModBB_main.cpp:86-90)."""
    from oracle import wmod
    Nz = 4000
    print("case ModBB wmod dispersion Nz=%d" % Nz, flush=True)
    out = run("ModBB", ["--dispersion", "--dispersion_file", "disp.dat", "--atmosfile", synth_profile,
                        "--azimuth", "90", "--f_min", "0.02", "--f_step", "0.02", "--f_max", "0.1",
                        "--Nz_grid", str(Nz), "--method", "wmod"], tmp_path)
    assert out.returncode == 0 and "validation failed" not in out.stdout, out.stdout + out.stderr
    lines = (tmp_path / "disp.dat").read_text().splitlines()
    assert len(lines) == 5
    for i in (0, 4):
        v = lines[i].split()
        f = float(v[0])
        sol = wmod.solve(synth_profile, freq=f, azimuth=90.0, Nz=Nz, turnoff_wkb=True)
        q = np.array(v[4:], dtype=float).reshape(int(v[1]), 4)
        assert q.shape[0] == sol.k.size
        assert np.allclose(q[:, 0], sol.k_pert.real, rtol=1e-9)
        ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
        assert np.max(np.abs(q[:, 2] * q[:, 3] - ref)) <= 1e-6 * np.max(np.abs(ref))
