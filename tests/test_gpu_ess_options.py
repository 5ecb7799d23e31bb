"""GPU parity of the ESS options that change the problem: the Lamb-wave ground condition
(nonzero admittance on row 0, ESSModeSolver.cpp:344-357, 558-559), the wavenumber filter
(ESSModeSolver.cpp:366-369), a user attenuation file (Atmosphere1D.cpp:307-424 via
ESSModeSolver.cpp:224-228) and the modal starter file (ESSModeSolver.cpp:670-723), through the
C ABI and through the Modess command line, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ncpaprop-release_amd", "bin")


def _compare(m, q, sol, g, dr=100.0, nr=5000, p_all=None):
    from oracle import ess
    k, psr, ps, pr = m.problem_modes(q)
    assert k.size == sol.k.size
    k2, k2o = (k ** 2).real, sol.k ** 2
    normT = 4.0 / g.dz ** 2
    assert np.all(np.abs(k2 - k2o) <= 1e-9 * k2o + 64 * 2.2e-16 * normT)
    assert np.allclose(k.imag, sol.k_pert.imag, rtol=1e-4, atol=1e-14)
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(psr - ref)) <= 1e-6 * np.max(np.abs(ref))
    if p_all is not None:
        r, po, _ = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], dr, nr)
        tl, tlo = ess.tl_db(p_all[q]), ess.tl_db(po)
        mask = tlo > -140.0
        assert mask.any()
        assert np.max(np.abs(tl - tlo)[mask]) <= 0.1


def test_lamb_wave_bc_vs_oracle(gpu_ctx, synth_profile):
    """Admittance -rho'/(2 rho) at the ground, then the ground row 1/((dz adm + 1) dz^2)."""
    import ncpamodes
    from oracle import ess
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    adm = gpu_ctx.admittance(pid, g.z_min)
    sol = ess.solve(synth_profile, freq=0.5, azimuth=90.0, turnoff_wkb=True, lamb_wave_bc=True)
    adm_o = ess.admittance(sol.grid, True)
    print("case Lamb-wave BC: admittance %.12e (oracle %.12e)" % (adm, adm_o), flush=True)
    assert adm == pytest.approx(adm_o, rel=1e-9) and adm > 0
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, 0.5, 90.0, admittance=adm), ncpamodes.problem(pid, 0.5, 90.0)])
    p_all, _ = m.tl1d(100.0, 5000)
    _compare(m, 0, sol, g, p_all=p_all)
    # the ground term is live: the admittance moves the wavenumbers of modes with ground amplitude
    k_adm, _, _, _ = m.problem_modes(0)
    k_rig, _, _, _ = m.problem_modes(1)
    assert k_adm.size == k_rig.size
    assert np.max(np.abs(k_adm.real / k_rig.real - 1)) > 1e-7
    m.free()


def test_wvnum_filter_vs_oracle(gpu_ctx, synth_profile):
    """k_min = omega/c_max, k_max = omega/c_min replace the modal-trace interval."""
    import ncpamodes
    from oracle import ess
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    f, cmin, cmax = 0.5, 300.0, 345.0
    print("case wavenumber filter c in [%g, %g] at %g Hz" % (cmin, cmax, f), flush=True)
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, f, 270.0, wvnum_filter=(cmin, cmax))])
    sol = ess.solve(synth_profile, freq=f, azimuth=270.0, turnoff_wkb=True, wvnum_filter=(cmin, cmax))
    kmin, kmax = m.interval()
    omega = 2 * np.pi * f
    assert kmin[0] == pytest.approx(omega / cmax, rel=1e-15) and kmax[0] == pytest.approx(omega / cmin, rel=1e-15)
    p_all, _ = m.tl1d(100.0, 5000)
    _compare(m, 0, sol, g, p_all=p_all)
    assert np.all(2 * np.pi * f / sol.k <= cmax + 1e-9) and np.all(2 * np.pi * f / sol.k >= cmin - 1e-9)
    m.free()


def _attn_file(path):
    """A table from 10 to 120 km every 5 km (short of the profile at both ends)."""
    zt = np.arange(10.0, 121.0, 5.0)
    at = 2e-7 * np.exp(zt / 25.0)
    with open(path, "w") as fh:
        fh.write("# altitude [km]  attenuation [Np/m]\n")
        for z, a in zip(zt, at):
            fh.write("%.3f %.10e\n" % (z, a))
    return path


def test_modess_attn_file_lamb_starter_cli(tmp_path, synth_profile):
    """Modess with --use_attn_file, --Lamb_wave_BC and --modal_starter_file, source at 1 km:
    tloss_1d.nm and the starter file against the oracle with the same attenuation table,
    admittance and source row."""
    from oracle import ess
    from oracle.atmos import attenuation_file_nodes
    f, az, Nz = 0.3, 60.0, 6000
    attn = _attn_file(str(tmp_path / "attn.dat"))
    args = ["--atmosfile", synth_profile, "--freq", str(f), "--singleprop", "--azimuth", str(az),
            "--Nz_grid", str(Nz), "--sourceheight_km", "1", "--receiverheight_km", "0", "--maxrange_km", "500",
            "--Nrng_steps", "500", "--use_attn_file", attn, "--Lamb_wave_BC", "--modal_starter_file", "starter.pe"]
    print("case Modess --use_attn_file --Lamb_wave_BC --modal_starter_file", flush=True)
    out = subprocess.run([os.path.join(BIN, "Modess")] + args, cwd=tmp_path, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    nodes = ess.AtmosphereNodes(ess.read_profile(synth_profile))
    a_nodes = attenuation_file_nodes(attn, nodes.z / 1000.0)
    sol = ess.solve(synth_profile, freq=f, azimuth=az, Nz=Nz, zs_km=1.0, zr_km=0.0, lamb_wave_bc=True,
                    attn_nodes=a_nodes)
    assert not sol.used_wkb
    a = np.loadtxt(tmp_path / "tloss_1d.nm")
    r, po, inco = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], 1000.0, 500)
    tl, tlo = ess.tl_db(a[:, 1] + 1j * a[:, 2]), ess.tl_db(po)
    mask = tlo > -140
    assert np.max(np.abs(tl - tlo)[mask]) <= 0.1
    assert np.allclose(a[:, 3], inco, rtol=1e-5, atol=0)
    # the table is live: Sutherland-Bass at this frequency gives a different incoherent TL
    sol_sb = ess.solve(synth_profile, freq=f, azimuth=az, Nz=Nz, zs_km=1.0, zr_km=0.0, lamb_wave_bc=True)
    _, _, inc_sb = ess.tl1d(sol_sb.k_pert, sol_sb.psi[sol_sb.i_s], sol_sb.psi[sol_sb.i_r], 1000.0, 500)
    assert np.max(np.abs(inc_sb / inco - 1)) > 1e-2
    # modal starter: rows every int((340/f)/10/dz), "%10.3f   %16.12e   %16.12e", closing blank line
    text = (tmp_path / "starter.pe").read_text()
    z_o, st_o = ess.modal_starter(sol.psi, sol.i_s, sol.k_pert, sol.grid.dz, f)
    lines = text.split("\n")
    assert lines[-1] == "" and lines[-2] == ""
    rows = [ln for ln in lines if ln]
    assert len(rows) == z_o.size
    ref = ess.format_modal_starter(z_o, st_o).split("\n")
    for got_ln, ref_ln in zip(rows[:3], ref[:3]):
        assert len(got_ln) == len(ref_ln) and got_ln[:10] == ref_ln[:10], (got_ln, ref_ln)
    vals = np.array([[float(x) for x in ln.split()] for ln in rows])
    assert np.allclose(vals[:, 0], z_o / 1000.0, atol=5e-4)
    # each eigenvector row agrees with LAPACK's to ~1e-7 of max |psi|; summed over the modes
    # that is ~1e-6 of the starter's largest value (6.8e-7 measured)
    assert np.max(np.abs(vals[:, 1] - st_o)) <= 2e-6 * np.max(np.abs(st_o))
    assert np.all(vals[:, 2] == 0.0)


def test_modess_wvnum_filter_cli(tmp_path, synth_profile):
    from oracle import ess
    f, az, Nz = 0.4, 270.0, 8000
    args = ["--atmosfile", synth_profile, "--freq", str(f), "--singleprop", "--azimuth", str(az),
            "--Nz_grid", str(Nz), "--sourceheight_km", "0", "--receiverheight_km", "0", "--maxrange_km", "800",
            "--Nrng_steps", "800", "--wvnum_filter", "--c_min", "305", "--c_max", "350"]
    print("case Modess --wvnum_filter", flush=True)
    out = subprocess.run([os.path.join(BIN, "Modess")] + args, cwd=tmp_path, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    # the filter replaces the interval after the WKB cut-off (zs = zr = 0 keeps WKB on)
    sol = ess.solve(synth_profile, freq=f, azimuth=az, Nz=Nz, turnoff_wkb=False, wvnum_filter=(305.0, 350.0))
    a = np.loadtxt(tmp_path / "tloss_1d.nm")
    r, po, _ = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], 1000.0, 800)
    tl, tlo = ess.tl_db(a[:, 1] + 1j * a[:, 2]), ess.tl_db(po)
    mask = tlo > -140
    assert mask.any() and np.max(np.abs(tl - tlo)[mask]) <= 0.1
