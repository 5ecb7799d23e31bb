"""The oracle's C restatement (oracle/ess_port.c, the timed CPU baseline) against the Python
oracle on the same inputs: modes, perturbed wavenumbers, psi(zs) psi(zr), 1-D TL, with WKB on
and off, the Lamb-wave ground term and the wavenumber filter."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _port():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build as obuild
    return obuild.port_path()


CASES = [  # (freq, azimuth, Nz, zs_m, zr_m, wkb, lamb, filter)
    (0.1, 90.0, 20000, 0.0, 0.0, True, False, None),
    (0.3, 60.0, 6000, 1000.0, 0.0, False, False, None),
    (0.5, 270.0, 4000, 0.0, 500.0, False, True, None),
    (0.4, 200.0, 8000, 0.0, 0.0, True, False, (305.0, 350.0)),
]


@pytest.mark.parametrize("f,az,Nz,zs,zr,wkb,lamb,filt", CASES)
def test_port_matches_python_oracle(synth_profile, f, az, Nz, zs, zr, wkb, lamb, filt):
    from oracle import ess
    nr, dr = 1000, 1000.0
    cmd = [_port(), "solve", "--profile", synth_profile, "--freq", str(f), "--azimuth", str(az), "--Nz", str(Nz),
           "--zs_m", str(zs), "--zr_m", str(zr), "--nr", str(nr), "--dr", str(dr)]
    cmd += (["--wkb"] if wkb else []) + (["--lamb"] if lamb else [])
    if filt:
        cmd += ["--filter", str(filt[0]), str(filt[1])]
    out = json.loads(subprocess.run(cmd, capture_output=True, text=True, check=True, timeout=120).stdout)
    sol = ess.solve(synth_profile, freq=f, azimuth=az, Nz=Nz, zs_km=zs / 1e3, zr_km=zr / 1e3, turnoff_wkb=not wkb,
                    lamb_wave_bc=lamb, wvnum_filter=filt)
    assert out["used_wkb"] == int(sol.used_wkb)
    assert out["k_min"] == pytest.approx(sol.k_min, rel=1e-14)
    assert out["k_max"] == pytest.approx(sol.k_max, rel=1e-14)
    assert out["nm"] == sol.k.size
    k = np.array(out["k"])
    assert np.max(np.abs(k / sol.k - 1)) <= 1e-10
    assert np.allclose(np.array(out["kp_im"]), sol.k_pert.imag, rtol=1e-8, atol=0)
    prod = np.array(out["psi_s"]) * np.array(out["psi_r"])
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(prod - ref)) <= 1e-8 * np.max(np.abs(ref))
    r, po, inco = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], dr, nr)
    tl = ess.tl_db(np.array(out["tl_re"]) + 1j * np.array(out["tl_im"]))
    tlo = ess.tl_db(po)
    mask = tlo > -140
    assert mask.any() and np.max(np.abs(tl - tlo)[mask]) <= 1e-4
    assert np.allclose(np.array(out["tl_inc"]), inco, rtol=1e-8, atol=0)


def test_port_bench_mode(synth_profile):
    """bench mode: jobs on threads, per-job stage times and the field checksum (all jobs equal
    problems give equal mode counts and checksums)."""
    job = "%s:90:0.5" % synth_profile
    cmd = [_port(), "bench", "--threads", "2", "--Nz", "4000", "--nr", "100", "--dr", "1000", "--rows", "50",
           "--stride", "40", "--field_ranges", "3", "--job", job, "--job", job, "--job", job]
    out = json.loads(subprocess.run(cmd, capture_output=True, text=True, check=True, timeout=120).stdout)
    assert out["threads"] == 2 and len(out["jobs"]) == 3
    nm = {j["nm"] for j in out["jobs"]}
    chk = {j["check"] for j in out["jobs"]}
    assert len(nm) == 1 and nm.pop() > 0 and len(chk) == 1
    assert all(len(j["times"]) == 4 and min(j["times"]) >= 0 for j in out["jobs"])


def test_port_rejects_bad_input(tmp_path):
    bad = tmp_path / "bad.dat"
    bad.write_text("#% 1, Z, km\n#% 2, U, m/s\n0 1 2\n")
    r = subprocess.run([_port(), "solve", "--profile", str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "wrong number of columns" in r.stderr
    r = subprocess.run([_port(), "solve", "--profile", str(tmp_path / "missing.dat")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "cannot open" in r.stderr
