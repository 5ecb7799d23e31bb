"""Pins of the oracle's optional ESS outputs and inputs: group speeds (ESSModeSolver.cpp:726-752),
the modal starter (ESSModeSolver.cpp:670-723) and the attenuation-file path
(Atmosphere1D.cpp:307-424), each against what the mathematics fixes."""
import math

import numpy as np
import pytest
import scipy.linalg

from oracle import ess
from oracle.atmos import PI, Steffen, attenuation_file_nodes, read_attenuation_file


def _modes_by_index(diag, dz, i0, i1):
    a, e = ess.matrix(diag, dz)
    w, v = scipy.linalg.eigh_tridiagonal(a, e, select="i", select_range=(i0, i1))
    return w, v / math.sqrt(dz)


def test_group_speed_constant_ceff_closed_form():
    """Constant c_eff: sum psi^2 dz = 1 makes v_g = 1/(v_ph dz sum psi^2 / c^2) = c^2 / v_ph."""
    c, freq, Nz, dz = 340.0, 0.5, 800, 50.0
    ceff = np.full(Nz, c)
    omega = 2 * PI * freq
    diag = (omega / ceff) ** 2
    diag[0] += 1.0 / dz ** 2
    k2, psi = ess.eigen(diag, dz, 0.5 * omega / c, omega / c)
    k = np.sqrt(k2)
    vph, vg = ess.group_speeds(psi, ceff, dz, freq, k.astype(np.complex128))
    assert k.size > 10
    assert np.allclose(vg, c * c / vph, rtol=1e-12)
    assert np.all(vg < c) and np.all(vph > c)


def test_group_speed_is_domega_dk(synth_profile):
    """v_g = d omega / dk of the discrete dispersion relation: the matrix depends on omega only
    through diag = (omega/c_eff)^2, so (Hellmann-Feynman) dk^2/domega = 2 omega sum psi^2/c^2 dz
    and 1/v_g = dk/domega = sum psi^2 dz/c^2 / v_ph.  Checked against a central difference of
    k_n(omega) at fixed eigenvalue index n."""
    sol = ess.solve(synth_profile, freq=0.3, azimuth=90.0, Nz=4000, turnoff_wkb=True)
    g = sol.grid
    # the unperturbed k: the reference's v_phase = omega / Re k_pert differs from omega / k for
    # the modes with the most absorption (those reaching the thermosphere), by up to 1e-4 here
    vph, vg = ess.group_speeds(sol.psi, sol.ceffz, g.dz, 0.3, sol.k.astype(np.complex128))
    a, e = ess.matrix(sol.diag, g.dz)
    n_below = int(np.sum(scipy.linalg.eigh_tridiagonal(a, e, eigvals_only=True, select="v",
                                                        select_range=(-1e300, sol.k[0] ** 2 * (1 - 1e-12)))
                         .size))
    i0, i1 = n_below, n_below + sol.k.size - 1
    df = 1e-5
    ks = []
    for f in (0.3 - df, 0.3 + df):
        diag = (2 * PI * f / sol.ceffz) ** 2
        diag[0] += sol.diag[0] - (2 * PI * 0.3 / sol.ceffz[0]) ** 2   # the same ground term
        w, _ = _modes_by_index(diag, g.dz, i0, i1)
        ks.append(np.sqrt(w))
    vg_fd = (2 * PI * 2 * df) / (ks[1] - ks[0])
    assert np.allclose(vg, vg_fd, rtol=1e-6)
    assert np.all(vg < vph)


def test_modal_starter_is_field_at_source_range():
    """The starter is the lossless getTLoss2D depth profile with the range factor taken out:
    starter(z) = pi sqrt(k0) lim_{r->0} sqrt(r) p(r, z) / (1000 expov8pi) (e^{i k r} -> 1), on the
    rows 0, z_cnd, .. with z_cnd = int((340/f)/10/dz)."""
    rng = np.random.default_rng(3)
    Nz, dz, freq = 3000, 5.0, 0.5
    nm = 40
    psi = rng.standard_normal((Nz, nm)) * 1e-3
    k_pert = (2 * PI * freq / rng.uniform(300, 360, nm)) + 1j * rng.uniform(0, 1e-6, nm)
    i_s = 17
    z, st = ess.modal_starter(psi, i_s, k_pert, dz, freq)
    z_cnd = int((340.0 / freq) / 10 / dz)
    assert z_cnd == 13 and np.array_equal(z, np.arange(0, Nz, z_cnd) * dz)
    r = np.array([1e-12])
    rows = np.arange(0, Nz, z_cnd)
    p = ess.tl2d_at(k_pert, psi[rows], psi[i_s], r, lossless=True)[0]
    k0 = 2 * PI * freq / 340.0
    expect = (PI * math.sqrt(k0) * math.sqrt(r[0]) * p / (1000.0 * ess.EXPOV8PI))
    assert np.allclose(st, expect.real, rtol=1e-10, atol=1e-18)
    assert np.max(np.abs(expect.imag)) <= 1e-10 * np.max(np.abs(expect.real))
    text = ess.format_modal_starter(z[:2], st[:2])
    assert text.splitlines()[1].startswith("     0.065   ") and text.endswith("\n\n")


def test_attenuation_file_nodes(tmp_path):
    """Outside the table its end values are held (the table is extended by an end node carrying
    them), inside it the Steffen spline of the extended table; comments, blank lines and commas
    are accepted; a line with three fields is an error."""
    p = tmp_path / "attn.dat"
    zt = np.arange(10.0, 121.0, 5.0)
    at = 1e-6 + 2e-8 * zt
    p.write_text("# altitude km, attenuation Np/m\n\n" +
                 "".join("%.3f, %.10e\n" % (z, a) for z, a in zip(zt, at)))
    z, a = read_attenuation_file(str(p))
    assert np.array_equal(z, zt) and np.allclose(a, at, rtol=1e-10, atol=0)
    nodes = np.arange(0.0, 150.05, 0.1)
    an = attenuation_file_nodes(str(p), nodes)
    # between interior table nodes Steffen reproduces linear data exactly (both end slopes are
    # the secant); in the first and last table intervals the extension's flat segment sets the
    # end node's slope to 0, so the spline runs below (above) the chord there
    inner = (nodes >= 15.0) & (nodes <= 115.0)
    assert np.allclose(an[inner], 1e-6 + 2e-8 * nodes[inner], rtol=1e-12)
    first = (nodes > 10.0) & (nodes < 15.0)
    assert np.all(an[first] < 1e-6 + 2e-8 * nodes[first]) and np.all(an[first] > a[0])
    assert np.all(an[nodes <= 10.0] == a[0]) and np.all(an[nodes >= 120.0] == a[-1])
    # the extension changes the Steffen slope at the first table node: 0 (flat pad segment)
    # instead of the one-sided end slope, so a curved table differs from the unextended spline
    at2 = 1e-6 * np.exp(zt / 30.0)
    p.write_text("".join("%.3f %.10e\n" % (z, a) for z, a in zip(zt, at2)))
    an2 = attenuation_file_nodes(str(p), nodes)
    plain = Steffen(zt, np.array([float("%.10e" % v) for v in at2])).get(nodes)
    first = (nodes > 10.0) & (nodes < 15.0)
    assert np.all(an2[first] < plain[first])
    assert np.allclose(an2[(nodes >= 20.0) & (nodes <= 110.0)], plain[(nodes >= 20.0) & (nodes <= 110.0)],
                       rtol=1e-14)
    p.write_text("10 1e-6\n20 2e-6 3\n")
    with pytest.raises(ValueError, match="Error parsing attenuation line"):
        read_attenuation_file(str(p))
