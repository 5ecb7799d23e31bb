"""Oracle pinned by closed forms, brute force and invariants."""
import math

import numpy as np
import pytest

from oracle import ess, atmos


def test_constant_ceff_closed_form():
    """rigid-ground ghost node + Dirichlet top: lambda_j = k0^2 - (4/h^2) sin^2((2j+1) pi / (2(2n+1)))."""
    n, dz, c, f = 3000, 50.0, 340.0, 0.3
    k0 = 2 * math.pi * f / c
    diag = np.full(n, k0 * k0)
    diag[0] += 1.0 / dz ** 2
    k2, _ = ess.eigen(diag, dz, 0.5 * k0, k0, vectors=False)
    j = np.arange(n)
    lam = k0 ** 2 - (4 / dz ** 2) * np.sin((2 * j + 1) * math.pi / (2 * (2 * n + 1))) ** 2
    lam = np.sort(lam[(lam >= 0.25 * k0 * k0) & (lam <= k0 * k0)])
    assert k2.size == lam.size
    assert np.max(np.abs(k2 - lam)) < 64 * 2.2e-16 * 4 / dz ** 2


def test_dense_brute_force(synth_profile):
    sol = ess.solve(synth_profile, freq=0.05, azimuth=30.0, Nz=600, turnoff_wkb=True)
    a, e = ess.matrix(sol.diag, sol.grid.dz)
    T = np.diag(a) + np.diag(e, 1) + np.diag(e, -1)
    w, V = np.linalg.eigh(T)
    sel = (w >= sol.k_min ** 2) & (w <= sol.k_max ** 2)
    assert np.allclose(np.sqrt(w[sel]), sol.k, rtol=1e-12)
    ref = V[0, sel] ** 2 / sol.grid.dz
    assert np.allclose(sol.psi[0] ** 2, ref, rtol=1e-8, atol=1e-12 * ref.max())


def test_last_row_quirk():
    d = np.arange(1.0, 7.0)
    a, e = ess.matrix(d, 1.0)
    assert a[-1] == -2.0 + d[-2] and a[-2] == -2.0 + d[-2] and a[0] == -2.0 + d[0]


def test_sturm_count_matches(synth_profile):
    sol = ess.solve(synth_profile, freq=0.05, azimuth=0.0, Nz=500, turnoff_wkb=True)
    a, e = ess.matrix(sol.diag, sol.grid.dz)
    lo, hi = sol.k_min ** 2, sol.k_max ** 2
    n = ess.sturm_count(a, e[0] ** 2, hi) - ess.sturm_count(a, e[0] ** 2, lo)
    assert n == sol.k.size


def test_normalisation_and_perturbation_closed_form(synth_profile):
    sol = ess.solve(synth_profile, freq=0.1, azimuth=90.0, Nz=2000, turnoff_wkb=True)
    dz = sol.grid.dz
    assert np.allclose((sol.psi ** 2).sum(axis=0) * dz, 1.0, atol=1e-10)
    c = 3e-7
    kp = ess.perturb(sol.k, sol.psi, np.full(sol.psi.shape[0], c * dz))
    assert np.allclose(kp, np.sqrt(sol.k ** 2 + 1j * c), rtol=1e-12)


def test_tl1d_single_mode_closed_form():
    k = np.array([0.01 + 2e-6j])
    ps, pr = np.array([0.003]), np.array([0.002])
    r, p, inc = ess.tl1d(k, ps, pr, 500.0, 7)
    ev = 4 * math.pi * 1j * np.exp(-1j * math.pi / 4) / math.sqrt(8 * math.pi)
    ref = 1000 * ev * ps * pr * np.exp(1j * k * r) / np.sqrt(k) / np.sqrt(r)
    assert np.allclose(p, ref, rtol=1e-13)
    refi = 1000 * 4 * math.pi * np.abs(ps * pr) * np.exp(-k.imag * r) / np.sqrt(abs(k[0])) / np.sqrt(8 * math.pi * r)
    assert np.allclose(inc, refi, rtol=1e-13)


def test_tl2d_row_equals_tl1d(synth_profile):
    sol = ess.solve(synth_profile, freq=0.1, azimuth=90.0, Nz=2000, turnoff_wkb=True)
    rows = np.array([0, 7, 31])
    r, f2 = ess.tl2d(sol.k_pert, sol.psi[rows], sol.psi[0], 2000.0, 50)
    _, p1, _ = ess.tl1d(sol.k_pert, sol.psi[0], sol.psi[0], 2000.0, 50)
    assert np.allclose(f2[:, 0], p1, rtol=1e-12)


def test_steffen_properties():
    x = np.array([0.0, 1.0, 2.5, 3.0, 4.5, 6.0])
    y = np.array([0.0, 0.5, 0.6, 2.0, 2.1, 5.0])   # monotone
    s = atmos.Steffen(x, y)
    assert np.allclose(s.get(x), y, rtol=0, atol=1e-14)
    xx = np.linspace(0, 6, 2001)
    assert np.all(np.diff(s.get(xx)) >= -1e-14)
    lin = atmos.Steffen(x, 3 * x - 1)
    assert np.allclose(lin.get(xx), 3 * xx - 1, atol=1e-12)
    eps = 1e-7   # C1 at interior nodes
    for xi in x[1:-1]:
        assert abs(s.deriv(xi - eps) - s.deriv(xi + eps)) < 1e-5
    assert s.get(-1.0) == y[0] and s.get(7.0) == y[-1]


def test_ideal_gas_sound_speed(synth_profile):
    nodes = atmos.AtmosphereNodes(atmos.read_profile(synth_profile))
    assert np.allclose(nodes.c0, np.sqrt(1.4 * 287.0 * nodes.T), rtol=1e-9)
    assert abs(nodes.T[-1] - 440.5) < 1e-6


def test_wind_component_east(synth_profile):
    nodes = atmos.AtmosphereNodes(atmos.read_profile(synth_profile))
    assert np.allclose(nodes.ceff_nodes(90.0) - nodes.c0, nodes.u, atol=1e-9)
    assert np.allclose(nodes.ceff_nodes(0.0) - nodes.c0, nodes.v, atol=1e-9)
