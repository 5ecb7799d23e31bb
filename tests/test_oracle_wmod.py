"""Wide-angle (WMod) oracle pinned against the reference-held goldens, the reference's own
linearisation solved densely, and closed forms."""
import math
import os

import numpy as np
import pytest

from oracle import atmos, ess, wmod

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PROFILE = os.path.join(G, "canonical_profile.dat")


@pytest.fixture(scope="module")
def nodes():
    return atmos.AtmosphereNodes(atmos.read_profile(PROFILE))


def test_wtloss_1d_goldens(nodes):
    """WMod --singleprop --freq 0.1 --azimuth 90 --write_2d_tloss (WKB off), 1000 ranges.

    Pre-2.0.1 scaling (x1000) and an older attenuation model, as for the Modess goldens:
    lossless coherent TL to 0.01 dB, lossless incoherent to 1e-6 relative, lossy incoherent
    to 1e-3, lossy coherent to 0.1 dB away from interference nulls.  The lossless
    incoherent sum pins the reference's vector normalisation sum psi^2 dz = 1/(1 + k^2) (the
    first half of a unit linearised vector): with sum psi^2 dz = 1 it is off by k^2 = 3.4e-6."""
    grid = wmod.WGrid(nodes, 20000, 150.0, 0.0)
    sol = wmod.solve(nodes=nodes, grid=grid, freq=0.1, azimuth=90.0, turnoff_wkb=True)
    g = np.loadtxt(os.path.join(G, "wtloss_1d.nm"))
    gl = np.loadtxt(os.path.join(G, "wtloss_1d.lossless.nm"))
    assert np.allclose(g[:, 0], np.arange(1, 1001))
    ps, pr = sol.psi[sol.i_s], sol.psi[sol.i_r]
    r, p, inc = ess.tl1d(sol.k_pert, ps, pr, 1000.0, 1000)
    r, pl, incl = ess.tl1d(sol.k_pert, ps, pr, 1000.0, 1000, lossless=True)
    pg = 1000 * (g[:, 1] + 1j * g[:, 2])
    pgl = 1000 * (gl[:, 1] + 1j * gl[:, 2])
    m = ess.tl_db(pgl) > -140
    assert np.max(np.abs(ess.tl_db(pl) - ess.tl_db(pgl))[m]) < 0.01
    assert np.max(np.abs(incl / (1000 * gl[:, 3]) - 1)) < 1e-6
    assert np.max(np.abs(inc / (1000 * g[:, 3]) - 1)) < 1e-3
    notnull = ess.tl_db(pg) > ess.tl_db(1000 * g[:, 3]) - 10
    assert np.max(np.abs(ess.tl_db(p) - ess.tl_db(pg))[notnull]) < 0.1
    # the check above would catch a unit-norm (ESS-style) normalisation
    s = np.sqrt(1.0 + sol.k ** 2)
    r, pl1, incl1 = ess.tl1d(sol.k_pert, ps * s, pr * s, 1000.0, 1000, lossless=True)
    assert np.max(np.abs(incl1 / (1000 * gl[:, 3]) - 1)) > 2e-6


@pytest.mark.parametrize("Nz,f,az", [(300, 0.1, 90.0), (400, 0.2, 270.0)])
def test_dense_linearisation(nodes, Nz, f, az):
    """Sturm/Brent branch crossings == scipy.linalg.eig of the reference's 2N x 2N pair
    (EigenEngine.cpp:252-302): same eigenvalues in [k_min, k_max], all real, same vectors."""
    grid = wmod.WGrid(nodes, Nz, 150.0, 0.0)
    kd, md, cd, diag, ceffz, kmin, kmax, used = wmod.modal_trace(grid, f, az, 0.0, 0.0, True)
    k, v = wmod.eigen(kd, md, cd, grid.dz_m, kmin, kmax)
    kdn, vdn, im = wmod.eigen_dense(kd, md, cd, grid.dz_m, kmin, kmax)
    assert k.size == kdn.size > 20
    assert np.all(im == 0.0)
    assert np.max(np.abs(k / kdn - 1)) < 1e-10
    assert np.max(np.abs(np.abs(v) - np.abs(vdn))) < 1e-7 * np.abs(v).max()
    assert np.allclose((v * v).sum(axis=0) * grid.dz_m * (1 + k * k), 1.0, rtol=1e-12)


def uniform_pencil(N, dz, f, c, M):
    """Constant sound speed c and wind Mach number M (wind = M c) on N rows, rigid ground."""
    om = 2 * math.pi * f
    kd = np.full(N, (om / c) ** 2)
    kd[0] += 1.0 / dz ** 2
    md = np.full(N, M * M - 1.0)
    cd = np.full(N, -2 * om * M / c)
    return kd, md, cd


def uniform_wavenumbers(N, dz, f, c, M):
    """Closed form: dz^2 T(k) = L + dz^2 ((om - k M c)^2/c^2 - k^2) with L the second
    difference with the reflecting ghost row and a Dirichlet top, eigenvalues
    -4 sin^2 th_n, th_n = (2n+1) pi / (2 (2N+1)).  Zero crossing of branch n:
    k_n = (sqrt(om^2/c^2 - (1 - M^2) s_n) - om M / c) / (1 - M^2), s_n = 4 sin^2 th_n / dz^2."""
    om = 2 * math.pi * f
    th = (2 * np.arange(N) + 1) * math.pi / (2 * (2 * N + 1))
    s = 4 * np.sin(th) ** 2 / dz ** 2
    disc = (om / c) ** 2 - (1 - M * M) * s
    k = (np.sqrt(disc[disc > 0]) - om * M / c) / (1 - M * M)
    return np.sort(k[k > 0])


@pytest.mark.parametrize("M", [0.0, 0.3, -0.3])
def test_uniform_wind_closed_form(M):
    N, dz, f, c = 1500, 50.0, 0.5, 340.0
    kd, md, cd = uniform_pencil(N, dz, f, c, M)
    ref = uniform_wavenumbers(N, dz, f, c, M)
    om = 2 * math.pi * f
    kmin, kmax = 0.5 * om / c, 1.5 * om / c
    ref = ref[(ref >= kmin) & (ref <= kmax)]
    k, v = wmod.eigen(kd, md, cd, dz, kmin, kmax)
    assert k.size == ref.size > 10
    assert np.max(np.abs(k / ref - 1)) < 1e-12


def test_zero_wind_is_ess(nodes):
    """Without wind T(k) = D - k^2 I: k^2 are the eigenvalues of the ESS operator D (here
    with a regular last row) and psi sqrt(1 + k^2) its unit eigenvectors / sqrt(dz)."""
    grid = wmod.WGrid(nodes, 2000, 150.0, 0.0)
    f = 0.2
    om = 2 * math.pi * f
    c0 = grid.c0()
    kd = (om / c0) ** 2
    kd[0] += 1.0 / grid.dz_m ** 2
    md = -np.ones_like(kd)
    cd = np.zeros_like(kd)
    kmin, kmax = om / c0.max(), om / c0.min()
    k, v = wmod.eigen(kd, md, cd, grid.dz_m, kmin, kmax)
    import scipy.linalg
    a = -2.0 / grid.dz_m ** 2 + kd
    e = np.full(a.size - 1, 1.0 / grid.dz_m ** 2)
    lam, u = scipy.linalg.eigh_tridiagonal(a, e, select="v", select_range=(kmin ** 2, kmax ** 2))
    assert k.size == lam.size > 10
    assert np.max(np.abs(k * k / lam - 1)) < 1e-11
    ue = np.abs(u) / math.sqrt(grid.dz_m)
    assert np.max(np.abs(np.abs(v) * np.sqrt(1 + k * k) - ue)) < 1e-8 * ue.max()
