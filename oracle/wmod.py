"""Oracle: WModeSolver (wide-angle, high-Mach normal modes) restated in FP64 numpy/scipy.

Test infrastructure only (see oracle/__init__.py).

The reference solves the quadratic eigenvalue problem (QEP)
    (k^2 M + k C + D) v = 0                              (EigenEngine.cpp:220-250)
with D = tridiag(1/h^2, -2/h^2 + kd_i, 1/h^2), M = diag(md), C = diag(cd), through the
2N x 2N linearisation A x = k B x, x = (v, k v), A = [[-D, 0], [0, M]],
B = [[C, M], [M, 0]] (EigenEngine.cpp:252-302), by Krylov-Schur shift-invert around
sigma = (k_min + k_max)/2, then keeps the eigenvalues with k_min <= Re k <= k_max
(WModeSolver.cpp:491-506).  SLEPc returns eigenvectors of unit 2-norm in 2N, so the first
N components v satisfy |v|^2 (1 + k^2) = 1, and the code divides them by sqrt(dz)
(EigenEngine.cpp:376-381).

This restatement finds the same eigenpairs by a different route, independent of the HIP
kernels under test: T(k) = D + k C + k^2 M is real symmetric tridiagonal for real k, and
for subsonic winds dT/dk = C + 2 k M is negative definite on the trapped interval, so each
eigenvalue branch mu_i(k) of T(k) (ascending index i) decreases strictly and crosses zero
once.  The QEP eigenvalues in (k_min, k_max] are the crossings of the branches
i = n(k_min) .. n(k_max) - 1, n(k) = number of negative eigenvalues of T(k).  Each crossing
is found with Brent's method on mu_i(k), mu_i from LAPACK bisection (dstebz), and the
eigenvector from inverse iteration (dstein).  `eigen_dense` is the reference's own
linearisation solved densely (scipy.linalg.eig), the brute-force check for small N.
"""
import math

import numpy as np
import scipy.linalg
import scipy.optimize

from .atmos import PI, Steffen, AtmosphereNodes, read_profile
from . import ess


class WGrid(ess.EssGrid):
    """Grid of WModeSolver::setParams (WModeSolver.cpp:143-210) plus the matrix spacing.

    Profile samples sit at Hgt[i] = z_min + i (maxheight - z_min)/(Nz - 1) (:150, :202), as in
    ESSModeSolver; WModeSolver::solve uses dz = (maxheight - z_min)/Nz for the matrix, the
    ground term, the WKB integral, the perturbation and the source/receiver rows (:297).
    """

    def __init__(self, nodes, Nz=20000, maxheight_km=150.0, zground_km=0.0, freq=None,
                 attn_nodes=None):
        super().__init__(nodes, Nz, maxheight_km, zground_km, freq, attn_nodes)
        self.dz_m = (self.maxheight - self.z_min) / self.Nz

    def wind_component(self, azimuth):
        """_WC_ on the grid: Steffen spline of the node-level wind component."""
        az = azimuth * PI / 180.0
        n = self.nodes
        return Steffen(n.z, n.ws * np.cos(n.wd * PI / 180.0 - az)).get(self.Hgt)

    def c0(self):
        """_C0_ on the grid."""
        return self.nodes.spline("c0").get(self.Hgt)


def modal_trace(grid, freq, azimuth, zs_m, zr_m, turnoff_wkb, adm=0.0):
    """WModeSolver::getModalTrace (WModeSolver.cpp:514-629).

    kd = (omega/c0)^2, md = (wc/c0)^2 - 1, cd = -2 omega wc/c0^2, diag = (omega/c_eff)^2 at
    Hgt[i] (:549-557); the ground term 1/((dz adm + 1) dz^2) is added to kd[0] and diag[0]
    (:568-570); k_max = omega/min c_eff or the WKB cut-off (:574-609, no fallback when it
    falls below k_min, unlike ESSModeSolver.cpp:616-621); k_min = omega/c_eff(Hgt[top+1])
    (:611-620).  Returns (kd, md, cd, diag, ceffz, k_min, k_max, used_wkb).
    """
    nz, dz = grid.Nz, grid.dz_m
    omega = 2 * PI * freq
    c0 = grid.c0()
    wc = grid.wind_component(azimuth)
    ceffz = grid.ceff(azimuth)
    kd = (omega / c0) ** 2
    md = (wc / c0) ** 2 - 1
    cd = -2 * omega * (wc / c0 ** 2)
    diag = (omega / ceffz) ** 2
    ceff_grnd = ceffz[0]
    ceffmin = min(ceff_grnd, ceffz.min())
    bnd = (1.0 / (dz * adm + 1)) / (dz ** 2)
    diag[0] = bnd + diag[0]
    kd[0] = bnd + kd[0]
    used_wkb = False
    if abs(zs_m) < 1.0e-3 and abs(zr_m) < 1.0e-3 and not turnoff_wkb:
        used_wkb = True
        k_max_full = omega / ceffmin
        k_gnd = omega / ceff_grnd
        dkk = (k_max_full ** 2 - k_gnd ** 2) / 100.0
        kk = k_gnd ** 2
        if dkk > 1e-10:
            keff2 = (omega / ceffz) ** 2
            while kk < k_max_full ** 2:
                term = np.abs(kk - keff2)
                stop = np.nonzero(term <= dkk)[0]
                last = stop[0] if stop.size else nz - 1
                integral = float(np.cumsum(dz * np.sqrt(term[:last + 1]))[-1])
                if integral >= 10.0:
                    break
                kk = kk + dkk
        k_max = math.sqrt(kk)
    else:
        k_max = omega / ceffmin
    top = nz - nz // 10
    k_min = omega / ceffz[top + 1]
    return kd, md, cd, diag, ceffz, k_min, k_max, used_wkb


def pencil(kd, md, cd, dz):
    """Rows of dz^2 T(k) = dz^2 (D + k C + k^2 M): diag A + k (B + k C'), off-diagonal 1
    (EigenEngine.cpp:257-281 with the ground term already in kd[0])."""
    h2 = dz * dz
    return -2.0 + h2 * kd, h2 * cd, h2 * md


def _diag_at(P, k):
    A, B, C = P
    return A + k * (B + k * C)


def _n_negative(P, k):
    a = _diag_at(P, k)
    e = np.ones(a.size - 1)
    # eigenvalues in (0, 1e3] are few (the propagating branches); the rest are negative
    npos = scipy.linalg.eigh_tridiagonal(a, e, eigvals_only=True, select="v",
                                         select_range=(0.0, 1.0e3), lapack_driver="stebz").size
    return a.size - npos


def _mu(P, k, i):
    a = _diag_at(P, k)
    e = np.ones(a.size - 1)
    return float(scipy.linalg.eigh_tridiagonal(a, e, eigvals_only=True, select="i",
                                               select_range=(i, i), lapack_driver="stebz")[0])


def eigen(kd, md, cd, dz, k_min, k_max, vectors=True, stride=1):
    """All real QEP eigenvalues in (k_min, k_max], ascending, with the reference's vector
    normalisation v = x_{0:N} / sqrt(dz), |x|_2 = 1 (EigenEngine.cpp:376-381).
    stride > 1 computes every stride-th branch only (a timing sample).

    Returns (k [nconv], v [Nz, nconv] or None)."""
    P = pencil(kd, md, cd, dz)
    n_lo, n_hi = _n_negative(P, k_min), _n_negative(P, k_max)
    branches = range(n_lo, n_hi, stride)
    ks = []
    for i in branches:
        ks.append(scipy.optimize.brentq(lambda k: _mu(P, k, i), k_min, k_max, xtol=1e-300,
                                        rtol=4 * np.finfo(float).eps, maxiter=200))
    ks = np.array(ks)
    if not vectors:
        return ks, None
    V = np.empty((kd.size, ks.size))
    e = np.ones(kd.size - 1)
    for j, (i, k) in enumerate(zip(branches, ks)):
        _, u = scipy.linalg.eigh_tridiagonal(_diag_at(P, k), e, select="i", select_range=(i, i),
                                             lapack_driver="stebz")
        V[:, j] = u[:, 0] / math.sqrt(dz * (1.0 + k * k))
    return ks, V


def linearisation(kd, md, cd, dz):
    """The 2N x 2N pair (A, B) assembled by EigenEngine::doWideAngleCalculation
    (EigenEngine.cpp:252-302), dense."""
    n = kd.size
    h2 = dz * dz
    A = np.zeros((2 * n, 2 * n))
    B = np.zeros((2 * n, 2 * n))
    i = np.arange(n)
    A[i, i] = 2.0 / h2 - kd
    A[i[:-1], i[:-1] + 1] = -1.0 / h2
    A[i[1:], i[1:] - 1] = -1.0 / h2
    A[n + i, n + i] = md
    B[i, i] = cd
    B[i, n + i] = md
    B[n + i, i] = md
    return A, B


def eigen_dense(kd, md, cd, dz, k_min, k_max):
    """Brute force for small N: every eigenpair of A x = k B x, kept when
    k_min <= Re k <= k_max (WModeSolver.cpp:491-506); x of unit 2-norm, v = x_{0:N}/sqrt(dz).
    Returns (k, v, imag parts) sorted by Re k."""
    A, B = linearisation(kd, md, cd, dz)
    w, X = scipy.linalg.eig(A, B)
    n = kd.size
    keep = np.isfinite(w) & (w.real >= k_min) & (w.real <= k_max)
    w, X = w[keep], X[:, keep]
    o = np.argsort(w.real)
    w, X = w[o], X[:, o]
    X = X / np.linalg.norm(X, axis=0)[None, :]
    # a real eigenvector up to a unit complex phase: rotate it real
    ph = X[np.argmax(np.abs(X), axis=0), np.arange(X.shape[1])]
    X = X * (np.abs(ph) / ph)[None, :]
    return w.real, X[:n].real / math.sqrt(dz), w.imag


def perturb_weights(grid, freq):
    """doPerturb weights with the matrix spacing: dz (omega/c_T) alpha 2 (ModeSolver.cpp:276-278)."""
    omega = 2 * PI * freq
    c_T = np.sqrt(1.4 * grid.Pr / grid.rho)
    return grid.dz_m * (omega / c_T) * grid.alpha * 2


class Solution:
    pass


def solve(profile_path=None, nodes=None, freq=0.1, azimuth=90.0, Nz=20000, maxheight_km=150.0,
          zground_km=0.0, zs_km=0.0, zr_km=0.0, turnoff_wkb=False, lamb_wave_bc=False,
          wvnum_filter=None, alpha_freq=None, grid=None):
    """Modes for one (profile, f, azimuth), as WModeSolver::solve (WModeSolver.cpp:322-383):
    trace, interval (optionally the wavenumber filter, :356-360), eigenpairs, selection,
    perturbation (ModeSolver.cpp:263-284)."""
    if nodes is None:
        nodes = AtmosphereNodes(read_profile(profile_path))
    if grid is None:
        grid = WGrid(nodes, Nz, maxheight_km, zground_km)
    grid.set_attenuation(freq if alpha_freq is None else alpha_freq)
    zs, zr = zs_km * 1000.0, zr_km * 1000.0
    adm = ess.admittance(grid, lamb_wave_bc)
    kd, md, cd, diag, ceffz, k_min, k_max, used = modal_trace(grid, freq, azimuth, zs, zr,
                                                            turnoff_wkb, adm)
    if wvnum_filter is not None:
        c_min, c_max = wvnum_filter
        k_min = 2 * PI * freq / c_max
        k_max = 2 * PI * freq / c_min
    k, v = eigen(kd, md, cd, grid.dz_m, k_min, k_max)
    w = perturb_weights(grid, freq)
    k_pert = ess.perturb(k, v, w)
    sol = Solution()
    sol.grid, sol.freq, sol.azimuth = grid, freq, azimuth
    sol.kd, sol.md, sol.cd, sol.diag, sol.ceffz = kd, md, cd, diag, ceffz
    sol.k_min, sol.k_max, sol.used_wkb = k_min, k_max, used
    sol.k, sol.psi, sol.k_pert, sol.weights = k, v, k_pert, w
    sol.i_s = ess.grid_index(zs, grid.dz_m)
    sol.i_r = ess.grid_index(zr, grid.dz_m)
    return sol
