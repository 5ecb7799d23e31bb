"""Oracle: ESSModeSolver restated in FP64 numpy/scipy.

Test infrastructure only (see oracle/__init__.py).  Eigenpairs come from LAPACK
(scipy.linalg.eigh_tridiagonal: dstebz bisection + dstein inverse iteration), which is
independent of the HIP kernels under test.
"""
import math
import numpy as np
import scipy.linalg

from .atmos import PI, Steffen, AtmosphereNodes, read_profile

MAX_MODES = 4000  # src/libmodes/ESSModeSolver.cpp:17


class EssGrid:
    """Grid samples built by ESSModeSolver::setParams (ESSModeSolver.cpp:30-240).

    maxheight is clamped to the profile top - 1e-6 m (:144-148), z_min is raised to the
    profile's Z0 (:212-220), dz = (maxheight - z_min)/(Nz - 1) (:151), and rho, P, T, u, v
    and alpha(freq) are Steffen-spline samples at Hgt[i] = z_min + i dz (:231-239).
    """

    def __init__(self, nodes, Nz=20000, maxheight_km=150.0, zground_km=0.0, freq=None,
                 attn_nodes=None, legacy_dz=False):
        self.nodes = nodes
        self.Nz = int(Nz)
        z_min = zground_km * 1000.0
        maxheight = maxheight_km * 1000.0
        ztop = nodes.z[-1]
        if maxheight >= ztop:
            maxheight = ztop - 1e-6
        if nodes.Z0 is not None and z_min < nodes.Z0:
            z_min = nodes.Z0
        self.z_min = z_min
        self.maxheight = maxheight
        # legacy_dz: dz = H/Nz, the grid of the older ncpaprop that wrote the reference's
        # dispersion goldens (it reproduces their low-frequency wavenumbers to ~2e-9)
        self.dz = (maxheight - z_min) / (self.Nz if legacy_dz else self.Nz - 1)
        self.Hgt = z_min + np.arange(self.Nz) * self.dz
        self.rho = nodes.spline("rho").get(self.Hgt)
        self.Pr = nodes.spline("P").get(self.Hgt)
        self.T = nodes.spline("T").get(self.Hgt)
        self.zw = nodes.spline("u").get(self.Hgt)
        self.mw = nodes.spline("v").get(self.Hgt)
        self.freq = freq
        self.alpha = None
        if freq is not None or attn_nodes is not None:
            self.set_attenuation(freq, attn_nodes)

    def set_attenuation(self, freq, attn_nodes=None):
        """_ALPHA_ sampled on the grid; attn_nodes overrides Sutherland-Bass at the nodes."""
        a_nodes = attn_nodes if attn_nodes is not None else self.nodes.alpha_nodes(freq)
        self.alpha = Steffen(self.nodes.z, a_nodes).get(self.Hgt)
        return self.alpha

    def ceff(self, azimuth):
        """c_eff on the grid: Steffen spline of the node-level _CEFF_ (ESSModeSolver.cpp:545)."""
        return Steffen(self.nodes.z, self.nodes.ceff_nodes(azimuth)).get(self.Hgt)


def admittance(grid, lamb_wave_bc):
    """Ground admittance (ESSModeSolver.cpp:344-357): 0, or -rho'/rho/2 at z_min."""
    if not lamb_wave_bc:
        return 0.0
    sp = grid.nodes.spline("rho")
    return float(-sp.deriv(grid.z_min) / sp.get(grid.z_min) / 2.0)


def modal_trace(grid, freq, azimuth, zs_m, zr_m, turnoff_wkb, adm=0.0):
    """ESSModeSolver::getModalTrace (ESSModeSolver.cpp:518-645).

    Returns (diag, ceffz, k_min, k_max, used_wkb).  diag[i] = (omega/c_eff)^2 plus the
    ghost-node term 1/((dz adm + 1) dz^2) on row 0; k_min = omega/c_eff(Hgt[top+1]) with
    top = nz - nz/10; k_max = omega/min c_eff or the WKB cut-off for ground-to-ground runs.
    """
    nz, dz = grid.Nz, grid.dz
    omega = 2 * PI * freq
    ceffz = grid.ceff(azimuth)
    diag = (omega / ceffz) ** 2
    ceff_grnd = ceffz[0]
    ceffmin = min(ceff_grnd, ceffz.min())
    diag[0] = (1.0 / (dz * adm + 1)) / (dz ** 2) + diag[0]
    used_wkb = False
    if abs(zs_m) < 1.0e-3 and abs(zr_m) < 1.0e-3 and not turnoff_wkb:
        used_wkb = True
        k_max = math.sqrt(wkb_cutoff(ceffz, omega, dz)[0])
    else:
        k_max = omega / ceffmin
    top = nz - nz // 10
    k_min = omega / ceffz[top + 1]
    if used_wkb and k_max < k_min:
        k_max = omega / ceffmin
    return diag, ceffz, k_min, k_max, used_wkb


def wkb_cutoff(ceffz, omega, dz):
    """The WKB cut-off loop of getModalTrace (ESSModeSolver.cpp:561-607).

    kk steps from k_gnd^2 by dkk = (k_max_full^2 - k_gnd^2)/100, k_gnd = omega/c_eff[0] and
    k_max_full = omega/min c_eff, and stops at the first kk whose integral
    sum dz sqrt|kk - k_eff^2| over rows 0 .. first (the lowest row with |kk - k_eff^2| <= dkk,
    inclusive, as `while (wkbTerm > dkk)` runs its body for it) reaches 10.  Returns (kk, steps)
    with steps = [(kk, first, integral), ...] in the order tried; k_max = sqrt(kk).
    """
    ceffz = np.asarray(ceffz, dtype=np.float64)
    nz = ceffz.size
    ceffmin = min(ceffz[0], ceffz.min())
    k_max_full = omega / ceffmin
    k_gnd = omega / ceffz[0]
    dkk = (k_max_full ** 2 - k_gnd ** 2) / 100.0
    kk = k_gnd ** 2
    steps = []
    if dkk > 1.0e-10:
        keff2 = (omega / ceffz) ** 2
        while kk < k_max_full ** 2:
            term = np.abs(kk - keff2)
            stop = np.nonzero(term <= dkk)[0]
            first = int(stop[0]) if stop.size else nz - 1
            if dkk >= 1.0:
                integral = 0.0   # `while (wkbTerm > dkk)` with wkbTerm = 1 never runs
            else:
                # sequential left-to-right sum, as the reference loop accumulates
                integral = float(np.cumsum(dz * np.sqrt(term[:first + 1]))[-1])
            steps.append((kk, first, integral))
            if integral >= 10.0:
                break
            kk = kk + dkk
    return kk, steps


def matrix(diag, dz):
    """Tridiagonal of EigenEngine::doESSCalculation (EigenEngine.cpp:57-79).

    a[i] = -2/h^2 + diag[i] for i < Nz-1; the last row re-uses diag[Nz-2] (the stale
    value[1] of the assembly loop); off-diagonal 1/h^2.
    """
    h2 = dz * dz
    a = -2.0 / h2 + diag
    a[-1] = -2.0 / h2 + diag[-2]
    e = np.full(diag.size - 1, 1.0 / h2)
    return a, e


def eigen(diag, dz, k_min, k_max, vectors=True):
    """All eigenpairs with k^2 in [k_min^2, k_max^2] (EigenEngine.cpp:94-161).

    Eigenvalues ascending (the order of both reference dispersion goldens);
    psi = x/sqrt(dz) so that sum(psi^2) dz = 1 (EigenEngine.cpp:158).
    Returns (k2 [nconv], psi [Nz, nconv] or None).
    """
    a, e = matrix(diag, dz)
    lo, hi = k_min ** 2, k_max ** 2
    if vectors:
        w, v = scipy.linalg.eigh_tridiagonal(a, e, select="v", select_range=(lo, hi),
                                             lapack_driver="stebz")
        return w, v / math.sqrt(dz)
    w = scipy.linalg.eigh_tridiagonal(a, e, eigvals_only=True, select="v",
                                      select_range=(lo, hi), lapack_driver="stebz")
    return w, None


def sturm_count(a, e2, x):
    """Number of eigenvalues of the tridiagonal below x (LDL^T inertia, numpy)."""
    d = a[0] - x
    cnt = int(d < 0)
    for i in range(1, a.size):
        d = (a[i] - x) - e2 / d if d != 0.0 else (a[i] - x) - e2 / -1e-300
        cnt += int(d < 0)
    return cnt


def select(k2, psi, k_min, k_max):
    """ESSModeSolver::doSelect (ESSModeSolver.cpp:647-665): keep k_min <= k <= k_max."""
    k = np.sqrt(k2)
    keep = (k >= k_min) & (k <= k_max)
    return k[keep], (psi[:, keep] if psi is not None else None)


def perturb_weights(grid, freq):
    """w_i = dz (omega/c_T,i) alpha_i 2 with c_T = sqrt(1.4 P/rho) (ModeSolver.cpp:276-278)."""
    omega = 2 * PI * freq
    c_T = np.sqrt(1.4 * grid.Pr / grid.rho)
    return grid.dz * (omega / c_T) * grid.alpha * 2


def perturb(k, psi, weights):
    """ModeSolver::doPerturb (ModeSolver.cpp:263-284): k_pert = sqrt(k^2 + i sum w psi^2)."""
    absorption = (psi * psi * weights[:, None]).sum(axis=0)
    return np.sqrt(k * k + 1j * absorption)


def perturb_from_absorption(k, absorption):
    return np.sqrt(np.asarray(k) ** 2 + 1j * np.asarray(absorption))


EXPOV8PI = 4 * PI * 1j * np.exp(-1j * PI * 0.25) / math.sqrt(8.0 * PI)


def grid_index(z_m, dz):
    """n = (int) ceil(z/dz) (ModeSolver.cpp:306-307)."""
    return int(math.ceil(z_m / dz))


def tl1d(k_pert, psi_s, psi_r, dr, n_r, lossless=False, chunk=2048):
    """ModeSolver::getTLoss1D sums (ModeSolver.cpp:325-356).

    Returns (r [m], p complex [n_r], incoherent [n_r]); lossless uses Re(k_pert).
    """
    k = np.asarray(k_pert, dtype=np.complex128)
    prod = np.asarray(psi_s) * np.asarray(psi_r)
    r = (np.arange(n_r) + 1) * dr
    p = np.empty(n_r, dtype=np.complex128)
    inc = np.empty(n_r)
    if lossless:
        kr = k.real
        c = prod / np.sqrt(kr)
        ci = prod ** 2 / kr
    else:
        c = prod / np.sqrt(k)
        ci = prod ** 2 / np.abs(k)
    for s in range(0, n_r, chunk):
        rr = r[s:s + chunk]
        if lossless:
            ph = np.exp(1j * np.outer(rr, kr))
            p[s:s + chunk] = ph @ c
            inc[s:s + chunk] = ci.sum() * np.ones_like(rr)
        else:
            ph = np.exp(1j * np.outer(rr, k))
            p[s:s + chunk] = ph @ c
            inc[s:s + chunk] = np.exp(-2 * np.outer(rr, k.imag)) @ ci
    p = 1000.0 * EXPOV8PI * p / np.sqrt(r)
    inc = 1000.0 * 4 * PI * np.sqrt(inc) * np.sqrt(1. / 8. / PI / r)
    return r, p, inc


def tl2d_rows(nz):
    """Rows written by getTLoss2D: j = 0, stepj, ... with stepj = nz/500 or 10 (ModeSolver.cpp:513-526)."""
    stepj = nz // 500
    if stepj == 0:
        stepj = 10
    return np.arange(0, nz, stepj)


def tl2d(k_pert, psi_rows, psi_s, dr, n_r, lossless=False, chunk=1024):
    """ModeSolver::getTLoss2D field (ModeSolver.cpp:524-544) at r = dr, 2 dr, ..: p[r, row]."""
    r = (np.arange(n_r) + 1) * dr
    return r, tl2d_at(k_pert, psi_rows, psi_s, r, lossless, chunk)


def tl2d_at(k_pert, psi_rows, psi_s, r, lossless=False, chunk=1024):
    """The getTLoss2D sum at the ranges r [m] (any subset of the reference's range steps):
    p[r, row] = 1000 expov8pi sum_m psi_m(zs) psi_m(row) e^{i k_m r} / sqrt(k_m) / sqrt(r)."""
    k = np.asarray(k_pert, dtype=np.complex128)
    kk = k.real if lossless else k
    r = np.asarray(r, dtype=np.float64)
    out = np.empty((r.size, psi_rows.shape[0]), dtype=np.complex128)
    c = np.asarray(psi_s) / np.sqrt(kk.astype(np.complex128))
    for s in range(0, r.size, chunk):
        rr = r[s:s + chunk]
        E = np.exp(1j * np.outer(rr, kk)) * c[None, :]
        out[s:s + chunk] = E @ psi_rows.T
    out *= (1000.0 * EXPOV8PI / np.sqrt(r))[:, None]
    return out


def tl_db(p):
    return 20.0 * np.log10(np.maximum(np.abs(p), 1e-300))


# --------------------------------------------------------------------------------------
# writers (exact printf formats of the reference)
# --------------------------------------------------------------------------------------
def format_tl1d(r, p, inc):
    """'%f %20.12e %20.12e %20.12e' per range (ModeSolver.cpp:367-368)."""
    return "".join("%f %20.12e %20.12e %20.12e\n" % (ri / 1000.0, pi.real, pi.imag, ii)
                   for ri, pi, ii in zip(r, p, inc))


def format_tl1d_nx2(az, r, p, inc):
    """'%10.3f %8.3f %20.12e %20.12e %20.12e' + blank line per azimuth (ModeSolver.cpp:475-482)."""
    return "".join("%10.3f %8.3f %20.12e %20.12e %20.12e\n" % (ri / 1000.0, az, pi.real,
                                                               pi.imag, ii)
                   for ri, pi, ii in zip(r, p, inc)) + "\n"


def format_tl2d(r, z_rows_m, field):
    """'%f %f %15.8e %15.8e' with a blank line per range (ModeSolver.cpp:546-553)."""
    out = []
    for i, ri in enumerate(r):
        for j, zj in enumerate(z_rows_m):
            v = field[i, j]
            out.append("%f %f %15.8e %15.8e\n" % (ri / 1000.0, zj / 1000.0, v.real, v.imag))
        out.append("\n")
    return "".join(out)


def format_dispersion(freq, k_pert, psi_s, psi_r, rho_s, rho_r):
    """One writeDispersion line (ModeSolver.cpp:580-585)."""
    s = ["%.7e   %d    %.7e   %.7e" % (freq, len(k_pert), rho_s, rho_r)]
    for kp, a, b in zip(k_pert, psi_s, psi_r):
        s.append("   %.12e   %.12e" % (kp.real, kp.imag))
        s.append("   %.12e   %.12e" % (a, b))
    return "".join(s) + "\n"


def format_phase_speeds(freq, k_pert):
    """writePhaseSpeeds (ModeSolver.cpp:608-610)."""
    return "".join("%d %f %15.8e\n" % (j, (2 * PI * freq) / kp.real, kp.imag)
                   for j, kp in enumerate(k_pert))


def modal_starter(psi, i_s, k_pert, dz, freq):
    """ESSModeSolver::getModalStarter (ESSModeSolver.cpp:670-723): rows j = 0, z_cnd, 2 z_cnd ..
    with z_cnd = int((340/f)/10/dz); starter(z_j) = pi sqrt(k0) sum_m psi_m(zs) psi_m(z_j) /
    sqrt(Re k_m), k0 = 2 pi f/340.  Returns (z [m], starter [rows]) (real: the reference's
    complex sum has a zero imaginary part)."""
    k0 = 2 * PI * freq / 340.0
    z_cnd = int((340.0 / freq) / 10 / dz)
    rows = np.arange(0, psi.shape[0], z_cnd)
    s = psi[rows] @ (psi[i_s] / np.sqrt(np.real(k_pert)))
    return rows * dz, s * PI * math.sqrt(k0)


def format_modal_starter(z_m, starter):
    """'%10.3f   %16.12e   %16.12e' per row and a closing blank line (ESSModeSolver.cpp:718-720)."""
    return "".join("%10.3f   %16.12e   %16.12e\n" % (z / 1000.0, v, 0.0) for z, v in zip(z_m, starter)) + "\n"


def group_speeds(psi, ceffz, dz, freq, k_pert):
    """writePhaseAndGroupSpeeds (ESSModeSolver.cpp:735-745): (v_phase, v_group)."""
    omega = 2 * PI * freq
    vph = omega / np.real(k_pert)
    vg = (psi * psi / (ceffz * ceffz)[:, None]).sum(axis=0)
    vg = 1.0 / (vg * vph * dz)
    return vph, vg


# --------------------------------------------------------------------------------------
# one full single-frequency solve
# --------------------------------------------------------------------------------------
class Solution:
    pass


def solve(profile_path=None, nodes=None, freq=0.1, azimuth=90.0, Nz=20000, maxheight_km=150.0,
          zground_km=0.0, zs_km=0.0, zr_km=0.0, turnoff_wkb=False, lamb_wave_bc=False,
          wvnum_filter=None, alpha_freq=None, vectors=True, grid=None, attn_nodes=None):
    """Modes for one (profile, f, azimuth), as ESSModeSolver::solve (ESSModeSolver.cpp:320-382)."""
    if nodes is None:
        nodes = AtmosphereNodes(read_profile(profile_path))
    if grid is None:
        grid = EssGrid(nodes, Nz, maxheight_km, zground_km)
    # attn_nodes: _ALPHA_ at the profile nodes from --use_attn_file (ESSModeSolver.cpp:224-228)
    grid.set_attenuation(freq if alpha_freq is None else alpha_freq, attn_nodes)
    zs, zr = zs_km * 1000.0, zr_km * 1000.0
    adm = admittance(grid, lamb_wave_bc)
    diag, ceffz, k_min, k_max, used = modal_trace(grid, freq, azimuth, zs, zr, turnoff_wkb, adm)
    if wvnum_filter is not None:
        c_min, c_max = wvnum_filter
        k_min = 2 * PI * freq / c_max
        k_max = 2 * PI * freq / c_min
    k2, psi = eigen(diag, grid.dz, k_min, k_max, vectors=True)
    k, psi_s = select(k2, psi, k_min, k_max)
    w = perturb_weights(grid, freq)
    k_pert = perturb(k, psi_s, w)
    sol = Solution()
    sol.grid, sol.freq, sol.azimuth = grid, freq, azimuth
    sol.diag, sol.ceffz, sol.k_min, sol.k_max, sol.used_wkb = diag, ceffz, k_min, k_max, used
    sol.k, sol.psi, sol.k_pert, sol.weights = k, psi_s, k_pert, w
    sol.i_s = grid_index(zs, grid.dz)
    sol.i_r = grid_index(zr, grid.dz)
    return sol
