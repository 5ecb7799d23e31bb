"""GPU parity: batched ESS modes and 1-D TL through the C ABI vs the FP64 oracle."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle(path, f, az, zs=0.0, zr=0.0, wkb=False):
    from oracle import ess
    return ess.solve(path, freq=f, azimuth=az, zs_km=zs, zr_km=zr, turnoff_wkb=not wkb)


@pytest.mark.parametrize("f,az,zs", [(0.1, 90.0, 0.0), (0.5, 45.0, 0.0), (0.25, 270.0, 20.0)])
def test_modes_vs_oracle(gpu_ctx, synth_profile, f, az, zs):
    import ncpamodes
    print("case f=%g az=%g zs=%g" % (f, az, zs), flush=True)
    pid, info = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, f, az, zs * 1000, zs * 1000)])
    k, psr, ps, pr = m.problem_modes(0)
    sol = _oracle(synth_profile, f, az, zs, zs)
    assert k.size == sol.k.size, (k.size, sol.k.size)
    # unperturbed k: 1e-9 relative (plus the FP64 floor 64 u ||T|| on k^2)
    k2 = (k ** 2).real
    k2o = sol.k ** 2
    normT = 4.0 / g.dz ** 2
    assert np.all(np.abs(k2 - k2o) <= 1e-9 * k2o + 64 * 2.2e-16 * normT)
    # perturbed imaginary part
    assert np.allclose(k.imag, sol.k_pert.imag, rtol=1e-4, atol=1e-14)
    # psi(zs) psi(zr), sign-free, absolute to the largest product
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(psr - ref)) <= 1e-6 * np.max(np.abs(ref))


def test_tl1d_vs_oracle(gpu_ctx, synth_profile):
    import ncpamodes
    from oracle import ess
    f, az = 0.5, 90.0
    print("case tl1d f=%g az=%g" % (f, az), flush=True)
    pid, info = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, f, az)])
    nr, dr = 10000, 100.0
    p, inc = m.tl1d(dr, nr)
    p, inc = p[0], inc[0]
    sol = _oracle(synth_profile, f, az)
    r, po, inco = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], dr, nr)
    tl, tlo = ess.tl_db(p), ess.tl_db(po)
    mask = tlo > -140.0
    assert np.max(np.abs(tl - tlo)[mask]) <= 0.1
    assert np.max(np.abs(inc / inco - 1)) <= 1e-6


def test_rows_vs_oracle(gpu_ctx, synth_profile):
    """Eigenvectors at 500 rows (the reference's 2-D row set) vs LAPACK, to 1e-7 of max |psi|.

    Row 0 is the source row, where the convention psi(zs) >= 0 holds."""
    import ncpamodes
    f, az = 0.5, 90.0
    print("case rows f=%g az=%g" % (f, az), flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, f, az)])
    stride = g.Nz // 500
    rows = np.arange(0, g.Nz, stride)
    U = m.rows(0, stride, rows.size)
    sol = _oracle(synth_profile, f, az)
    ref = sol.psi[rows]
    s = np.sign((U * ref).sum(axis=0))
    err = np.abs(U * s - ref)
    kk, jj = np.unravel_index(np.argmax(err), err.shape)
    print("worst row %d mode %d: ours %.12e ref %.12e" % (rows[kk], jj, U[kk, jj] * s[jj], ref[kk, jj]), flush=True)
    assert np.max(err) <= 1e-7 * np.max(np.abs(ref))
    assert np.all(U[0] >= 0.0)
    assert np.allclose(U[0], np.abs(ref[0]), rtol=1e-6, atol=1e-9 * np.abs(ref).max())


def test_ess_eigen_dropin(gpu_ctx, synth_profile):
    """ncpa_ess_eigen == EigenEngine::doESSCalculation: k^2, orthonormal vectors, vs LAPACK."""
    import ncpamodes
    from oracle import ess
    f = 0.3
    print("case ess_eigen f=%g" % f, flush=True)
    sol = _oracle(synth_profile, f, 135.0)
    dz = sol.grid.dz
    k2, V = ncpamodes.ess_eigen(gpu_ctx, sol.diag, dz, sol.k_min, sol.k_max)
    assert k2.size == sol.k.size
    normT = 4.0 / dz ** 2
    assert np.all(np.abs(k2 - sol.k ** 2) <= 1e-9 * sol.k ** 2 + 64 * 2.2e-16 * normT)
    G = (V.T @ V) * dz
    assert np.max(np.abs(G - np.eye(k2.size))) < 1e-8
    s = np.sign((V * sol.psi).sum(axis=0))
    assert np.max(np.abs(V * s - sol.psi)) < 1e-7 * np.max(np.abs(sol.psi))
