"""GPU parity: wide-angle (WMod) modes through the C ABI vs the FP64 oracle, the reference
goldens and closed forms."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CANON = os.path.join(G, "canonical_profile.dat")


def _uniform(N, dz, f, c, M):
    from test_oracle_wmod import uniform_pencil, uniform_wavenumbers
    return uniform_pencil(N, dz, f, c, M), uniform_wavenumbers(N, dz, f, c, M)


@pytest.mark.parametrize("M", [0.0, 0.3, -0.3])
def test_eigen_uniform_wind_closed_form(gpu_ctx, M):
    """ncpa_wmod_eigen on constant c and wind: k to 1e-12 of the closed form, every vector
    normalised as the reference's (sum psi^2 dz (1 + k^2) = 1)."""
    import ncpamodes
    N, dz, f, c = 1500, 50.0, 0.5, 340.0
    print("case uniform M=%g" % M, flush=True)
    (kd, md, cd), ref = _uniform(N, dz, f, c, M)
    om = 2 * math.pi * f
    kmin, kmax = 0.5 * om / c, 1.5 * om / c
    ref = ref[(ref >= kmin) & (ref <= kmax)]
    k, v = ncpamodes.wmod_eigen(gpu_ctx, kd, md, cd, dz, kmin, kmax)
    assert k.size == ref.size > 10
    assert np.max(np.abs(k / ref - 1)) < 1e-12
    assert np.allclose((v * v).sum(axis=0) * dz * (1 + k * k), 1.0, rtol=1e-10)


def test_eigen_vs_oracle(gpu_ctx):
    """Canonical profile, 2000 rows, 0.2 Hz, az 270: k to 1e-11, |psi| to 1e-8 of max."""
    import ncpamodes
    from oracle import atmos, wmod
    print("case eigen canonical N=2000 f=0.2", flush=True)
    nodes = atmos.AtmosphereNodes(atmos.read_profile(CANON))
    grid = wmod.WGrid(nodes, 2000, 150.0, 0.0)
    kd, md, cd, diag, ceffz, kmin, kmax, used = wmod.modal_trace(grid, 0.2, 270.0, 0.0, 0.0, True)
    ko, vo = wmod.eigen(kd, md, cd, grid.dz_m, kmin, kmax)
    k, v = ncpamodes.wmod_eigen(gpu_ctx, kd, md, cd, grid.dz_m, kmin, kmax)
    assert k.size == ko.size > 20
    assert np.max(np.abs(k / ko - 1)) < 1e-11
    assert np.max(np.abs(np.abs(v) - np.abs(vo))) < 1e-8 * np.abs(vo).max()


def test_eigen_zero_wind_matches_ess_eigen(gpu_ctx):
    """Without wind the wide-angle pencil is D - k^2 I: ncpa_wmod_eigen k^2 == ncpa_ess_eigen
    eigenvalues and psi sqrt(1 + k^2) == the ESS vectors (the last ESS row re-uses diag[N-2],
    so the test operator has equal last two diagonal entries)."""
    import ncpamodes
    print("case zero wind", flush=True)
    N, dz, f = 3000, 40.0, 0.3
    om = 2 * math.pi * f
    z = np.arange(N) * dz
    c = 300.0 + 40.0 * np.exp(-((z - 50e3) / 10e3) ** 2) + 0.0008 * z
    kd = (om / c) ** 2
    kd[-1] = kd[-2]
    diag = kd.copy()
    kd[0] += 1.0 / dz ** 2
    diag[0] += 1.0 / dz ** 2
    kmin, kmax = om / c.max(), om / c.min()
    k, v = ncpamodes.wmod_eigen(gpu_ctx, kd, -np.ones(N), np.zeros(N), dz, kmin, kmax)
    k2, u = ncpamodes.ess_eigen(gpu_ctx, diag, dz, kmin, kmax)
    assert k.size == k2.size > 10
    assert np.max(np.abs(k * k / k2 - 1)) < 1e-12
    assert np.max(np.abs(np.abs(v) * np.sqrt(1 + k * k)[None, :] - np.abs(u))) < 1e-9 * np.abs(u).max()


@pytest.fixture(scope="module")
def canon_01(gpu_ctx):
    """WMod canonical run of the reference goldens: 0.1 Hz, az 90, WKB off, 20000 rows."""
    import ncpamodes
    from oracle import atmos, wmod
    pid, info = gpu_ctx.read_profile(CANON)
    g = gpu_ctx.grid(pid)
    m = gpu_ctx.solve_wide(g, [ncpamodes.problem(pid, 0.1, 90.0)])
    nodes = atmos.AtmosphereNodes(atmos.read_profile(CANON))
    sol = wmod.solve(nodes=nodes, freq=0.1, azimuth=90.0, turnoff_wkb=True)
    return m, sol


def test_solve_vs_oracle(canon_01):
    m, sol = canon_01
    print("case solve canonical 0.1 Hz", flush=True)
    k, psr, ps, pr = m.problem_modes(0)
    assert k.size == sol.k.size
    assert np.max(np.abs(np.sqrt((k ** 2).real) / sol.k - 1)) < 1e-10
    assert np.allclose(k.imag, sol.k_pert.imag, rtol=1e-6, atol=1e-16)
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(psr - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_tl1d_vs_goldens(canon_01):
    """The same bounds as the oracle's golden test: lossless coherent 0.01 dB, lossless
    incoherent 1e-6 relative (pins the 1/(1 + k^2) vector normalisation)."""
    from oracle import ess
    m, sol = canon_01
    print("case tl1d goldens", flush=True)
    g = np.loadtxt(os.path.join(G, "wtloss_1d.nm"))
    gl = np.loadtxt(os.path.join(G, "wtloss_1d.lossless.nm"))
    p, inc = m.tl1d(1000.0, 1000)
    pl, incl = m.tl1d(1000.0, 1000, lossless=True)
    pg = 1000 * (g[:, 1] + 1j * g[:, 2])
    pgl = 1000 * (gl[:, 1] + 1j * gl[:, 2])
    msk = ess.tl_db(pgl) > -140
    assert np.max(np.abs(ess.tl_db(pl[0]) - ess.tl_db(pgl))[msk]) < 0.01
    assert np.max(np.abs(incl[0] / (1000 * gl[:, 3]) - 1)) < 1e-6
    assert np.max(np.abs(inc[0] / (1000 * g[:, 3]) - 1)) < 1e-3
    notnull = ess.tl_db(pg) > ess.tl_db(1000 * g[:, 3]) - 10
    assert np.max(np.abs(ess.tl_db(p[0]) - ess.tl_db(pg))[notnull]) < 0.1
    r, po, inco = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], 1000.0, 1000)
    assert np.max(np.abs(ess.tl_db(p[0]) - ess.tl_db(po))[ess.tl_db(po) > -140]) < 1e-4


def test_rows_and_field_vs_oracle(canon_01):
    """Eigenvector rows (the reference's 500 2-D rows) to 1e-7 of max |psi|; the FP64 2-D
    field at the receiver row equals the 1-D TL to 1e-8: the rows pass forms psi(zs) psi(z)
    through a reference row, the eigenvalue kernel forms psi(zs) psi(zr) directly, and the two
    adjugate forms agree to the eigenvector's conditioning at the converged k (residual over
    the gap to the neighbouring branch, up to ~6e-10 per mode here)."""
    from oracle import ess
    m, sol = canon_01
    print("case rows + field", flush=True)
    rows = ess.tl2d_rows(20000)
    stride = int(rows[1] - rows[0])
    R = m.rows(0, stride, rows.size)
    ref = sol.psi[rows]
    assert np.max(np.abs(np.abs(R) - np.abs(ref))) < 1e-7 * np.abs(ref).max()
    f = m.field2d(0, stride, 2, 1000.0, 500, precision=0)
    p, inc = m.tl1d(1000.0, 500)
    assert np.allclose(f[0, 0], p[0], rtol=1e-8, atol=1e-10 * np.abs(p[0]).max())


@pytest.mark.parametrize("f,az,zs", [(0.5, 45.0, 0.0), (0.25, 200.0, 20.0)])
def test_synthetic_vs_oracle(gpu_ctx, synth_profile, f, az, zs):
    """Jet profile (winds to ~70 m/s), 6000 rows, elevated source: counts, k, Im k, psi products."""
    import ncpamodes
    from oracle import wmod
    print("case synthetic f=%g az=%g zs=%g" % (f, az, zs), flush=True)
    pid, info = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=6000)
    m = gpu_ctx.solve_wide(g, [ncpamodes.problem(pid, f, az, zs * 1000, zs * 1000)])
    k, psr, ps, pr = m.problem_modes(0)
    sol = wmod.solve(synth_profile, freq=f, azimuth=az, Nz=6000, zs_km=zs, zr_km=zs, turnoff_wkb=True)
    assert k.size == sol.k.size > 10
    assert np.max(np.abs(np.sqrt((k ** 2).real) / sol.k - 1)) < 1e-10
    assert np.allclose(k.imag, sol.k_pert.imag, rtol=1e-6, atol=1e-16)
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(psr - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_batch_equals_single(gpu_ctx, synth_profile):
    """A batch of (f, az) problems returns per problem what single solves return."""
    import ncpamodes
    print("case batch", flush=True)
    pid, info = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=6000)
    probs = [ncpamodes.problem(pid, f, az) for f, az in [(0.2, 0.0), (0.4, 90.0), (0.3, 180.0), (0.1, 270.0)]]
    mb = gpu_ctx.solve_wide(g, probs)
    for i, pb in enumerate(probs):
        ms = gpu_ctx.solve_wide(g, [pb])
        kb = mb.problem_modes(i)[0]
        ks = ms.problem_modes(0)[0]
        assert kb.size == ks.size
        assert np.array_equal(kb, ks)


@pytest.mark.parametrize("opts", [
    {"wkb": True},                                   # ground-to-ground WKB cut-off (dz = H/Nz in the integral)
    {"wvnum_filter": (300.0, 360.0)},                # k_min = w/c_max, k_max = w/c_min
    {"lamb": True},                                  # Lamb-wave ground admittance
])
def test_options_vs_oracle(gpu_ctx, synth_profile, opts):
    """Interval options of WModeSolver::solve (WModeSolver.cpp:340-360, 572-620): the same modes
    as the oracle with the WKB cut-off, the wavenumber filter and the Lamb-wave boundary term."""
    import ncpamodes
    from oracle import atmos, ess, wmod
    f, az, Nz = 0.4, 120.0, 5000
    print("case options %s" % opts, flush=True)
    pid, info = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=Nz)
    nodes = atmos.AtmosphereNodes(atmos.read_profile(synth_profile))
    grid = wmod.WGrid(nodes, Nz, 150.0, 0.0)
    adm = ess.admittance(grid, True) if opts.get("lamb") else 0.0
    if opts.get("lamb"):
        assert abs(gpu_ctx.admittance(pid, g.z_min) - adm) <= 1e-12 * abs(adm)
    m = gpu_ctx.solve_wide(g, [ncpamodes.problem(pid, f, az, wkb=opts.get("wkb", False), admittance=adm,
                                                 wvnum_filter=opts.get("wvnum_filter"))])
    k, psr, ps, pr = m.problem_modes(0)
    sol = wmod.solve(nodes=nodes, grid=grid, freq=f, azimuth=az, turnoff_wkb=not opts.get("wkb", False),
                     lamb_wave_bc=bool(opts.get("lamb")), wvnum_filter=opts.get("wvnum_filter"))
    kmin, kmax = m.interval()
    assert kmin[0] == pytest.approx(sol.k_min, rel=1e-13) and kmax[0] == pytest.approx(sol.k_max, rel=1e-12)
    if opts.get("wkb"):
        assert sol.used_wkb and sol.k_max < 2 * np.pi * f / sol.ceffz.min()
    assert k.size == sol.k.size > 5
    assert np.max(np.abs(np.sqrt((k ** 2).real) / sol.k - 1)) < 1e-10
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(psr - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_multiprop_cli_vs_oracle(tmp_path, synth_profile):
    """WMod --multiprop: Nby2D_tloss_1d.wnm blocks per azimuth against the oracle's TL."""
    import subprocess
    from oracle import ess, wmod
    f, Nz = 0.2, 4000
    print("case WMod multiprop", flush=True)
    exe = os.path.join(os.path.dirname(G), "..", "ncpaprop-release_amd", "bin", "WMod")
    out = subprocess.run([exe, "--atmosfile", synth_profile, "--freq", str(f), "--multiprop", "--azimuth_start", "0",
                          "--azimuth_end", "180", "--azimuth_step", "90", "--Nz_grid", str(Nz), "--maxrange_km", "300",
                          "--Nrng_steps", "300", "--turnoff_WKB"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    blocks = [b for b in (tmp_path / "Nby2D_tloss_1d.wnm").read_text().split("\n\n") if b.strip()]
    assert len(blocks) == 3
    for az, blk in zip((0.0, 90.0, 180.0), blocks):
        a = np.array([l.split() for l in blk.strip().splitlines()], dtype=float)
        assert a.shape == (300, 5) and np.all(a[:, 1] == az)
        sol = wmod.solve(synth_profile, freq=f, azimuth=az, Nz=Nz, turnoff_wkb=True)
        _, po, _ = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], 1000.0, 300)
        tl, tlo = ess.tl_db(a[:, 2] + 1j * a[:, 3]), ess.tl_db(po)
        assert np.max(np.abs(tl - tlo)[tlo > -140]) <= 0.01


def test_supersonic_wind_rejected(gpu_ctx):
    """The counts need dT/dk < 0 on [k_min, k_max] (every branch decreasing): a pencil whose
    cd + 2 k md changes sign is refused with an argument error, nothing is computed."""
    import ncpamodes
    print("case supersonic", flush=True)
    N, dz, f, c = 500, 50.0, 0.5, 340.0
    om = 2 * math.pi * f
    kd = np.full(N, (om / c) ** 2)
    kd[0] += 1.0 / dz ** 2
    M = np.where(np.arange(N) > 300, -1.2, 0.2)     # a supersonic head-wind layer
    with pytest.raises(ncpamodes.NcpaError):
        ncpamodes.wmod_eigen(gpu_ctx, kd, M * M - 1.0, -2 * om * M / c, dz, 0.5 * om / c, 1.5 * om / c)


def test_full_size_residuals(gpu_ctx, synth_profile):
    """The benchmark's size (1 Hz, Nz = 20000, ~500 modes): every mode satisfies
    (D + k C + k^2 M) psi = 0 to 1e-9 relative (operator assembled on the host from the
    oracle's trace, EigenEngine.cpp:257-302) and the reference's normalisation
    sum psi^2 dz (1 + k^2) = 1 to 1e-9; eigenvalues strictly ascending inside [k_min, k_max]."""
    import ncpamodes
    from oracle import atmos, wmod
    f, az, Nz = 1.0, 45.0, 20000
    print("case full size f=%g az=%g" % (f, az), flush=True)
    pid, info = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=Nz)
    m = gpu_ctx.solve_wide(g, [ncpamodes.problem(pid, f, az)])
    k2 = (m.problem_modes(0)[0] ** 2).real
    k = np.sqrt(k2)
    psi = m.rows(0, 1, Nz)                      # [Nz, modes]
    grid = wmod.WGrid(atmos.AtmosphereNodes(atmos.read_profile(synth_profile)), Nz, 150.0, 0.0)
    kd, md, cd, _, _, kmin, kmax, _ = wmod.modal_trace(grid, f, az, 0.0, 0.0, True)
    dz = grid.dz_m
    assert k.size > 300 and np.all(np.diff(k) > 0) and k[0] >= kmin and k[-1] <= kmax
    assert np.allclose((psi * psi).sum(axis=0) * dz * (1 + k2), 1.0, rtol=1e-9, atol=0)
    # residual of the quadratic problem, rows scaled by dz^2: (A + k B + k^2 C) psi + off-diagonals
    A, B, C = -2.0 + dz * dz * kd, dz * dz * cd, dz * dz * md
    diag = A[:, None] + k[None, :] * (B[:, None] + k[None, :] * C[:, None])
    r = diag * psi
    r[:-1] += psi[1:]
    r[1:] += psi[:-1]
    scale = 4.0 * np.abs(psi).max(axis=0)
    res = float(np.max(np.abs(r).max(axis=0) / scale))
    print("modes %d, max relative residual %.2e" % (k.size, res), flush=True)
    assert res < 1e-9


@pytest.mark.parametrize("f,Nz,canon", [(0.02, 20000, True), (1.5, 6000, False)])
def test_frequency_extremes_vs_oracle(gpu_ctx, synth_profile, f, Nz, canon):
    """Low frequency (0.02 Hz, few modes, k ~ 4e-4) on the canonical profile and 1.5 Hz on the jet
    profile: counts, k to 1e-10, psi products to 1e-7 of max."""
    import ncpamodes
    from oracle import wmod
    path = CANON if canon else synth_profile
    print("case extremes f=%g Nz=%d" % (f, Nz), flush=True)
    pid, info = gpu_ctx.read_profile(path)
    g = gpu_ctx.grid(pid, Nz=Nz)
    m = gpu_ctx.solve_wide(g, [ncpamodes.problem(pid, f, 90.0)])
    k, psr, ps, pr = m.problem_modes(0)
    sol = wmod.solve(path, freq=f, azimuth=90.0, Nz=Nz, turnoff_wkb=True)
    assert k.size == sol.k.size > 2
    assert np.max(np.abs(np.sqrt((k ** 2).real) / sol.k - 1)) < 1e-10
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(psr - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_compact_geometry_matches(gpu_ctx, synth_profile):
    """Batches of more than 4096 work items run the compact staging (128-row tiles, 8-row
    reads); the arithmetic per lane is the same as with 256-row tiles, so a problem's modes
    from a 600-problem batch equal, bit for bit, those of its solve alone."""
    import ncpamodes
    print("case compact geometry", flush=True)
    pid, info = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    azs = np.linspace(0.0, 359.0, 600)
    mb = gpu_ctx.solve_wide(g, [ncpamodes.problem(pid, 1.0, a) for a in azs])
    assert int(np.sum((mb.counts + 63) // 64)) >= 4096
    kb, psrb, _, _ = mb.download()
    for i in (0, 311, 599):
        ms = gpu_ctx.solve_wide(g, [ncpamodes.problem(pid, 1.0, azs[i])])
        ks, psrs, _, _ = ms.download()
        s = slice(mb.offsets[i], mb.offsets[i + 1])
        assert np.array_equal(kb[s], ks) and np.array_equal(psrb[s], psrs)
        ms.free()
    mb.free()
