"""GPU parity at the edges of the batched solvers: the 1 Hz ESS problem on the full 20000-row
grid against LAPACK, an empty batch, and a problem past MAX_MODES (ESSModeSolver.cpp:17,
EigenEngine.cpp:155 stores into k2[MAX_MODES]) rejected without harming the context."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_modes_1hz_full_grid_vs_oracle(gpu_ctx, synth_profile):
    """The configuration-1 problem size (1 Hz, Nz 20000, ~500 modes): k^2 to 1e-9 relative
    (plus the FP64 floor 64 u ||T||), the perturbed Im k and psi(zs) psi(zr) vs the oracle."""
    import ncpamodes
    from oracle import ess
    f, az = 1.0, 90.0
    print("case 1 Hz full grid az=%g" % az, flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    assert g.Nz == 20000
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, f, az)])
    k, psr, _, _ = m.problem_modes(0)
    sol = ess.solve(synth_profile, freq=f, azimuth=az, turnoff_wkb=True)
    assert k.size == sol.k.size and k.size > 400, (k.size, sol.k.size)
    k2 = (k ** 2).real
    k2o = sol.k ** 2
    normT = 4.0 / g.dz ** 2
    assert np.all(np.abs(k2 - k2o) <= 1e-9 * k2o + 64 * 2.2e-16 * normT)
    assert np.allclose(k.imag, sol.k_pert.imag, rtol=1e-4, atol=1e-14)
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(psr - ref)) <= 1e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("wide", [False, True])
def test_empty_batch_rejected(gpu_ctx, synth_profile, wide):
    from ncpamodes._lib import NcpaError, NCPA_ERR_ARG
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=4000)
    with pytest.raises(NcpaError) as ei:
        (gpu_ctx.solve_wide if wide else gpu_ctx.solve)(g, [])
    assert ei.value.code == NCPA_ERR_ARG


@pytest.mark.parametrize("wide", [False, True])
def test_max_modes_rejected_context_usable(gpu_ctx, synth_profile, wide):
    """At 12 Hz the interval holds more than MAX_MODES = 4000 eigenvalues: the solve fails with
    NCPA_ERR_MAX_MODES before any eigenvalue kernel runs, and the same context then solves a
    normal batch to the same modes as before the failure."""
    import ncpamodes
    from ncpamodes._lib import NcpaError, NCPA_ERR_MAX_MODES
    solve = gpu_ctx.solve_wide if wide else gpu_ctx.solve
    print("case max modes wide=%s" % wide, flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    ok = [ncpamodes.problem(pid, 0.2, 90.0)]
    before = solve(g, ok).download()
    with pytest.raises(NcpaError) as ei:
        solve(g, [ncpamodes.problem(pid, 0.2, 90.0), ncpamodes.problem(pid, 12.0, 90.0)])
    assert ei.value.code == NCPA_ERR_MAX_MODES
    after = solve(g, ok).download()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("Nz", [100, 300])
def test_small_grids_vs_oracle(gpu_ctx, synth_profile, wide, Nz):
    """Grids smaller than one 256-row LDS tile (100 rows) and not a multiple of the 16-row
    chunk (300 rows): every sweep starts in a partial tile.  Counts, k, Im k and psi(zs)psi(zr)
    against the oracle on the same grid (the matrix is coarse, the parity is exact)."""
    import ncpamodes
    from oracle import ess, wmod
    print("case small grid Nz=%d wide=%s" % (Nz, wide), flush=True)
    f, az = 0.1, 90.0
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=Nz)
    m = (gpu_ctx.solve_wide if wide else gpu_ctx.solve)(g, [ncpamodes.problem(pid, f, az)])
    k, psr, _, _ = m.problem_modes(0)
    sol = (wmod if wide else ess).solve(synth_profile, freq=f, azimuth=az, Nz=Nz, turnoff_wkb=True)
    assert k.size == sol.k.size > 0, (k.size, sol.k.size)
    assert np.max(np.abs(np.sqrt((k ** 2).real) / sol.k - 1)) < 1e-10
    assert np.allclose(k.imag, sol.k_pert.imag, rtol=1e-6, atol=1e-16)
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(psr - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_eigenvalues_to_working_precision(gpu_ctx, synth_profile, tmp_path):
    """Every eigenvalue of an 8-problem 1 Hz batch (Nz 20000, ~4000 modes) within 8 tolerances
    (8 x 8 eps ||dz^2 T||) of LAPACK's bisection, in dz^2 k^2: the early Newton acceptance must
    not stop a lane short of working precision (accepting the predicted next step at the
    tolerance itself left one of these eigenvalues 300 tolerances off)."""
    import ncpamodes
    from ncpamodes import synth
    from oracle import ess
    from oracle.atmos import read_profile
    print("case eigenvalues at working precision, 1 Hz x 8", flush=True)
    paths = [synth_profile, synth.write_profile(str(tmp_path / "P3.dat"), seed=3)]
    pids = [gpu_ctx.read_profile(p)[0] for p in paths]
    g = gpu_ctx.grid(pids[0])
    azs = [0.0, 90.0, 180.0, 270.0]
    probs = [ncpamodes.problem(pid, 1.0, az) for pid in pids for az in azs]
    m = gpu_ctx.solve(g, probs)
    worst = 0.0
    for i, (path, az) in enumerate([(p, az) for p in paths for az in azs]):
        grid = ess.EssGrid(ess.AtmosphereNodes(read_profile(path)), g.Nz, 150.0, 0.0)
        grid.set_attenuation(1.0, None)
        diag, _, kmin, kmax, _ = ess.modal_trace(grid, 1.0, az, 0.0, 0.0, True, 0.0)
        w, _ = ess.eigen(diag, grid.dz, kmin, kmax, vectors=False)
        k, _, _, _ = m.problem_modes(i)
        assert k.size == w.size
        h2 = grid.dz ** 2
        tol = 8 * np.finfo(float).eps * (4.0 + h2 * max(abs(kmin ** 2), abs(kmax ** 2)))
        err = np.max(np.abs((k ** 2).real - w)) * h2
        worst = max(worst, err / tol)
    m.free()
    print("largest error %.2f tolerances" % worst, flush=True)
    assert worst <= 8.0
