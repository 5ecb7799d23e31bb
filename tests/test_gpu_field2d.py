"""GPU parity: 2-D field (eigenvector rows + MFMA contraction) in FP64, FP32, split BF16 (3 and 6
products) and split FP16 (3 and 4 products), plain and FP64-folded, vs the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def field_case(gpu_ctx, synth_profile):
    import ncpamodes
    from oracle import ess
    f, az = 0.5, 90.0
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, f, az)])
    stride = g.Nz // 500                      # the reference's 2-D row set (ModeSolver.cpp:513)
    nrows = len(range(0, g.Nz, stride))
    nr, dr = 2000, 500.0
    sol = ess.solve(synth_profile, freq=f, azimuth=az, turnoff_wkb=True)
    rows = np.arange(nrows) * stride
    r, ref = ess.tl2d(sol.k_pert, sol.psi[rows], sol.psi[sol.i_s], dr, nr)
    return m, stride, nrows, nr, dr, ref.T   # ref [rows, ranges]


@pytest.mark.parametrize("prec,tol_db", [(0, 1e-2), (1, 0.1), (2, 0.1), (3, 0.1), (4, 0.1), (5, 0.1), (6, 0.1),
                                         (7, 0.1)])
def test_field2d(field_case, prec, tol_db):
    m, stride, nrows, nr, dr, ref = field_case
    print("case field2d precision=%d" % prec, flush=True)
    F = m.field2d(0, stride, nrows, dr, nr, precision=prec)[0]
    from oracle import ess
    tl, tlo = ess.tl_db(F), ess.tl_db(ref)
    mask = tlo > -140.0
    err = np.abs(tl - tlo)[mask]
    print("precision %d: max |dTL| %.3e dB, rms %.3e dB over %d points" % (prec, err.max(), np.sqrt((err ** 2).mean()), mask.sum()), flush=True)
    assert err.max() <= tol_db
    if prec == 0:
        assert np.max(np.abs(F - ref)) <= 1e-6 * np.max(np.abs(ref))
