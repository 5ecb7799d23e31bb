"""Grid samples of k_setup_grid (c_eff on the rows, and through the eigenvalues the matrix
diagonal) against the oracle's Steffen splines (ESSModeSolver.cpp:225-239, 544-559).

Two paths in the kernel: a block's node window staged in LDS (nodes coarser than ~3 grid rows,
the usual case), and global reads when the window does not fit (much finer nodes); grids whose
row count is not a multiple of the 1024-row block, and the last-row quirk of the matrix
(EigenEngine.cpp:57-71: row Nz-1 takes diag[Nz-2]) through the eigenvalues.
"""
import numpy as np
import pytest

import ncpamodes
from ncpamodes import synth
from oracle import ess
from oracle.atmos import AtmosphereNodes, read_profile

pytestmark = pytest.mark.gpu


def fine_profile(path, dz_km):
    z, u, v, w, T, rho, P = synth.profile_columns(seed=5, dz_km=dz_km)
    with open(path, "w") as fh:
        fh.write(synth.HEADER)
        for row in zip(z, u, v, w, T, rho, P):
            fh.write(" ".join("%.10e" % x for x in row) + "\n")
    return path


@pytest.mark.parametrize("dz_km,Nz,freq,az", [(0.1, 20000, 1.0, 90.0),    # LDS window
                                             (0.1, 7001, 0.5, 225.0),     # partial last block
                                             (0.005, 3000, 0.3, 45.0)])   # window too wide: global
def test_grid_samples_vs_oracle(gpu_ctx, tmp_path, dz_km, Nz, freq, az):
    print("case grid samples dz_km=%g Nz=%d f=%g az=%g" % (dz_km, Nz, freq, az), flush=True)
    path = fine_profile(str(tmp_path / "fine.dat"), dz_km)
    pid, _ = gpu_ctx.read_profile(path)
    g = gpu_ctx.grid(pid, Nz=Nz)
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, freq, az)])
    ce = m.ceff(0)
    nodes = AtmosphereNodes(read_profile(path))
    grid = ess.EssGrid(nodes, Nz=Nz)
    ce_o = grid.ceff(az)
    assert np.max(np.abs(ce - ce_o) / ce_o) < 1e-13
    # the eigenvalues see every row of the diagonal, the last one included
    k, _, _, _ = m.problem_modes(0)
    sol = ess.solve(path, freq=freq, azimuth=az, Nz=Nz, turnoff_wkb=True, vectors=False)
    assert k.size == sol.k.size
    assert np.allclose(k.real ** 2 - k.imag ** 2, sol.k ** 2, rtol=1e-9, atol=0)
    m.free()
