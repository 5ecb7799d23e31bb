"""The eigenvalue sweeps' two ways of reading the diagonal (LDS ring, scalar loads) give the
same modes to the bit: the per-row arithmetic is the same, only the source of a_i differs.
The eigenvector rows of the two forms (k_rows: the LDS record buffer, or the recorded rows
stored as the sweep passes them) are the same to the bit too.

Grids of 20000 rows (whole 16-row chunks), 7001 (9 top rows before the first chunk) and 100
(4 top rows, then six chunks), sources and receivers on the ground and aloft (the final sweep
records rows inside a chunk, the forward sweep climbs to them), frequencies from a few modes to
~500 per problem, several problems per launch; row sets with strides 1, 13 and 61, starting at
row 0 or inside the grid and reaching past the evanescent cut; the scalar form is also checked
against the oracle at 1 Hz.  Each test leaves the context in the default form (1).
"""
import numpy as np
import pytest

import ncpamodes
from ncpamodes import synth
from oracle import ess

pytestmark = pytest.mark.gpu


def solve_both(ctx, g, probs):
    out = []
    for form in (0, 1):
        ctx.set_modes_form(form)
        m = ctx.solve(g, probs)
        k, psr, ps, pr = m.download()
        out.append((m.counts.copy(), k, psr, ps, pr, m.iterations()))
        m.free()
    ctx.set_modes_form(1)
    return out


@pytest.mark.parametrize("Nz,freqs,zs_m,zr_m", [(20000, (0.05, 0.4, 1.0, 1.7), 0.0, 0.0),
                                                (20000, (0.3, 1.2), 10000.0, 1000.0),   # rows recorded aloft
                                                (7001, (0.1, 0.5), 2000.0, 0.0),
                                                (100, (0.3, 1.0), 0.0, 0.0)])
def test_forms_bit_equal(gpu_ctx, tmp_path, Nz, freqs, zs_m, zr_m):
    print("case modes form Nz=%d freqs=%s zs=%g zr=%g" % (Nz, freqs, zs_m, zr_m), flush=True)
    paths = [synth.write_profile(str(tmp_path / ("P%d.dat" % s)), seed=s) for s in (0, 2)]
    pids = [gpu_ctx.read_profile(p)[0] for p in paths]
    g = gpu_ctx.grid(pids[0], Nz=Nz)
    probs = [ncpamodes.problem(pid, f, az, zs_m=zs_m, zr_m=zr_m) for pid in pids for f in freqs
             for az in (30.0, 200.0)]
    a, b = solve_both(gpu_ctx, g, probs)
    assert np.array_equal(a[0], b[0])
    assert a[1].size > 0
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("Nz,freqs,zs_m,row0,stride,nrows", [(20000, (0.4, 1.0), 0.0, 0, 13, 1500),
                                                          (20000, (0.3, 1.2), 10000.0, 7, 61, 328),
                                                          (7001, (0.1, 0.5), 2000.0, 0, 1, 7001),
                                                          (100, (0.3, 1.0), 0.0, 3, 13, 8)])
def test_rows_forms_bit_equal(gpu_ctx, tmp_path, Nz, freqs, zs_m, row0, stride, nrows):
    print("case rows form Nz=%d freqs=%s zs=%g rows %d+%dk x %d" % (Nz, freqs, zs_m, row0, stride, nrows),
          flush=True)
    paths = [synth.write_profile(str(tmp_path / ("P%d.dat" % s)), seed=s) for s in (0, 2)]
    pids = [gpu_ctx.read_profile(p)[0] for p in paths]
    g = gpu_ctx.grid(pids[0], Nz=Nz)
    probs = [ncpamodes.problem(pid, f, az, zs_m=zs_m) for pid in pids for f in freqs for az in (30.0, 200.0)]
    out = []
    try:
        for form in (0, 1):
            gpu_ctx.set_modes_form(form)
            m = gpu_ctx.solve(g, probs)
            rows = m.rows_device(row0, stride, nrows)
            out.append(rows.numpy().copy())
            rows.free()
            m.free()
    finally:
        gpu_ctx.set_modes_form(1)
    assert out[0].size > 0 and np.count_nonzero(out[0]) > 0
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))


def test_scalar_form_vs_oracle(gpu_ctx, synth_profile):
    print("case modes form 1 vs oracle", flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=20000)
    gpu_ctx.set_modes_form(1)
    try:
        m = gpu_ctx.solve(g, [ncpamodes.problem(pid, 1.0, 90.0)])
        k, _, _, _ = m.problem_modes(0)
        m.free()
    finally:
        gpu_ctx.set_modes_form(1)
    sol = ess.solve(synth_profile, freq=1.0, azimuth=90.0, Nz=20000, turnoff_wkb=True, vectors=False)
    assert k.size == sol.k.size
    assert np.allclose(k.real ** 2 - k.imag ** 2, sol.k ** 2, rtol=1e-9, atol=0)
