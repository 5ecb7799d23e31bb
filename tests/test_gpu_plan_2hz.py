"""GPU parity: a broadband plan at the configuration-2 shape, up to 2 Hz on the full grid.

64 frequencies uniformly spaced to 2 Hz x 2 azimuths at Nz = 20000: an uneven work list from a
handful of modes per problem to ~1100, one persistent launch under the plan's graph, with the
autotuner picking the eigenvalue form (multisection depth and the diagonal through LDS or by
scalar loads).  The plan's modes equal the eager path's on the same context to the bit (the
tuner sets the context's eager form and depth to its own), and the 2 Hz and lowest-frequency
problems match the oracle (k^2 to 1e-9 relative, Im k to 1e-4, psi(zs) psi(zr) to 1e-7 of its
largest magnitude).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NZ, NF, FMAX = 20000, 64, 2.0
AZS = (90.0, 270.0)


@pytest.fixture(scope="module")
def plan_2hz(synth_profile):
    import ncpamodes
    from ncpamodes import broadband as bb
    ctx = ncpamodes.Context(0)
    f = np.arange(NF + 1) * (FMAX / NF)
    pid, _ = ctx.read_profile(synth_profile)
    g = ctx.grid(pid, Nz=NZ)
    nfft = 4 * f.size
    S, _ = bb.source(ctx, "pulse1", f, nfft)
    probs = [ncpamodes.problem(pid, f[i], az) for az in AZS for i in range(1, f.size)]
    print("case plan 2 Hz: %d problems, Nz %d" % (len(probs), NZ), flush=True)
    plan = bb.BroadbandPlan(ctx, g, probs, len(AZS), f, S, np.array([200e3, 400e3]), 1.0, 320.0, nfft,
                            tune=True)
    plan.run()
    offs, kre, kim, psr = plan.modes()
    tuning = plan.tuning()
    m = ctx.solve(g, probs)
    # k_pert of the eager modes by the plan's own device function (ncpa_kpert): the host
    # download forms it with the host's complex square root, which may differ in the last bit
    offs_d, k2, ab, psr_d = (C.c_void_p() for _ in range(4))
    ctx.check(ctx.lib.ncpa_modes_device(m.h, C.byref(offs_d), C.byref(k2), C.byref(ab), C.byref(psr_d)))
    kre_d, kim_d = ctx.empty(m.total), ctx.empty(m.total)
    ctx.check(ctx.lib.ncpa_kpert(ctx.h, k2, ab, m.total, C.c_void_p(kre_d.ptr), C.c_void_p(kim_d.ptr)))
    ke = kre_d.numpy() + 1j * kim_d.numpy()
    _, pse, _, _ = m.download()
    counts = m.counts.copy()
    kre_d.free()
    kim_d.free()
    m.free()
    plan.close()
    ctx.close()
    yield f, offs, kre, kim, psr, tuning, ke, pse, counts


def test_plan_2hz_equals_eager(plan_2hz):
    f, offs, kre, kim, psr, tuning, ke, pse, counts = plan_2hz
    print("tuned: passes %d, form %d" % (tuning["ms_passes"], tuning["grid_blocks"]), flush=True)
    assert np.array_equal(np.diff(offs), counts)
    assert counts.min() >= 1 and counts.max() > 1000          # uneven: a few modes to ~1100
    assert np.array_equal(kre, ke.real) and np.array_equal(kim, ke.imag)
    assert np.array_equal(psr, pse)


@pytest.mark.parametrize("i,az_index", [(NF, 1), (1, 0)])
def test_plan_2hz_vs_oracle(plan_2hz, synth_profile, i, az_index):
    from oracle import ess
    f, offs, kre, kim, psr, _, _, _, _ = plan_2hz
    p = az_index * NF + (i - 1)
    s = slice(offs[p], offs[p + 1])
    print("case plan 2 Hz vs oracle: f %.4f az %g, %d modes" % (f[i], AZS[az_index], offs[p + 1] - offs[p]),
          flush=True)
    sol = ess.solve(synth_profile, freq=f[i], azimuth=AZS[az_index], Nz=NZ, turnoff_wkb=True)
    k = kre[s] + 1j * kim[s]
    assert k.size == sol.k.size
    assert np.allclose(k.real ** 2 - k.imag ** 2, sol.k ** 2, rtol=1e-9, atol=0)
    assert np.allclose(k.imag, sol.k_pert.imag, rtol=1e-4, atol=0)
    ref = sol.psi[sol.i_s] * sol.psi[sol.i_r]
    assert np.max(np.abs(psr[s] - ref)) <= 1e-7 * np.max(np.abs(ref))
