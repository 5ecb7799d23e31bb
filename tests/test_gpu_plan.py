"""GPU parity: the broadband job plan (batched dispersion -> transfer functions -> waveforms as
one persistent launch under a replayed HIP graph) vs the eager C-ABI path and the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NFP, DF, NZ = 24, 0.02, 4000
RANGES = np.array([150e3, 230e3, 310e3])


def _problems(ncpamodes, pid, f, azs):
    return [ncpamodes.problem(pid, f[i], az) for az in azs for i in range(1, f.size)]


def _eager(ctx, grid, probs, n_sets, f, S, nfft):
    from ncpamodes import broadband as bb
    m = ctx.solve(grid, probs)
    W, kb = bb.synthesize_from_modes(ctx, m, n_sets, f, RANGES, np.ones(n_sets), S, 320.0, nfft)
    w = W.numpy().reshape(n_sets, RANGES.size, nfft)
    counts = m.counts.copy()
    for b in (W,) + kb:
        b.free()
    m.free()
    return w, counts


@pytest.fixture(scope="module")
def plan_case(gpu_ctx, synth_profile):
    import ncpamodes
    from ncpamodes import broadband as bb
    f = np.arange(NFP + 1) * DF
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=NZ)
    nfft = 4 * f.size
    S, _ = bb.source(gpu_ctx, "pulse1", f, nfft)
    azs = [90.0, 250.0]
    probs = _problems(ncpamodes, pid, f, azs)
    plan = bb.BroadbandPlan(gpu_ctx, g, probs, len(azs), f, S, RANGES, 1.0, 320.0, nfft, tune=True)
    yield plan, g, pid, f, S, nfft, azs
    plan.close()


def test_plan_replays_match_eager(gpu_ctx, plan_case):
    """Three replays, each after eager work and a synchronisation, all equal to the eager path."""
    import ncpamodes
    plan, g, pid, f, S, nfft, azs = plan_case
    print("case plan replays, tuning %s" % plan.tuning(), flush=True)
    probs = _problems(ncpamodes, pid, f, azs)
    w_e, counts = _eager(gpu_ctx, g, probs, len(azs), f, S, nfft)
    peak = np.max(np.abs(w_e))
    for rep in range(3):
        plan.run()
        w_p = plan.wave()
        offs, kre, kim, psr = plan.modes()
        assert np.array_equal(np.diff(offs), counts), rep
        err = np.max(np.abs(w_p - w_e)) / peak
        print("replay %d: %d modes, max |w_plan - w_eager| %.2e of peak" % (rep, kre.size, err), flush=True)
        assert err <= 1e-9
        _eager(gpu_ctx, g, probs, len(azs), f, S, nfft)   # eager work between replays
        gpu_ctx.synchronize()


def test_plan_vs_oracle(gpu_ctx, plan_case, synth_profile):
    """Plan waveforms vs the oracle's per-frequency solves -> compute_modal_sum -> FFT."""
    from oracle import broadband as obb
    from oracle import ess
    plan, g, pid, f, S, nfft, azs = plan_case
    print("case plan vs oracle", flush=True)
    plan.run()
    w = plan.wave()
    for s, az in enumerate(azs):
        modes = [np.zeros((0, 4))]
        for i in range(1, f.size):
            sol = ess.solve(synth_profile, freq=f[i], azimuth=az, Nz=NZ, turnoff_wkb=True)
            kp = sol.k_pert
            modes.append(np.stack([kp.real, kp.imag, sol.psi[sol.i_s], sol.psi[sol.i_r]], axis=1))
        d = {"f": f, "counts": np.array([q.shape[0] for q in modes]), "modes": modes, "rho_s": 1.0, "rho_r": 1.0}
        _, _, wo = obb.waveforms(d, RANGES, "pulse1", nfft=nfft, max_cel=320.0)
        err = np.max(np.abs(w[s] - wo)) / np.max(np.abs(wo))
        print("az %g: max waveform error %.2e of peak" % (az, err), flush=True)
        assert err <= 1e-5


def test_plan_set_problems(gpu_ctx, plan_case):
    """New problem records (other azimuths) replayed through the same graph."""
    import ncpamodes
    plan, g, pid, f, S, nfft, azs = plan_case
    print("case plan set_problems", flush=True)
    azs2 = [0.0, 135.0]
    probs2 = _problems(ncpamodes, pid, f, azs2)
    plan.set_problems(probs2)
    plan.run()
    w_p = plan.wave()
    w_e, _ = _eager(gpu_ctx, g, probs2, len(azs2), f, S, nfft)
    assert np.max(np.abs(w_p - w_e)) <= 1e-9 * np.max(np.abs(w_e))
    plan.set_problems(_problems(ncpamodes, pid, f, azs))


def test_plan_capacity_overflow(gpu_ctx, synth_profile):
    """A batch with more modes than the plan was sized for is reported, not computed."""
    import ncpamodes
    from ncpamodes import broadband as bb
    from ncpamodes._lib import NcpaError, NCPA_ERR_MAX_MODES
    print("case plan overflow", flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid, Nz=NZ)
    f = np.arange(5) * 0.01
    S, _ = bb.source(gpu_ctx, "impulse", f, 4 * f.size)
    low = _problems(ncpamodes, pid, f, [90.0])
    m = gpu_ctx.solve(g, low)
    cap = m.total
    m.free()
    plan = bb.BroadbandPlan(gpu_ctx, g, low, 1, f, S, RANGES, 1.0, 320.0, 4 * f.size, max_modes_total=cap,
                            tune=False)
    plan.run()
    assert plan.result()[1] == cap
    plan.set_problems([ncpamodes.problem(pid, 10 * fi, 90.0) for fi in f[1:]])   # ~10x the modes
    plan.run()
    with pytest.raises(NcpaError) as ei:
        plan.result()
    assert ei.value.code == NCPA_ERR_MAX_MODES
    plan.close()
