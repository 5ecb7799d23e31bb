"""GPU parity of the full-size 2-D field: 1500 depths (every 13th row of Nz = 20000) x 10000
ranges (0.1 km steps to 1000 km) on profile P0, at 2 Hz (azimuth 270, the most modes of the
eight azimuths: 1094) and at 1 Hz (azimuth 90, 443 modes).

* The FP64 contraction (getTLoss2D's sum, ModeSolver.cpp:492-565) against the oracle on a
  sampled block (every 10th depth x every 7th range), to 1e-6 of the block's largest |p|.
* At 2 Hz, the reduced precisions against the FP64 contraction at every point with
  TL > -140 dB, computed on the device (ncpa_field_compare): FP32 MFMA, split-BF16 with three
  products (hi*hi + hi*lo + lo*hi, FP32 accumulation) and split-BF16 with six products folded
  into FP64 every 32 modes; max and RMS |dTL| asserted.
* The 1 Hz 1-D TL at 10000 ranges against the oracle (getTLoss1D, ModeSolver.cpp:292-381).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NR, ROWS, STRIDE, DR = 10000, 1500, 13, 100.0
ROW_SAMPLE, RANGE_SAMPLE = 10, 7


def _field_case(gpu_ctx, path, f, az):
    import ncpamodes
    from oracle import ess
    pid, _ = gpu_ctx.read_profile(path)
    g = gpu_ctx.grid(pid, Nz=20000)
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, f, az)])
    rows = m.rows_device(0, STRIDE, ROWS)
    o64 = m.field2d_device(rows, ROWS, DR, NR, 0)
    sol = ess.solve(path, freq=f, azimuth=az, Nz=20000, turnoff_wkb=True)
    return m, rows, o64, sol


@pytest.fixture(scope="module")
def field_2hz(gpu_ctx, synth_profile):
    print("case 2 Hz field: P0 az 270, %d x %d" % (ROWS, NR), flush=True)
    m, rows, o64, sol = _field_case(gpu_ctx, synth_profile, 2.0, 270.0)
    yield m, rows, o64, sol
    o64.free()
    rows.free()
    m.free()


@pytest.fixture(scope="module")
def field_1hz(gpu_ctx, synth_profile):
    print("case 1 Hz field: P0 az 90, %d x %d" % (ROWS, NR), flush=True)
    m, rows, o64, sol = _field_case(gpu_ctx, synth_profile, 1.0, 90.0)
    yield m, rows, o64, sol
    o64.free()
    rows.free()
    m.free()


def _check_fp64_vs_oracle(case):
    from oracle import ess
    m, rows, o64, sol = case
    assert m.counts[0] == sol.k.size
    F = o64.numpy().reshape(ROWS, NR, 2)
    jr = np.arange(0, ROWS, ROW_SAMPLE)
    ir = np.arange(0, NR, RANGE_SAMPLE)
    got = F[jr][:, ir, 0] + 1j * F[jr][:, ir, 1]                       # [rows, ranges]
    ref = ess.tl2d_at(sol.k_pert, sol.psi[jr * STRIDE], sol.psi[sol.i_s], (ir + 1) * DR).T
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    tl, tlo = ess.tl_db(got), ess.tl_db(ref)
    mask = tlo > -140.0
    dtl = np.max(np.abs(tl - tlo)[mask])
    print("%d modes: FP64 field vs oracle on %d x %d: max |dp| %.2e of max |p|, max |dTL| %.2e dB over %d points"
          % (sol.k.size, jr.size, ir.size, err, dtl, mask.sum()), flush=True)
    assert err <= 1e-6
    assert dtl <= 1e-2


def test_fp64_field_2hz_vs_oracle(field_2hz):
    assert field_2hz[3].k.size > 1000
    _check_fp64_vs_oracle(field_2hz)


def test_fp64_field_1hz_vs_oracle(field_1hz):
    _check_fp64_vs_oracle(field_1hz)


# (precision, max dB, rms dB).  BF16x6F (six products of 3-part BF16 splits, FP32 sums over 32
# modes folded into FP64) meets 0.1 dB everywhere.  FP32 MFMA and the 3-term split BF16 carry
# ~24 and ~16 bits through running FP32 sums over all 1094 modes and lose the deepest
# interference nulls, where the field is ~1e-4 of the terms summed; their bounds are regression
# limits (about 1.3x the errors measured on this field: FP32 0.099 dB max / 1.2e-3 dB rms,
# BF16x3 9.6 dB / 0.12 dB; BF16x6F 0.030 dB / 3.8e-4 dB), and BF16x3 does not meet 0.1 dB.
@pytest.mark.parametrize("prec,max_db,rms_db", [(6, 0.1, 2e-3), (1, 0.13, 5e-3), (2, 12.5, 0.16)])
def test_reduced_precision_2hz(gpu_ctx, field_2hz, prec, max_db, rms_db):
    m, rows, o64, sol = field_2hz
    oc = m.field2d_device(rows, ROWS, DR, NR, prec)
    mx, rms, cnt = gpu_ctx.field_compare(o64, oc, prec, ROWS * NR)
    oc.free()
    print("2 Hz precision %d: max |dTL| %.4f dB, rms %.3e dB over %d of %d points"
          % (prec, mx, rms, cnt, ROWS * NR), flush=True)
    # at 2 Hz most of the 150 km x 1000 km field lies below -140 dB (absorption above the
    # stratospheric duct and beyond a few hundred km): ~2.9 % of the points count
    assert cnt > 0.02 * ROWS * NR
    assert mx <= max_db
    assert rms <= rms_db


def test_tl1d_1hz_vs_oracle(field_1hz):
    from oracle import ess
    m, rows, o64, sol = field_1hz
    p, inc = m.tl1d(DR, NR)
    r, po, inco = ess.tl1d(sol.k_pert, sol.psi[sol.i_s], sol.psi[sol.i_r], DR, NR)
    tl, tlo = ess.tl_db(p[0]), ess.tl_db(po)
    mask = tlo > -140.0
    err = np.max(np.abs(tl - tlo)[mask])
    print("1 Hz 1-D TL: max |dTL| %.2e dB over %d ranges" % (err, mask.sum()), flush=True)
    assert err <= 0.1
    assert np.max(np.abs(inc[0] / inco - 1)) <= 1e-6
