"""GPU: the device-side TL comparison (ncpa_field_compare) against the same statistic on the
host, and the reduced-precision contractions against the FP64 contraction at the 1 Hz field
shape (1500 depths x 10000 ranges, Nz = 20000) over 64 solves (8 profiles x 8 azimuths).  The
FP64 contraction itself is pinned to the oracle in test_gpu_field2d.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NR, ROWS, STRIDE, DR = 10000, 1500, 13, 100.0


def _host_db(a, b, floor=-140.0):
    p0 = a[..., 0] ** 2 + a[..., 1] ** 2
    p1 = np.maximum(b[..., 0] ** 2 + b[..., 1] ** 2, 1e-300)
    mask = 10.0 * np.log10(np.maximum(p0, 1e-300)) > floor
    d = np.abs(10.0 * np.log10(p1[mask] / p0[mask]))
    return d.max(), np.sqrt(np.mean(d ** 2)), int(mask.sum())


def test_field_compare_matches_host(gpu_ctx, synth_profile):
    import ncpamodes
    print("case field_compare vs host", flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    m = gpu_ctx.solve(g, [ncpamodes.problem(pid, 0.5, 90.0), ncpamodes.problem(pid, 0.5, 270.0)])
    nrows, nr = 300, 3000
    rows = m.rows_device(0, g.Nz // 500, nrows)
    o64 = m.field2d_device(rows, nrows, 300.0, nr, 0)
    a = o64.numpy().reshape(-1, 2)
    n = 2 * nrows * nr
    for prec in (0, 1, 4):
        oc = m.field2d_device(rows, nrows, 300.0, nr, prec)
        b = oc.numpy().astype(np.float64).reshape(-1, 2)
        mx, rms, cnt = gpu_ctx.field_compare(o64, oc, prec, n)
        hmx, hrms, hcnt = _host_db(a, b)
        print("precision %d: device %.4e / %.4e dB over %d, host %.4e / %.4e over %d"
              % (prec, mx, rms, cnt, hmx, hrms, hcnt), flush=True)
        assert cnt == hcnt
        assert mx == pytest.approx(hmx, rel=1e-9, abs=1e-12)
        assert rms == pytest.approx(hrms, rel=1e-6, abs=1e-12)
        oc.free()
    # a planted 1 dB error at one point above the floor is found exactly
    i = int(np.argmax(a[:, 0] ** 2 + a[:, 1] ** 2))
    c = a.copy()
    c[i] *= 10.0 ** (1.0 / 20.0)
    oc = gpu_ctx.empty(c.size)
    oc.upload(c.ravel())
    mx, _, _ = gpu_ctx.field_compare(o64, oc, 0, n)
    assert mx == pytest.approx(1.0, abs=1e-9)
    mx, rms, cnt = gpu_ctx.field_compare(o64, oc, 0, 0)
    assert (mx, rms, cnt) == (0.0, 0.0, 0)
    for x in (oc, o64, rows):
        x.free()
    m.free()


@pytest.fixture(scope="module")
def field_1hz(gpu_ctx, tmp_path_factory):
    import ncpamodes
    from ncpamodes import synth
    d = tmp_path_factory.mktemp("gate")
    pids = [gpu_ctx.read_profile(synth.write_profile(str(d / ("P%d.dat" % s)), seed=s))[0] for s in range(8)]
    g = gpu_ctx.grid(pids[0], Nz=20000)
    probs = [ncpamodes.problem(pid, 1.0, az) for pid in pids for az in range(0, 360, 45)]
    m = gpu_ctx.solve(g, probs)
    rows = m.rows_device(0, STRIDE, ROWS)
    o64 = m.field2d_device(rows, ROWS, DR, NR, 0)
    yield m, rows, o64, len(probs)
    for x in (o64, rows):
        x.free()
    m.free()


# Split BF16 with 6 products and FP64 folding (24-bit operands, FP32 sums over 32 modes only):
# within 0.1 dB at every point.  The other forms miss 0.1 dB at a few deep interference nulls
# of some fields, where the result is ~1e-4 of the terms summed: FP32 running sums over ~700
# modes (FP32, split BF16 x6), or 22-bit FP16 operand splits (FP16 x3 / x4, folded or not).
# Their bounds here are regression limits, about 1.3x the errors measured on these 64 solves.
@pytest.mark.parametrize("prec,tol_db", [(6, 0.1), (3, 0.17), (1, 0.2), (5, 0.25), (7, 0.27), (4, 0.33)])
def test_precision_1hz_all_solves(gpu_ctx, field_1hz, prec, tol_db):
    m, rows, o64, P = field_1hz
    print("case 1 Hz gate precision=%d over %d solves" % (prec, P), flush=True)
    oc = m.field2d_device(rows, ROWS, DR, NR, prec)
    mx, rms, cnt = gpu_ctx.field_compare(o64, oc, prec, P * ROWS * NR)
    oc.free()
    print("precision %d: max |dTL| %.4f dB, rms %.2e dB over %d points" % (prec, mx, rms, cnt), flush=True)
    assert cnt > 0.4 * P * ROWS * NR
    assert mx <= tol_db
