"""GPU: the launch variants of each contraction (block tile, LDS pipeline depth, row tiles per
XCD) and the contraction autotuner (ncpa_field_tune).

Within a precision every variant sums each output element with the same operations in the
same order, so each must return variant 0's field to the bit.  The batch has uneven mode
counts (three problems at 0.3, 0.7 and 1 Hz), a row count (300) that is not a multiple of the
128-row tiles and a column count (2 x 1100) that is not a multiple of 64 or 128, so the
padded edges of every tile shape are crossed.  Its top row tile (rows above ~117 km) holds
all-zero 128 x 32 chunks of Psi (modes above their evanescent cut), which the folded form
(k_gemm_fold, variants 0-2 of precisions 6 and 7) skips and the three-stage form (variants
3-4) does not: their equality checks that the skip changes no bit.  Variant 0 itself is
checked against the oracle by test_gpu_field2d.py and test_gpu_field_2hz.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NROWS, STRIDE, NR, DR = 300, 61, 1100, 500.0


@pytest.fixture(scope="module")
def batch(gpu_ctx, synth_profile, tmp_path_factory):
    import ncpamodes
    from ncpamodes import synth
    d = tmp_path_factory.mktemp("variants")
    paths = [synth_profile] + [synth.write_profile(str(d / ("P%d.dat" % s)), seed=s) for s in (3, 5)]
    pids = [gpu_ctx.read_profile(p)[0] for p in paths]
    g = gpu_ctx.grid(pids[0], Nz=20000)
    probs = [ncpamodes.problem(pids[0], 0.3, 90.0), ncpamodes.problem(pids[1], 1.0, 270.0),
             ncpamodes.problem(pids[2], 0.7, 45.0)]
    m = gpu_ctx.solve(g, probs)
    rows = m.rows_device(0, STRIDE, NROWS)
    yield m, rows
    rows.free()
    m.free()


def _bits(a):
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def test_batch_has_zero_chunks(batch):
    """The batch crosses the zero-chunk skip: some (problem, 128-row tile, 32-mode k-step) chunk
    of Psi is all zero (+0 or -0), and some chunk of the same row tile is not."""
    m, rows = batch
    psi = rows.numpy().reshape(NROWS, m.total)
    mp = -(-NROWS // 128) * 128
    zero = nonzero = 0
    off = 0
    for k in m.counts:
        k = int(k)
        kp = -(-k // 32) * 32
        nz = np.zeros((mp, kp), dtype=bool)
        nz[:NROWS, :k] = psi[:, off:off + k] != 0.0
        off += k
        chunk = nz.reshape(mp // 128, 128, kp // 32, 32).any(axis=(1, 3))
        zero += int((~chunk).sum())
        nonzero += int(chunk.sum())
    print("case zero chunks: %d zero, %d nonzero" % (zero, nonzero), flush=True)
    assert zero > 0 and nonzero > 0


@pytest.mark.parametrize("prec", range(8))
def test_variants_agree_to_the_bit(gpu_ctx, batch, prec):
    m, rows = batch
    nv = gpu_ctx.field_variants(prec)
    print("case contraction variants, precision %d: %d variants, modes %s" % (prec, nv, list(m.counts)), flush=True)
    assert nv >= 1
    outs = []
    try:
        for v in range(nv):
            gpu_ctx.set_field_variant(prec, v)
            assert gpu_ctx.field_variant(prec) == v
            o = m.field2d_device(rows, NROWS, DR, NR, prec)
            outs.append(o.numpy())
            o.free()
    finally:
        gpu_ctx.set_field_variant(prec, 0)
    assert np.count_nonzero(outs[0]) > 0.9 * outs[0].size
    for v in range(1, nv):
        assert np.array_equal(_bits(outs[v]), _bits(outs[0])), (prec, v)


def test_set_variant_rejects_out_of_range(gpu_ctx):
    import ncpamodes
    nv = gpu_ctx.field_variants(6)
    with pytest.raises(ncpamodes.NcpaError):
        gpu_ctx.set_field_variant(6, nv)
    with pytest.raises(ncpamodes.NcpaError):
        gpu_ctx.set_field_variant(9, 0)
    assert gpu_ctx.field_variant(6) == 0


@pytest.mark.parametrize("prec", [0, 6])
def test_tune_on_another_context(gpu_ctx, batch, prec):
    """ncpa_field_tune on a contractor context: it times every variant, keeps the fastest for
    that context only, and the field it leaves in the output is the variant-0 field."""
    import ncpamodes
    m, rows = batch
    print("case contraction autotune, precision %d" % prec, flush=True)
    ref = m.field2d_device(rows, NROWS, DR, NR, prec)
    other = ncpamodes.Context(0)
    try:
        st = m.field2d_stage(rows, NROWS, DR, NR, prec)
        ev = gpu_ctx.record()
        other.wait(ev)
        ev.free()
        out = other.empty(ref.n, ref.dtype)
        best, ms = st.tune(other, out, reps=2)
        print("precision %d: variant ms %s, best %d" % (prec, ["%.3f" % t for t in ms], best), flush=True)
        assert 0 <= best < other.field_variants(prec) and len(ms) == other.field_variants(prec)
        assert all(t > 0 for t in ms)
        # the fastest, or variant 0 when no other beats it by more than 2 %
        if best == 0:
            assert min(ms) >= 0.98 * ms[0]
        else:
            assert ms[best] == min(ms) and ms[best] < 0.98 * ms[0]
        assert other.field_variant(prec) == best and gpu_ctx.field_variant(prec) == 0
        got = st.contract(other)
        done = other.record()
        done.synchronize()
        gpu_ctx.wait(done)
        done.free()
        st.free()
        assert np.array_equal(_bits(got.numpy()), _bits(ref.numpy()))
        assert np.array_equal(_bits(out.numpy()), _bits(ref.numpy()))
        for x in (got, out):
            x.free()
    finally:
        ref.free()
        other.close()
