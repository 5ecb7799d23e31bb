"""GPU: stream ordering between contexts (ncpa_event_*) and the split field API
(ncpa_field2d_stage / _contract).  A modes handle solved on one context is contracted on
another after an event, and a two-stage pipeline (eigenpairs, rows and staging of step i on
one stream, the field contraction of step i-1 on the other, buffers reused two steps later)
gives the same fields as running every step serially on one context.  A batch whose last
problem has no modes gives a zero field for it in every precision."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROWS, NR, DR, PREC = 300, 2000, 500.0, 6


def _probs(ncpamodes, pid, azs, f=0.5):
    return [ncpamodes.problem(pid, f, az) for az in azs]


def test_field_on_second_context(gpu_ctx, synth_profile):
    import ncpamodes
    print("case field on a second context", flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    other = ncpamodes.Context(0)
    m = gpu_ctx.solve(g, _probs(ncpamodes, pid, [30.0, 200.0]))
    rows = m.rows_device(0, g.Nz // 500, ROWS)
    ev = gpu_ctx.record()
    other.wait(ev)
    ev.free()
    f_other = m.field2d_device(rows, ROWS, DR, NR, PREC, ctx=other)
    f_own = m.field2d_device(rows, ROWS, DR, NR, PREC)
    other.synchronize()
    a, b = f_other.numpy(), f_own.numpy()
    assert np.array_equal(a, b)       # same kernels, same inputs: bit-identical
    for x in (f_other, f_own, rows):
        x.free()
    m.free()
    other.close()


def test_pipelined_steps_match_serial(gpu_ctx, synth_profile):
    import ncpamodes
    print("case pipelined steps vs serial", flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    stride = g.Nz // 500
    sets = [_probs(ncpamodes, pid, azs) for azs in ([0.0, 90.0], [45.0, 300.0], [180.0, 270.0],
                                                      [10.0, 120.0], [60.0, 240.0])]
    serial = []
    for pr in sets:
        m = gpu_ctx.solve(g, pr)
        serial.append(m.field2d(0, stride, ROWS, DR, NR, precision=PREC))
        m.free()
    contractor = ncpamodes.Context(0)
    depth = 2
    rows = [None] * (depth + 1)
    inflight, got = [], []
    for i, pr in enumerate(sets):
        m = gpu_ctx.solve(g, pr)
        slot = i % (depth + 1)
        if rows[slot] is None:
            rows[slot] = gpu_ctx.empty(ROWS * 4000)
        gpu_ctx.check(gpu_ctx.lib.ncpa_modes_rows(gpu_ctx.h, m.h, 0, stride, ROWS,
                                                  ctypes.c_void_p(rows[slot].ptr)))
        # odd steps: staging on the solver, contraction on the other context; even steps:
        # the whole field on the other context
        if i % 2:
            st = m.field2d_stage(rows[slot], ROWS, DR, NR, PREC)
        ready = gpu_ctx.record()
        contractor.wait(ready)
        ready.free()
        if i % 2:
            out = st.contract(contractor)
        else:
            st = None
            out = m.field2d_device(rows[slot], ROWS, DR, NR, PREC, ctx=contractor)
        done = contractor.record()
        inflight.append((m, st, done, out))
        while len(inflight) > (depth if i + 1 < len(sets) else 0):
            m0, st0, ev0, out0 = inflight.pop(0)
            ev0.synchronize()
            got.append(out0.numpy().reshape(2, ROWS, NR, 2))
            out0.free()
            gpu_ctx.wait(ev0)
            if st0 is not None:
                st0.free()
            m0.free()
            ev0.free()
    for r in rows:
        if r is not None:
            r.free()
    contractor.close()
    for i, (F, G) in enumerate(zip(serial, got)):
        Gc = G[..., 0] + 1j * G[..., 1]
        print("step %d: max |pipelined - serial| %.3e" % (i, np.max(np.abs(Gc - F))), flush=True)
        assert np.array_equal(Gc, F)


@pytest.mark.parametrize("prec", [0, 1, 4, 6])
def test_problem_without_modes_last(gpu_ctx, synth_profile, prec):
    """The last problem of the batch has no modes (phase-speed window above every mode): its
    field is zero and the other problem's field equals its single-problem result."""
    import ncpamodes
    print("case last problem without modes, precision %d" % prec, flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    stride = g.Nz // 500
    full = ncpamodes.problem(pid, 0.5, 90.0)
    empty = ncpamodes.problem(pid, 0.5, 90.0, wvnum_filter=(1000.0, 1001.0))
    m = gpu_ctx.solve(g, [full, empty])
    assert m.counts[1] == 0 and m.counts[0] > 0
    F = m.field2d(0, stride, ROWS, DR, NR, precision=prec)
    m.free()
    m1 = gpu_ctx.solve(g, [full])
    F1 = m1.field2d(0, stride, ROWS, DR, NR, precision=prec)
    m1.free()
    assert np.all(F[1] == 0)
    assert np.array_equal(F[0], F1[0])


@pytest.mark.parametrize("prec", [0, 1, 4, 5, 6, 7])
def test_stage_contract_equals_field2d(gpu_ctx, synth_profile, prec):
    """ncpa_field2d_stage on one context + ncpa_field2d_contract on another after an event gives
    the same field, bit for bit, as the one-call ncpa_field2d (every precision family: FP64 and
    FP32 staging, the scaled FP16 splits, the folded forms)."""
    import ncpamodes
    print("case stage/contract precision %d" % prec, flush=True)
    pid, _ = gpu_ctx.read_profile(synth_profile)
    g = gpu_ctx.grid(pid)
    m = gpu_ctx.solve(g, _probs(ncpamodes, pid, [70.0, 160.0]))
    rows = m.rows_device(0, g.Nz // 500, ROWS)
    ref = m.field2d_device(rows, ROWS, DR, NR, prec)
    other = ncpamodes.Context(0)
    st = m.field2d_stage(rows, ROWS, DR, NR, prec)
    ev = gpu_ctx.record()
    other.wait(ev)
    ev.free()
    got = st.contract(other)
    done = other.record()
    done.synchronize()
    gpu_ctx.wait(done)
    done.free()
    st.free()
    assert np.array_equal(got.numpy(), ref.numpy())
    for x in (got, ref, rows):
        x.free()
    m.free()
    other.close()
