"""GPU: the serving loop (continuous batching over several device contexts, admission by the
device-memory footprint of each batch) at the production precision (BF16x6F).  Each result
must be the same alone and in the mix, with batches overlapping on the device."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PROD = 6   # NCPA_PREC_BF16X6F, the server's default


def _tl_db(p):
    return 20 * np.log10(np.maximum(np.abs(p), 1e-300))


def _results_equal(a, b):
    if isinstance(a, tuple):
        return all(np.array_equal(x, y) for x, y in zip(a, b))
    return np.array_equal(a, b)


def test_footprint_matches_allocator():
    """batch_footprint (from the allocation sites) against the allocator's measured peak for
    1-D, 2-D (BF16x6F and FP64 contraction) and mixed batches: never below it, within 5 %."""
    from ncpamodes import serve
    print("case serve footprint vs measured", flush=True)
    srv = serve.Server(n_workers=1, precision=PROD, max_batch=64)
    try:
        w = srv.workers[0]
        cases = [("1d", 8, 0.0, PROD), ("2d", 6, 1.0, PROD), ("2d fp64", 4, 1.0, 0), ("mix", 10, 0.4, PROD)]
        for name, n, frac, prec in cases:
            reqs = serve.make_requests(n, 100.0, seed=21 + n, n_profiles=8, frac_2d=frac)
            srv.count(reqs)
            assert all(r.nm and r.nm > 0 for r in reqs)
            live0, _, _ = w.ctx.memory(reset_peak=True)
            w.run_batch(reqs, precision=prec)
            live1, peak, _ = w.ctx.memory()
            used = peak - live0
            model = serve.batch_footprint(reqs, srv.Nz, prec)
            print("%s: %d requests, %d modes: measured peak %.2f MB, model %.2f MB"
                  % (name, n, sum(r.nm for r in reqs), used / 1e6, model / 1e6), flush=True)
            assert live1 == live0            # everything the batch allocated was returned
            assert used <= model <= 1.05 * used + (1 << 20)
    finally:
        srv.close()


def test_count_matches_solve():
    """ncpa_ess_count gives the mode counts the solve returns, without the modes."""
    import ncpamodes
    from ncpamodes import serve
    print("case ess_count", flush=True)
    srv = serve.Server(n_workers=1, precision=PROD)
    try:
        reqs = serve.make_requests(12, 100.0, seed=5, n_profiles=6)
        srv.count(reqs)
        w = srv.workers[0]
        for r in reqs:
            pid, g, _ = w.cache.get(r.seed)
            m = w.ctx.solve(g, [ncpamodes.problem(pid, r.freq, r.azimuth)])
            assert m.counts[0] == r.nm, (r.rid, m.counts[0], r.nm)
            m.free()
    finally:
        srv.close()


def test_alone_equals_mix():
    """Batches of up to 4 requests on 2 workers, queued at once so that two batches are on the
    device together: every result equals the request solved alone, to the bit."""
    from ncpamodes import serve
    print("case serve alone-vs-mix, 2 workers, BF16x6F", flush=True)
    reqs = serve.make_requests(16, 100.0, seed=7, n_profiles=5, frac_2d=0.4)
    srv = serve.Server(n_workers=2, precision=PROD, max_batch=4)
    try:
        w = srv.workers[0]
        alone = {}
        for r in reqs:
            w.run_batch([r])
            alone[r.rid] = r.result
            r.result = None
        srv.serve(reqs, realtime=False)
        assert srv.batches >= 4
        assert {r.kind for r in reqs} == {"1d", "2d"}
        # batches on different workers overlapped in time (they were on the device together)
        spans = srv.spans
        overlap = any(a[1] != b[1] and a[2] < b[3] and b[2] < a[3] for i, a in enumerate(spans) for b in spans[i + 1:])
        assert overlap, spans
        for r in reqs:
            assert _results_equal(alone[r.rid], r.result), r.rid
    finally:
        srv.close()


def test_serving_stream_vs_plain_fp64():
    from ncpamodes import serve
    print("case serve stream 150 req at 300/s, BF16x6F", flush=True)
    reqs = serve.make_requests(150, 300.0, seed=11, n_profiles=40, frac_2d=0.2)
    srv = serve.Server(n_workers=2, precision=PROD, max_batch=32)
    try:
        srv.serve(reqs, realtime=True)
        st = serve.latency_stats(reqs)
        print("p50 %.1f ms p99 %.1f ms over %d batches" % (st["p50_ms"], st["p99_ms"], srv.batches), flush=True)
        assert all(r.t_done is not None for r in reqs)
        assert srv.batches < len(reqs)            # requests were batched
    finally:
        srv.close()
    plain = serve.Server(n_workers=1, precision=0, max_batch=1)
    try:
        w = plain.workers[0]
        for r in reqs[:40]:
            got = r.result
            w.run_batch([r])
            ref = r.result
            if r.kind == "1d":
                a, b = _tl_db(got[0]), _tl_db(ref[0])
            else:
                a, b = _tl_db(got), _tl_db(ref)
            mask = b > -140
            assert np.max(np.abs(a - b)[mask]) <= 0.1, (r.rid, r.kind, r.freq)
    finally:
        plain.close()


def test_admission_by_memory():
    """A budget of 40 MB: the batches in flight never exceed it (a batch that alone needs more
    runs by itself), so the stream is split into several batches."""
    from ncpamodes import serve
    print("case serve admission", flush=True)
    reqs = serve.make_requests(30, 1000.0, seed=3, n_profiles=10, frac_2d=0.5)
    budget = 0.04e9
    srv = serve.Server(n_workers=2, precision=PROD, max_batch=64, hbm_budget_gb=budget / 1e9)
    try:
        srv.serve(reqs, realtime=False)
        single = max(serve.batch_footprint([r], srv.Nz, PROD) for r in reqs)
        assert srv.max_in_flight <= max(budget, single) + 1.0
        assert srv.batches >= 3
        assert all(r.result is not None for r in reqs)
        # what was admitted is what the allocator needed: each batch's footprint covers the
        # worker's measured peak (checked per batch in test_footprint_matches_allocator)
        assert all(s[4] > 0 for s in srv.spans)
    finally:
        srv.close()


def test_profiles_with_other_tops_get_their_own_grid(monkeypatch):
    """A profile ending at 120 km is solved on its own grid (maxheight clamped to its top), not
    on the grid of the 150 km profiles in the same batch."""
    import ncpamodes
    from ncpamodes import serve
    base = serve.profile_arrays

    def arrays(seed):
        z, c0, u, v, T, P, rho = base(seed if seed < 1000 else seed - 1000)
        if seed < 1000:
            return z, c0, u, v, T, P, rho
        keep = z <= 120e3
        return z[keep], c0[keep], u[keep], v[keep], T[keep], P[keep], rho[keep]

    monkeypatch.setattr(serve, "profile_arrays", arrays)
    print("case serve per-profile grids", flush=True)
    reqs = [serve.Request(0, 3, 0.4, 80.0, 500.0, "1d"), serve.Request(1, 1003, 0.4, 80.0, 500.0, "1d")]
    srv = serve.Server(n_workers=1, precision=PROD)
    try:
        w = srv.workers[0]
        _, g150, k150 = w.cache.get(3)
        pid120, g120, k120 = w.cache.get(1003)
        assert k150 != k120 and g120.maxheight < 120e3 + 1 and g150.maxheight > 149e3
        w.run_batch(reqs)
        m = w.ctx.solve(g120, [ncpamodes.problem(pid120, 0.4, 80.0)])
        p, inc = m.tl1d(serve.TL1D_DR_M, reqs[1].n_ranges())
        m.free()
        assert np.array_equal(reqs[1].result[0], p[0]) and np.array_equal(reqs[1].result[1], inc[0])
    finally:
        srv.close()
