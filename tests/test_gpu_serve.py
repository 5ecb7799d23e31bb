"""GPU: the serving loop (continuous batching over several device contexts, admission by
device-memory footprint).  Each result must be the same alone and in the mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tl_db(p):
    return 20 * np.log10(np.maximum(np.abs(p), 1e-300))


def _results_equal(a, b):
    if isinstance(a, tuple):
        return all(np.array_equal(x, y) for x, y in zip(a, b))
    return np.array_equal(a, b)


def test_alone_equals_mix():
    from ncpamodes import serve
    print("case serve alone-vs-mix", flush=True)
    reqs = serve.make_requests(10, 100.0, seed=7, n_profiles=5, frac_2d=0.4)
    srv = serve.Server(n_workers=1, precision=3, max_batch=16)
    try:
        w = srv.workers[0]
        alone = {}
        for r in reqs:
            w.run_batch([r])
            alone[r.rid] = r.result
            r.result = None
        srv.serve(reqs, realtime=False)
        assert srv.batches == 1
        kinds = {r.kind for r in reqs}
        assert kinds == {"1d", "2d"}
        for r in reqs:
            assert _results_equal(alone[r.rid], r.result), r.rid
    finally:
        srv.close()


def test_serving_stream_vs_plain_fp64():
    from ncpamodes import serve
    print("case serve stream 150 req at 300/s", flush=True)
    reqs = serve.make_requests(150, 300.0, seed=11, n_profiles=40, frac_2d=0.2)
    srv = serve.Server(n_workers=2, precision=3, max_batch=32)
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
    from ncpamodes import serve
    print("case serve admission", flush=True)
    reqs = serve.make_requests(30, 1000.0, seed=3, n_profiles=10, frac_2d=0.5)
    budget = 0.05e9
    srv = serve.Server(n_workers=2, precision=3, max_batch=64, hbm_budget_gb=budget / 1e9)
    try:
        srv.serve(reqs, realtime=False)
        fps = [r.footprint_bytes(srv.Nz) for r in reqs]
        assert srv.max_in_flight <= max(budget, max(fps)) + 1.0
        assert srv.batches >= int(np.ceil(sum(fps) / max(budget, max(fps))))
        assert srv.batches >= 3
        assert all(r.result is not None for r in reqs)
    finally:
        srv.close()
