"""Host-side pieces of the serving loop (no GPU): the relative L2 used for
served-against-plain comparisons and the summary of a run with failed
requests."""
import math

import numpy as np

from fd3d_tsn_amd import serve


def test_rel_l2_is_infinite_on_non_finite_values():
    a = np.ones(10)
    b = np.ones(10)
    assert serve.rel_l2(a, b) == 0.0
    a[3] = np.nan
    assert math.isinf(serve.rel_l2(a, b))
    assert math.isinf(serve.rel_l2(b, a))
    a[3] = np.inf
    assert math.isinf(serve.rel_l2(a, b))
    c = np.ones(10)
    c[0] = 2.0
    assert abs(serve.rel_l2(c, b) - 1.0 / math.sqrt(10.0)) < 1e-15
    assert serve.rel_l2(np.zeros(3), np.zeros(3)) == 0.0


def test_summary_counts_failed_requests_apart():
    ok = [serve.Result(q, 0.0, 0.1, 1.0 + q, 10, 1000) for q in range(4)]
    bad = serve.Result(9, 0.0, 0.1, 2.0, 10, 1000, error="FD3D_E_NEWTON: capped")
    st = serve.summary(ok + [bad], makespan=4.0)
    assert st["requests"] == 4 and st["failed"] == 1
    assert st["ruptures_per_s"] == 1.0
    assert st["Gcell_updates_per_s"] == 4 * 10 * 1000 / 4.0 / 1e9
    assert st["latency_p50_s"] == float(np.percentile([1.0, 2.0, 3.0, 4.0], 50))


def test_c5_stream_mix():
    reqs = serve.c5_requests(2000)
    fvw = sum(r.case.friction == "fvw" for r in reqs)
    assert 0.25 < fvw / len(reqs) < 0.35
    arrivals = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arrivals) > 0)
    rate = len(reqs) / arrivals[-1]
    assert 18.0 < rate < 22.0
    for r in reqs[:50]:
        assert 64 <= r.case.nxtT <= 384 and 64 <= r.case.nytT <= 192 and 64 <= r.case.nztT <= 192
        assert 800 <= r.steps <= 3000 and r.case.cfl() <= 0.25
