"""The serving loop (continuous batching) against back-to-back plain solves.

Requests of different sizes, friction laws and step counts arrive over
time; with three batch slots they queue, join the running batch at chunk
boundaries and leave when done.  Each request's final arrays must equal
the same request solved alone (bit for bit), whatever it shared the GPU
with.
"""
import sys

import numpy as np
import pytest

from fd3d_tsn_amd import capi
from fd3d_tsn_amd import cases as C
from fd3d_tsn_amd import serve

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def requests():
    specs = [
        (0.00, C.tiny("sw", True, seed=51, ntfd=40)),
        (0.00, C.tiny("fvw", False, nabc=4, n=(50, 12, 18), seed=52, ntfd=56)),
        (0.02, C.tiny("sw", False, nabc=5, n=(37, 13, 19), seed=53, ntfd=24)),
        (0.05, C.tiny("sw", True, nabc=4, n=(70, 9, 33), seed=54, ntfd=48)),
        (0.05, C.tiny("fvw", True, seed=55, ntfd=32)),
        (0.30, C.tiny("sw", True, seed=56, ntfd=20)),
    ]
    reqs = [serve.Request(q, t, c) for q, (t, c) in enumerate(specs)]
    # two requests also accumulate the inversion bins
    for r in (reqs[0], reqs[3]):
        nxT, nzT = r.case.fault_shape
        r.bins = (2, 2, nxT * r.case.dh / 2, nzT * r.case.dh / 2, 0.016)
    return reqs


def test_served_equals_alone():
    mode = capi.STAGED
    reqs = requests()
    say(f"[serve] {len(reqs)} requests, mode {mode}")
    plain = {r.rid: r for r in serve.run_plain(reqs)}
    srv = serve.Server(max_slots=3, chunk=8, tz=8, mode=mode, reserve_bytes=0, log=say)
    try:
        got = srv.run(reqs)
    finally:
        srv.close()
    assert sorted(r.rid for r in got) == [r.rid for r in reqs]
    for r in got:
        ref = plain[r.rid]
        for k, v in ref.final.items():
            assert np.array_equal(r.final[k], v), f"request {r.rid} final {k}"
        assert sorted(r.bins) == sorted(ref.bins)
        for k, v in ref.bins.items():
            assert np.array_equal(r.bins[k], v), f"request {r.rid} bins {k}"
        assert r.finished >= r.admitted >= 0.0 and r.latency > 0.0
    stats = serve.summary(got, srv.makespan)
    say(f"[serve] {stats}")
    assert stats["requests"] == len(reqs)



def test_failed_member_leaves_the_others_running():
    """A rate-and-state request whose Newton solve hits its cap (newton_max
    1) fails alone: it comes back with the error, and every other request
    of the mix still equals its solo run bit for bit."""
    reqs = requests()
    bad = 1
    assert reqs[bad].case.friction == "fvw"
    say(f"[serve] request {bad} runs with newton_max 1")
    plain = {r.rid: r for r in serve.run_plain([r for r in reqs if r.rid != bad])}
    srv = serve.Server(max_slots=3, chunk=8, tz=8, mode=capi.STAGED, reserve_bytes=0, log=say,
                       solver_args=lambda r: {"newton_max": 1} if r.rid == bad else {})
    try:
        got = {r.rid: r for r in srv.run(reqs)}
    finally:
        srv.close()
    assert sorted(got) == [r.rid for r in reqs]
    assert got[bad].error is not None and "FD3D_E_NEWTON" in got[bad].error, got[bad].error
    for rid, ref in plain.items():
        r = got[rid]
        assert r.error is None, (rid, r.error)
        for k, v in ref.final.items():
            assert np.array_equal(r.final[k], v), f"request {rid} final {k}"
        for k, v in ref.bins.items():
            assert np.array_equal(r.bins[k], v), f"request {rid} bins {k}"
    stats = serve.summary(list(got.values()), srv.makespan)
    assert stats["failed"] == 1 and stats["requests"] == len(reqs) - 1
    # the back-to-back baseline fails the same request alone too
    alone = {r.rid: r for r in serve.run_plain(reqs, solver_args=lambda r: {"newton_max": 1} if r.rid == bad else {})}
    assert alone[bad].error is not None and "FD3D_E_NEWTON" in alone[bad].error, alone[bad].error
    for rid, ref in plain.items():
        assert alone[rid].error is None, (rid, alone[rid].error)
        for k, v in ref.final.items():
            assert np.array_equal(alone[rid].final[k], v), f"request {rid} final {k} (back to back)"
