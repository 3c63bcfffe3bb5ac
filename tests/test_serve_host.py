"""Host-side pieces of the serving loop (no GPU): the relative L2 used for
served-against-plain comparisons and the summary of a run with failed
requests."""
import math

import numpy as np
import pytest

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


class _FakeError(Exception):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class _FakeSolver:
    """Stands in for capi.Solver: counts the steps the batch gives it."""
    ring_steps = 64

    def __init__(self, log, case, **kw):
        self.log, self.case, self.done, self.closed, self.batch = log, case, 0, False, None
        self.bad = kw.get("newton_max") == 1
        log.append(("made", case.rid))

    def set_source_bins(self, *a, **k):
        pass

    def set_tuning(self, *a):
        pass

    def sync(self):
        if self.bad:
            raise _FakeError(-6, "FD3D_E_NEWTON: capped")

    def final(self):
        assert self.batch is None and not self.closed, "results read while in the batch or after close"
        return {"steps": np.array([self.done])}

    def close(self):
        assert self.batch is None and not self.closed
        self.closed = True
        self.log.append(("close", self.case.rid))


class _FakeBatch:
    """Stands in for capi.Batch: records each chunk's members and advances them."""

    def __init__(self, log, max_slots, **kw):
        self.log, self.max_slots, self.members, self.in_flight = log, max_slots, [], False

    def add(self, s):
        assert not self.in_flight and len(self.members) < self.max_slots and s.batch is None
        s.batch = self
        self.members.append(s)

    def remove(self, s):
        assert not self.in_flight
        s.batch = None
        self.members.remove(s)

    def step(self, n):
        assert not self.in_flight and self.members
        for s in self.members:
            assert s.done + n <= s.case.ntfd, "a member stepped past its run"
            s.done += n
        self.in_flight = True
        self.log.append(("chunk", n, sorted(s.case.rid for s in self.members)))

    def sync(self):
        self.in_flight = False
        if any(s.bad for s in self.members):
            raise _FakeError(-6, "FD3D_E_NEWTON in a member")

    def elapsed(self):
        return (1.0,)

    def close(self):
        pass


class _Case:
    def __init__(self, rid, ntfd):
        self.rid, self.ntfd, self.cells = rid, ntfd, 1000


def _fake_server(monkeypatch, log, **kw):
    class FakeCapi:
        STAGED = 1
        FD3DError = _FakeError
        Batch = staticmethod(lambda max_slots, **k: _FakeBatch(log, max_slots, **k))
        Solver = staticmethod(lambda case, **k: _FakeSolver(log, case, **k))
        device_memory = staticmethod(lambda: (1 << 40, 1 << 40))
        footprint = staticmethod(lambda case: 1 << 20)

    monkeypatch.setattr(serve, "capi", FakeCapi)
    monkeypatch.setattr(serve.autotune, "choose", lambda case, tz: None)
    monkeypatch.setattr(serve.autotune, "batch_member", lambda *a: None)
    return serve.Server(mode=FakeCapi.STAGED, reserve_bytes=0, **kw)


@pytest.mark.parametrize("prefetch", [0, 2])
def test_server_refills_slots_at_the_boundary_they_free(monkeypatch, prefetch):
    """Scheduling of the serving loop, on fakes: every request gets exactly
    its steps; no chunk runs with a free slot while requests wait (a slot
    freed at a chunk's end is filled at that boundary); at most `prefetch`
    prepared requests wait beyond the slots; results are read only after
    the member left the batch, and every solver is freed once."""
    log = []
    steps = [20, 8, 35, 13, 8, 50, 27, 16, 9, 40]
    reqs = [serve.Request(q, 0.0, _Case(q, n)) for q, n in enumerate(steps)]
    srv = _fake_server(monkeypatch, log, max_slots=3, chunk=8, prefetch=prefetch)
    got = srv.run(reqs)
    assert sorted(r.rid for r in got) == list(range(len(reqs)))
    for r in got:
        assert r.error is None and int(r.final["steps"][0]) == steps[r.rid]
    assert sorted(e[1] for e in log if e[0] == "close") == list(range(len(reqs)))
    chunks = [e for e in log if e[0] == "chunk"]
    started = set()
    for _tag, n, members in chunks:
        started.update(members)
        waiting = len(reqs) - len(started)
        assert len(members) == 3 or waiting == 0, (members, waiting)
    assert srv.chunks == len(chunks)
    # prepared solvers not yet in the batch, at every chunk
    made = 0
    for e in log:
        if e[0] == "made":
            made += 1
        elif e[0] == "chunk":
            closed = sum(1 for f in log[:log.index(e)] if f[0] == "close")
            assert made - closed - len(e[2]) <= prefetch + 3, (made, closed, e)


def test_server_failed_member_leaves_alone(monkeypatch):
    """A member whose Newton solve hits its cap leaves with the error after
    the chunk that reported it; the others run to the end."""
    log = []
    steps = [24, 24, 24, 24]
    reqs = [serve.Request(q, 0.0, _Case(q, n)) for q, n in enumerate(steps)]
    srv = _fake_server(monkeypatch, log, max_slots=3, chunk=8,
                       solver_args=lambda r: {"newton_max": 1} if r.rid == 1 else {})
    got = {r.rid: r for r in srv.run(reqs)}
    assert got[1].error is not None and "FD3D_E_NEWTON" in got[1].error and got[1].final == {}
    for q in (0, 2, 3):
        assert got[q].error is None and int(got[q].final["steps"][0]) == 24
    assert sorted(e[1] for e in log if e[0] == "close") == [0, 1, 2, 3]
