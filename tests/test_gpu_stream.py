"""The overlapped output stream (fd3d_step_begin / fd3d_step_wait): the
chunks it hands out hold exactly the slabs and station samples fd3d_step
returns (which the oracle pins bit for bit), for chunk sizes that wrap
around the device ring, with the GPU running the next chunk while the host
reads one; misuse and the error flags are reported per chunk.
"""
import sys

import numpy as np
import pytest

from fd3d_tsn_amd import capi
from fd3d_tsn_amd import cases as C
import oracle_lib as O

pytestmark = pytest.mark.gpu

KEYS = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ", "psi", "seis")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def with_stations(case):
    nx, ny, nz = case.nxt, case.nyt, case.nzt
    return case.replace(stations=[(nx // 2, ny // 2, nz // 2), (nx // 2 + 1, ny, nz // 2), (2, 2, 2)])


@pytest.mark.parametrize("friction,fspace,chunk,ring", [("sw", True, 32, 64), ("sw", False, 7, 16),
                                                       ("fvw", False, 5, 12), ("fvw", True, 1, 2)],
                         ids=["sw_fs_32", "sw_free_7", "fvw_free_5", "fvw_fs_1"])
def test_stream_equals_step(friction, fspace, chunk, ring):
    case = with_stations(C.tiny(friction, fspace, ntfd=75, seed=71))
    say(f"[stream] {case.name}: chunks of {chunk} on a ring of {ring}, {case.ntfd} steps")
    with capi.Solver(case, ring_steps=ring) as s:
        ref = s.step(case.ntfd)
    got = {k: [] for k in KEYS}
    with capi.Solver(case, ring_steps=ring) as s:
        for res in s.stream(case.ntfd, chunk):
            for k in KEYS:
                got[k].append(res[k])
        assert s.steps_done() == case.ntfd
    for k in KEYS:
        a = np.concatenate(got[k])
        assert np.array_equal(a, ref[k]), k
    if friction == "sw":
        orc = O.Oracle(case, O.F32)
        ro = orc.step(case.ntfd)
        for k in ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ", "seis"):
            assert np.array_equal(np.concatenate(got[k]).astype(np.float64), ro[k]), k


def test_stream_misuse_and_flags():
    case = C.tiny("sw", True, ntfd=20, seed=72)
    with capi.Solver(case, ring_steps=16) as s:
        for bad in (0, 9):
            with pytest.raises(capi.FD3DError) as e:
                s.step_begin(bad)         # 1 <= n <= ring/2
            assert e.value.code == -1
        with pytest.raises(capi.FD3DError) as e:
            s.step_wait()                 # nothing in flight
        assert e.value.code == -5
        s.step_begin(8)
        s.step_begin(8)
        with pytest.raises(capi.FD3DError) as e:
            s.step_begin(1)               # two chunks in flight
        assert e.value.code == -5
        with pytest.raises(capi.FD3DError) as e:
            s.step(1)
        assert e.value.code == -5
        with capi.Batch(1, tz=8) as b:
            with pytest.raises(capi.FD3DError) as e:
                b.add(s)
            assert e.value.code == -5
        s.step_wait()
        s.step_wait()
        s.step(2)                         # usable again
    fvw = C.tiny("fvw", False, nabc=4, n=(50, 12, 18), seed=73, ntfd=16)
    with capi.Solver(fvw, newton_max=1, ring_steps=16) as s:
        s.step_begin(4)
        with pytest.raises(capi.FD3DError) as e:
            s.step_wait()
        assert e.value.code == -6         # FD3D_E_NEWTON from the chunk's own flags


# Calls of 1, 2, 3 and 4 steps, streamed chunks, steps without outputs and
# kernel-by-kernel (profiled) steps, mixed: 40 steps in all.
CALLS = [1, 2, ("stream", 5, 2), 3, ("prof", 2), ("quiet", 3), 1, ("stream", 8, 8), 4, ("prof", 1),
         ("stream", 3, 1), 2, ("quiet", 1), 4]


@pytest.mark.parametrize("friction,fspace", [("sw", True), ("fvw", False)], ids=["sw_fs", "fvw_free"])
def test_call_pattern_equals_one_call(friction, fspace):
    """A run of steps replays a head graph, rotated graphs (each step's
    fault-plane kernels beside the next step's velocity sweeps) and a tail
    graph, so every call returns with its last step complete.  Any mix of
    call sizes and forms must then give, bit for bit, the slabs, stations,
    final arrays and the nine 3D fields of one call of the same length."""
    case = with_stations(C.tiny(friction, fspace, ntfd=40, seed=74))
    assert sum(c if isinstance(c, int) else c[1] for c in CALLS) == case.ntfd
    say(f"[calls] {case.name}: {len(CALLS)} calls, {case.ntfd} steps")
    with capi.Solver(case, ring_steps=16) as s:
        ref = s.step(case.ntfd)
        ref_fin = s.final()
        ref_f = {n: s.field(n) for n in capi.FIELDS}
    got = {k: [] for k in KEYS}
    with capi.Solver(case, ring_steps=16) as s:
        for c in CALLS:
            if isinstance(c, int):
                parts = [s.step(c)]
            elif c[0] == "stream":
                parts = list(s.stream(c[1], c[2]))
            elif c[0] == "prof":
                s.set_profiling(True)
                parts = [s.step(c[1])]
                s.set_profiling(False)
            else:
                s.step(c[1], outputs=False)
                parts = [s.read_outputs(c[1])]
            for r in parts:
                for k in KEYS:
                    got[k].append(r[k])
        assert s.steps_done() == case.ntfd
        fin = s.final()
        for n in capi.FIELDS:
            assert np.array_equal(s.field(n), ref_f[n]), n
    for k in KEYS:
        assert np.array_equal(np.concatenate(got[k]), ref[k]), k
    for n, v in ref_fin.items():
        assert np.array_equal(fin[n], v), n
    say(f"[calls] {case.name}: bit-identical to one call")
