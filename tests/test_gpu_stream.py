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
