"""The batch that allocates no device memory (fd3d_batch_create_static)
against the allocating batch (fd3d_batch_create): two batches in a row on
the fixed device tables, the second with fewer and smaller problems (so it
runs on the first one's leftover table contents), each member's outputs,
stations and final arrays bit-identical to the same problems advanced by
an allocating batch.
"""
import sys

import numpy as np
import pytest

from fd3d_tsn_amd import capi
from fd3d_tsn_amd import cases as C

pytestmark = pytest.mark.gpu

KEYS = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ", "psi", "seis")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(cases, nsteps, static, fused):
    solvers = [capi.Solver(c) for c in cases]
    try:
        if fused:
            for s in solvers:
                s.set_tuning(1 | 4, 8, 16, 8, 16)
        with capi.Batch(len(cases), tz=8, mode=capi.STAGED, static=static) as b:
            if fused:
                b.set_fused_items(0)
            for s in solvers:
                b.add(s)
            outs = [{k: [] for k in KEYS} for _ in cases]
            for _ in range(nsteps // 8):
                b.step(8)
                b.sync()
                for q, s in enumerate(solvers):
                    o = s.read_outputs(8)
                    for k in KEYS:
                        outs[q][k].append(o[k])
            return [({k: np.concatenate(v) for k, v in o.items()}, s.final()) for o, s in zip(outs, solvers)]
    finally:
        for s in solvers:
            s.close()


def check(tag, got, ref):
    for q, ((go, gf), (ro, rf)) in enumerate(zip(got, ref)):
        for k in KEYS:
            assert np.array_equal(go[k], ro[k], equal_nan=True), f"{tag} member {q} {k}"
        for k, v in rf.items():
            assert np.array_equal(gf[k], v), f"{tag} member {q} final {k}"


def test_static_batch_equals_allocating_batch_twice():
    first = [C.tiny("sw", True, seed=61), C.tiny("fvw", False, nabc=4, n=(50, 12, 18), seed=62),
             C.tiny("sw", False, nabc=5, n=(70, 13, 19), seed=63)]
    second = [C.tiny("sw", True, nabc=4, n=(30, 9, 12), seed=64), C.tiny("fvw", True, nabc=4, n=(24, 8, 10), seed=65)]
    for tag, cases, fused in (("first", first, True), ("second", second, False)):
        say(f"[no-alloc] {tag}: {len(cases)} problems, fused {fused}")
        ref = run(cases, 24, static=False, fused=fused)
        got = run(cases, 24, static=True, fused=fused)
        check(tag, got, ref)
        say(f"[no-alloc] {tag}: bit-identical to the allocating batch")


def test_static_batch_limits():
    with pytest.raises(capi.FD3DError) as e:
        capi.Batch(65, tz=8, static=True)
    assert e.value.code == -1
    with capi.Batch(2, tz=8, static=True):
        with pytest.raises(capi.FD3DError) as e:
            capi.Batch(2, tz=8, static=True)    # the device's fixed tables are taken
        assert e.value.code == -5
    with capi.Batch(2, tz=8, static=True):       # released with the first batch
        pass
