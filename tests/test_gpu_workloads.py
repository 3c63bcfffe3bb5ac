"""The batched (64 uneven domains) and served (Poisson stream) workloads at
their real sizes, against the oracle.

* Four of the 64 batched domains (interior sides 96-320: the smallest, the
  largest and two between) advance together in one batch, the fused
  interior sweep forced on as a batch stage (partial edge tiles in x and z
  of uneven domains, several tile strips), for 16 steps.  Every output
  slab, station sample, final array and the nine 3D fields equal the
  oracle's FP32 run bit for bit.  The same four run again with the separate
  sweeps as batch stages.
* All 64 domains advance 3 steps in one batch with the batch plan's shapes;
  each equals the same domain stepped alone with its own best shape.
* The first 8 requests of the 2000-request stream, at their real grid
  sizes and fields (step counts cut to 24), go through the serving loop;
  slip-weakening results equal the FP32 oracle bit for bit, rate-and-state
  results are within 1e-9 relative L2 of the mixed-precision oracle (float
  wavefield, double Newton solve and state: what the GPU computes; the gap
  is the last ulp of double transcendentals, device against host libm).

The per-rupture contract is the reference's time loop
(src/fd3d_theo.f90:187-602): each domain is an independent fd3d() run.
"""
import functools
import sys

import numpy as np
import pytest

from fd3d_tsn_amd import autotune
from fd3d_tsn_amd import capi
from fd3d_tsn_amd import cases as C
from fd3d_tsn_amd import serve
import oracle_lib as O

pytestmark = pytest.mark.gpu

OUTS = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ")
KEYS = OUTS + ("psi", "seis")
TOL_MIXED = 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def stations(case):
    """Stations in the interior, next to the fault and in the PML."""
    nx, ny, nz, a = case.nxt, case.nyt, case.nzt, case.nabc
    return [(nx // 2, ny // 2, nz // 2), (nx // 2 + 1, ny, nz // 2), (a // 2, ny - 3, nz // 3),
            (nx - a // 2, a // 2, nz - a // 2), (nx // 3, a + 1, a + 2)]


def c4_four():
    """Smallest, largest and the two thirds points (by cells) of the 64 domains."""
    cs = sorted(C.c4_batched(64), key=lambda c: c.cells)
    pick = [cs[0], cs[len(cs) // 3], cs[2 * len(cs) // 3], cs[-1]]
    return [c.replace(stations=stations(c)) for c in pick]


C4_STEPS = 16


@functools.lru_cache(maxsize=None)
def c4_oracle(q):
    case = c4_four()[q]
    say(f"[c4] oracle {case.name}: {case.nxt}x{case.nyt}x{case.nzt}, {C4_STEPS} steps")
    orc = O.Oracle(case, O.F32)
    ro = orc.step(C4_STEPS)
    fo = orc.final()
    fields = {n: orc.field(n).astype(np.float32) for n in capi.FIELDS}
    return ro, fo, fields


@pytest.mark.parametrize("form", ["fused", "separate"])
def test_c4_domains_in_a_batch_equal_the_oracle(form):
    cases = c4_four()
    for c in cases:
        nxi = c.nxt - 2 * c.nabc
        say(f"[c4] {form}: {c.name} interior x {nxi} ({-(-nxi // 64)} fused tiles across)")
    solvers = [capi.Solver(c) for c in cases]
    try:
        tz = 8
        for s in solvers:
            nyi = s.case.nyt - 2 - s.case.nabc
            if form == "fused":
                s.set_tuning(1 | 4, tz, min(nyi, autotune.FUSED_BATCH_NY), tz, 32)
            else:
                s.set_tuning(1, tz, 24, tz, 40)
        got = [{k: [] for k in KEYS} for _ in cases]
        with capi.Batch(len(cases), tz=tz, mode=capi.STAGED) as b:
            if form == "fused":
                b.set_fused_items(0)     # four members hold too few fused items for the default threshold
            for s in solvers:
                b.add(s)
            for _ in range(C4_STEPS // 8):
                b.step(8)
                b.sync()
                for q, s in enumerate(solvers):
                    o = s.read_outputs(8)
                    for k in KEYS:
                        got[q][k].append(o[k])
            say(f"[c4] {form}: batch {b.info()}")
            for q, (s, c) in enumerate(zip(solvers, cases)):
                ro, fo, fields = c4_oracle(q)
                res = {k: np.concatenate(v) for k, v in got[q].items()}
                for n in OUTS:
                    assert np.array_equal(res[n].astype(np.float64), ro[n]), f"{c.name} {n}"
                assert np.array_equal(res["seis"].astype(np.float64), ro["seis"]), f"{c.name} stations"
                fin = s.final()
                for n, v in fo.items():
                    assert np.array_equal(fin[n].astype(np.float64), v), f"{c.name} final {n}"
                for n in capi.FIELDS:
                    bad = np.argwhere(s.field(n) != fields[n])
                    assert bad.size == 0, f"{c.name} field {n}: {len(bad)} cells differ, first {bad[0].tolist()}"
                say(f"[c4] {form}: {c.name} bit-exact against the oracle")
    finally:
        for s in solvers:
            s.close()


def test_c4_all_64_domains_batched_equal_alone():
    nsteps = 3
    cases = C.c4_batched(64)
    tz, shapes = autotune.batch_plan(cases)
    say(f"[c4] 64 domains, {sum(c.cells for c in cases) / 1e6:.0f}M cells, tz {tz}, {nsteps} steps")
    refs = []
    for c in cases:
        with capi.Solver(c) as s:
            autotune.apply(s)
            refs.append((s.step(nsteps), s.final()))
    say("[c4] alone: done")
    solvers = []
    try:
        with capi.Batch(len(cases), tz=tz, mode=capi.STAGED) as b:
            for c, t in zip(cases, shapes):
                s = capi.Solver(c)
                solvers.append(s)
                if t is not None:
                    s.set_tuning(*t)
                b.add(s)
            b.step(nsteps)
            b.sync()
            say(f"[c4] batch {b.info()}")
            for s, (ro, fo) in zip(solvers, refs):
                o = s.read_outputs(nsteps)
                for k in KEYS:
                    assert np.array_equal(o[k], ro[k]), f"{s.case.name} {k}"
                fin = s.final()
                for n, v in fo.items():
                    assert np.array_equal(fin[n], v), f"{s.case.name} final {n}"
    finally:
        for s in solvers:
            s.close()


C5_STEPS = 24
C5_N = 8


def test_c5_first_requests_served_against_the_oracle():
    reqs = serve.c5_requests(2000)[:C5_N]
    reqs = [serve.Request(r.rid, r.arrival, r.case.replace(ntfd=C5_STEPS, stations=stations(r.case)))
            for r in reqs]
    for r in reqs:
        say(f"[c5] request {r.rid}: {r.case.name} ({r.case.friction}, {'full space' if r.case.fspace else 'free surface'}), "
            f"arrives {r.arrival:.3f} s")
    srv = serve.Server(max_slots=8, chunk=8, tz=16, mode=capi.STAGED, reserve_bytes=0, keep_outputs=True, log=say)
    try:
        got = {r.rid: r for r in srv.run(reqs)}
    finally:
        srv.close()
    assert sorted(got) == [r.rid for r in reqs]
    for r in reqs:
        res = got[r.rid]
        assert res.error is None, res.error
        fvw = r.case.friction == "fvw"
        orc = O.Oracle(r.case, O.MIXED if fvw else O.F32)
        ro = orc.step(C5_STEPS)
        fo = orc.final()
        if not fvw:
            for n in OUTS:
                assert np.array_equal(res.outputs[n].astype(np.float64), ro[n]), f"request {r.rid} {n}"
            assert np.array_equal(res.outputs["seis"].astype(np.float64), ro["seis"]), f"request {r.rid} stations"
            for n, v in fo.items():
                assert np.array_equal(res.final[n].astype(np.float64), v), f"request {r.rid} final {n}"
            say(f"[c5] request {r.rid}: bit-exact against the FP32 oracle")
        else:
            # the slabs are float (psi.res included); the mixed oracle's psi is double
            errs = {n: serve.rel_l2(res.outputs[n], ro[n].astype(np.float32)) for n in OUTS + ("psi", "seis")}
            errs.update({"final " + n: serve.rel_l2(res.final[n], v) for n, v in fo.items()})
            say(f"[c5] request {r.rid}: relative L2 against the mixed oracle, worst "
                f"{max(errs.values()):.2e} ({max(errs, key=errs.get)})")
            for n, e in errs.items():
                assert e <= TOL_MIXED, f"request {r.rid} {n}: relative L2 {e:.3e}"
