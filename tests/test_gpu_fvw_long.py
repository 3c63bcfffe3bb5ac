"""Rate-and-state accuracy over a long run, per output, against the oracle
in its all-double mode (which matches the reference built with
-fdefault-real-8, the -r8 build of examples/TPV104/compile.sh:13, to
1e-13: tests/test_oracle_vs_ref.py) and in its mixed mode (float
wavefield, double Newton solve and state update: what the GPU computes).

96x48x48 interior, 10-node PML, 1500 steps, free surface and full space.
Bar: every per-step output history (slip rates, shear tractions, the state
psi of result/psi.res, fd3d_theo.f90:546-553) and the final slips within
relative L2 1e-4 of the FP64 oracle; the same nodes rupture, at times
within one time step of the FP64 oracle's.  The error against FP64 is the
FP32 bulk wavefield, as the single-precision target specifies.
"""
import json
import sys

import numpy as np
import pytest

from fd3d_tsn_amd import capi
from fd3d_tsn_amd import cases as C
from fd3d_tsn_amd import serve
import oracle_lib as O

pytestmark = pytest.mark.gpu

OUTS = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ", "psi")
TOL_FP64 = 1e-4
STEPS = 1500


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("fspace", [False, True], ids=["free", "fs"])
def test_fvw_long_run_against_fp64(fspace):
    case = C.tiny(friction="fvw", fspace=fspace, nabc=10, n=(96, 48, 48), ntfd=STEPS, seed=21)
    say(f"[fvw-long] {case.name}: {case.nxt}x{case.nyt}x{case.nzt}, {STEPS} steps")
    fault, scal = C.fault_fvw(case)
    f64 = O.Oracle(case, O.F64, fault=fault, fvw_scalars=scal)
    r64 = f64.step(STEPS)
    fin64 = f64.final()
    mx = O.Oracle(case, O.MIXED, fault=fault, fvw_scalars=scal)
    rmx = mx.step(STEPS)
    with capi.Solver(case, fault=fault, fvw_scalars=scal) as s:
        rg = s.step(STEPS)
        fg = s.final()
    rep = {"case": case.name, "steps": STEPS}
    for n in OUTS:
        # vs_mixed against the mixed oracle's values rounded to the float slabs
        rep[n] = {"vs_fp64": serve.rel_l2(rg[n], r64[n]),
                  "vs_mixed": serve.rel_l2(rg[n], rmx[n].astype(np.float32))}
    for n in ("slipX", "slipZ"):
        rep["final " + n] = {"vs_fp64": serve.rel_l2(fg[n], fin64[n])}
    ok = fin64["ruptime"] < 1e4
    same = bool(np.array_equal(fg["ruptime"] < 1e4, ok))
    dtmax = float(np.max(np.abs(fg["ruptime"][ok] - fin64["ruptime"][ok])) / case.dt) if ok.any() else 0.0
    rep["ruptime"] = {"ruptured_nodes": int(ok.sum()), "same_nodes": same, "max_diff_in_dt": dtmax}
    say("[fvw-long] " + json.dumps(rep))
    assert ok.sum() > 0, "no node ruptured: the run does not exercise the rupture front"
    for n, e in rep.items():
        if isinstance(e, dict) and "vs_fp64" in e:
            assert e["vs_fp64"] <= TOL_FP64, f"{n}: relative L2 {e['vs_fp64']:.3e} against FP64"
    assert same, "the GPU and the FP64 oracle rupture different nodes"
    assert dtmax <= 1.0, f"rupture time differs by {dtmax:.3f} time steps"
