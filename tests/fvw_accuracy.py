"""Error of the GPU rate-and-state path, per output, against the oracle in
its all-double mode (which matches the reference built with
-fdefault-real-8 to 1e-13) and in its mixed mode (float wavefield, double
Newton solve and state update: what the GPU computes), on a 96x48x48 grid
with 10-node PML over 1500 steps, free surface and full space.

    python tests/fvw_accuracy.py     (on a GPU host; prints one JSON line per case)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from fd3d_tsn_amd import capi  # noqa: E402
from fd3d_tsn_amd import cases as C  # noqa: E402
import oracle_lib as O  # noqa: E402


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / den)


def main(nsteps=1500):
    for fspace in (False, True):
        case = C.tiny(friction="fvw", fspace=fspace, nabc=10, n=(96, 48, 48), ntfd=nsteps, seed=21)
        fault, scal = C.fault_fvw(case)
        f64 = O.Oracle(case, O.F64, fault=fault, fvw_scalars=scal)
        r64 = f64.step(case.ntfd)
        fin64 = f64.final()
        mx = O.Oracle(case, O.MIXED, fault=fault, fvw_scalars=scal)
        rmx = mx.step(case.ntfd)
        with capi.Solver(case, fault=fault, fvw_scalars=scal) as s:
            rg = s.step(case.ntfd)
            fg = s.final()
        out = {"case": case.name, "grid": [case.nxt, case.nyt, case.nzt], "steps": case.ntfd}
        for n in ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ", "psi"):
            out[n] = {"relL2_vs_fp64": rel_l2(rg[n], r64[n]),
                      "relL2_vs_mixed": rel_l2(rg[n], rmx[n].astype(np.float32))}
        ok = fin64["ruptime"] < 1e4
        out["ruptime"] = {"ruptured_nodes": int(ok.sum()), "same_nodes": bool(np.array_equal(fg["ruptime"] < 1e4, ok)),
                          "max_abs_in_dt_vs_fp64":
                              float(np.max(np.abs(fg["ruptime"][ok] - fin64["ruptime"][ok])) / case.dt) if ok.any() else None}
        for n in ("slipX", "slipZ"):
            out[n] = {"relL2_vs_fp64": rel_l2(fg[n], fin64[n])}
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
