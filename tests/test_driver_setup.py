"""The Fortran driver's input readers and fault set-ups against the
reference's own routines (fd3d_init, readcrustalmodel,
forwardspecialTPV5 / forwardspecialTPV104, readinversionresult), on the
reference's example inputs.  Both write the set-up arrays to
setup_dump.bin without running the time loop, so no GPU is needed."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import driver_lib as D
import refrun

REF = refrun.REF_DIR
EXAMPLES = [
    # example, reference variant, driver args, FVW?
    ("TPV5", "fd3d_ref_sw_free", ["sw", "free", "tpv5", "dump"], False),
    ("TPV104", "fd3d_ref_fvw_free", ["fvw", "free", "tpv104", "dump"], True),
    ("Amatrice2016", "fd3d_ref_sw_free", ["sw", "free", "inversion", "dump"], False),
]


def grid(wd):
    with open(os.path.join(wd, "inputfd3d.dat")) as f:
        lines = f.read().split("\n")
    nxT, nyT, nzT = (int(v) for v in lines[0].split()[:3])
    nabc = int(lines[5].split()[0])
    return nxT + 2 * nabc, nyT + nabc, nzT + nabc + 2


@pytest.mark.parametrize("example,variant,args,fvw", EXAMPLES, ids=[e[0] for e in EXAMPLES])
def test_driver_setup_matches_reference(example, variant, args, fvw):
    exe = os.path.join(REF, variant)
    if not os.access(exe, os.X_OK):
        pytest.skip("reference build oracle/_ref not present")
    if not os.access(D.DRIVER, os.X_OK):
        pytest.fail("Fortran driver not built")
    wref, wdrv = D.scratch(), D.scratch()
    try:
        for wd in (wref, wdrv):
            D.copy_example(example, wd)
            os.makedirs(os.path.join(wd, "result"), exist_ok=True)
        with open(os.path.join(wref, "harness_setup.txt"), "w") as f:
            f.write(args[2] + " dump\n")
        p = subprocess.run([exe], cwd=wref, capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        D.run_driver(wdrv, *args)
        nxt, nyt, nzt = grid(wref)
        n2, n3 = nxt * nzt, nxt * nyt * nzt
        ref = np.fromfile(os.path.join(wref, "setup_dump.bin"), dtype=np.uint8)
        drv = np.fromfile(os.path.join(wdrv, "setup_dump.bin"), dtype=np.uint8)
        if not fvw:
            # 6 fault arrays + lam1, mu1, d1, all float32: identical bytes
            assert ref.size == drv.size == 4 * (6 * n2 + 3 * n3)
            assert np.array_equal(ref, drv)
        else:
            nf = 11 * n2 + 12
            r = ref[: 8 * nf].view(np.float64)
            d = drv[: 8 * nf].view(np.float64)
            assert np.array_equal(r, d), "rate-and-state arrays differ"
            # the medium: FP64 in the reference build, FP32 here
            rm = ref[8 * nf:].view(np.float64)
            dm = drv[8 * nf:].view(np.float32).astype(np.float64)
            assert rm.size == dm.size == 3 * n3
            assert np.allclose(dm, rm, rtol=1e-6)
    finally:
        shutil.rmtree(wref, ignore_errors=True)
        shutil.rmtree(wdrv, ignore_errors=True)
