"""The Fortran driver end to end on the GPU: every file it writes to
result/ must equal, byte for byte, the file the reference build wrote for
the same inputs (slip weakening; FP32).  Covers the seeded-field path
(fault_in.bin) in a full space and with a free surface, and the
reference's own TPV5 and Amatrice 2016 example inputs (cut to 200 steps)."""
import hashlib
import os
import shutil
import sys

import numpy as np
import pytest

import driver_lib as D
import golden_lib as G

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def check_files(wd, digests):
    bad = []
    for f, want in digests.items():
        got = hashlib.sha256(open(os.path.join(wd, "result", f), "rb").read()).hexdigest()
        if got != want:
            bad.append(f)
    assert not bad, f"files differing from the reference: {bad}"


@pytest.mark.parametrize("name,args", [("files_sw_fs", ["sw", "fspace", "file"]),
                                       ("files_sw_free", ["sw", "free", "file"])])
def test_driver_files_identical_to_reference(name, args):
    m = G.manifest()[name]
    wd = D.scratch()
    try:
        src = os.path.join(G.GOLDEN, m["inputs"])
        for f in os.listdir(src):
            shutil.copy(os.path.join(src, f), os.path.join(wd, f))
        say(f"[driver] {name} {' '.join(args)}")
        D.run_driver(wd, *args)
        check_files(wd, m["result_sha256"])
    finally:
        shutil.rmtree(wd, ignore_errors=True)


@pytest.mark.parametrize("name,args", [("files_tpv5_200", ["sw", "free", "tpv5"]),
                                       ("files_amatrice_200", ["sw", "free", "inversion"])])
def test_driver_on_reference_examples(name, args):
    m = G.manifest()[name]
    wd = D.scratch()
    try:
        D.copy_example(m["example"], wd)
        path = os.path.join(wd, "inputfd3d.dat")
        lines = open(path).read().split("\n")
        lines[2] = str(m["steps"])
        open(path, "w").write("\n".join(lines))
        say(f"[driver] {name} {' '.join(args)}")
        D.run_driver(wd, *args)
        check_files(wd, m["result_sha256"])
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if not (np.all(np.isfinite(a)) and np.all(np.isfinite(b))):
        return float("inf")
    return float(np.sqrt(np.sum((a - b) ** 2)) / np.sqrt(np.sum(b ** 2)))


def test_driver_rate_and_state_tpv104_against_fp64_reference():
    """The driver's rate-and-state run of the reference's TPV104 example (free
    surface, smooth nucleation, 100 steps): every per-step stream it writes
    (sliprateX/Z, shearstressX/Z and psi.res, fd3d_theo.f90:546-553) and
    the final slips within relative L2 1e-4 of the reference's FP64 build
    (-fdefault-real-8, as examples/TPV104/compile.sh:13), at every 10th
    fault node (the fixture's sampling)."""
    m = G.manifest()["files_tpv104_100"]
    ref = np.load(os.path.join(G.GOLDEN, m["file"]))
    every, steps = m["node_stride"], m["steps"]
    wd = D.scratch()
    try:
        D.copy_example(m["example"], wd)
        path = os.path.join(wd, "inputfd3d.dat")
        lines = open(path).read().split("\n")
        lines[2] = str(steps)
        open(path, "w").write("\n".join(lines))
        nxT, _nyT, nzT = (int(v) for v in lines[0].split()[:3])
        say(f"[driver] files_tpv104_100 fvw free tpv104 ({nxT}x{nzT} fault, {steps} steps)")
        D.run_driver(wd, "fvw", "free", "tpv104")
        errs = {}
        for n in ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ", "psi"):
            a = np.fromfile(os.path.join(wd, "result", f"{n}.res"), dtype=np.float32)
            assert a.size == steps * nzT * nxT, (n, a.size)
            a = a.reshape(steps, nzT, nxT)[:, ::every, ::every]
            errs[n] = rel_l2(a, ref[f"hist_{n}"])
        for n in ("slipX", "slipZ"):
            a = np.asarray([float(x) for x in open(os.path.join(wd, "result", f"{n}.res")).read().split()])
            assert a.size == ref[f"text_{n}"].size, (n, a.size)
            errs[n] = rel_l2(a, ref[f"text_{n}"])
        say(f"[driver] relative L2 against the FP64 reference: " +
            ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        for n, e in errs.items():
            assert e <= 1e-4, f"{n}: relative L2 {e:.3e} against the FP64 reference"
    finally:
        shutil.rmtree(wd, ignore_errors=True)
