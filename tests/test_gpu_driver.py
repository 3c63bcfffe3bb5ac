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
