"""Run the flang-built reference (oracle/_ref/fd3d_ref_*) on a Case.

The binaries are built from the unmodified reference sources by
``make -C oracle ref`` (only where the reference tree is present) and are
driven by oracle/ref_harness.f90.  This module writes the reference's
input files into a scratch directory, runs the binary there and reads back
its outputs: the per-step result/*.res streams (fd3d_theo.f90:546-553),
the station text files (:555-557) and the final 2D arrays.
"""
from __future__ import annotations

import os
import resource
import subprocess
import tempfile
import time
from typing import Dict, Optional

import numpy as np

from fd3d_tsn_amd import cases as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def variant(case: C.Case, fast: bool = False) -> str:
    v = ("fvw" if case.friction == "fvw" else "sw") + ("_fs" if case.fspace else "_free")
    return os.path.join(REF_DIR, f"fd3d_ref_{v}" + ("_fast" if fast else ""))


def available(case: C.Case, fast: bool = False) -> bool:
    return os.access(variant(case, fast), os.X_OK)


def _f64(case: C.Case) -> bool:
    return case.friction == "fvw"


def write_inputs(case: C.Case, wd: str, fault: Optional[Dict] = None, fvw_scalars=None):
    os.makedirs(os.path.join(wd, "result"), exist_ok=True)
    with open(os.path.join(wd, "inputfd3d.dat"), "w") as f:
        f.write(case.inputfd3d_text())
    with open(os.path.join(wd, "crustal.dat"), "w") as f:
        f.write(case.crustal_text())
    with open(os.path.join(wd, "inputinv.dat"), "w") as f:
        f.write("0\n1 1\n")
    with open(os.path.join(wd, "fault_in.bin"), "wb") as f:
        if case.friction == "fvw":
            if fault is None:
                fault, fvw_scalars = C.fault_fvw(case)
            for n in C.FVW_FIELDS:
                f.write(np.asfortranarray(fault[n], dtype=np.float64).tobytes(order="F"))
            f.write(np.array([fvw_scalars[n] for n in C.FVW_SCALARS], dtype=np.float64).tobytes())
        else:
            if fault is None:
                fault = C.fault_sw(case, np.float32)
            for n in C.SW_FIELDS:
                f.write(np.asfortranarray(fault[n], dtype=np.float32).tobytes(order="F"))


def read_outputs(case: C.Case, wd: str) -> Dict[str, np.ndarray]:
    dt = np.float64 if _f64(case) else np.float32
    nxT, nzT = case.fault_shape
    res = {}
    names = ["sliprateX", "sliprateZ", "shearstressX", "shearstressZ"]
    if case.friction == "fvw":
        names.append("psi")
    for n in names:
        a = np.fromfile(os.path.join(wd, "result", f"{n}.res"), dtype=dt)
        res[n] = a.reshape(-1, nzT, nxT).transpose(0, 2, 1).astype(np.float64)
    fin = np.fromfile(os.path.join(wd, "result", "final.bin"), dtype=dt)
    fin = fin.reshape(7, case.nzt, case.nxt).transpose(0, 2, 1).astype(np.float64)
    # final.bin holds ruptime, rise(-ruptime applied), sliptime, slipX, slipZ,
    # schangeX-T0X, schangeZ-T0Z (the harness writes them after fd3d() returns)
    for q, n in enumerate(("ruptime", "rise", "sliptime", "slipX", "slipZ", "stressdropX",
                           "stressdropZ")):
        res[n] = fin[q]
    seis = []
    for s in range(len(case.stations)):
        rows = []
        with open(os.path.join(wd, "result", f"stan{s}.txt")) as f:
            for line in f:
                if line.strip():
                    rows.append([float(x) for x in line.split()])
        seis.append(rows)
    if seis:
        res["seis"] = np.asarray(seis, dtype=np.float64).transpose(1, 0, 2)
    return res


def _big_stack():
    """fd3d() keeps several (nxt, nzt) automatic arrays on the stack
    (fd3d_theo.f90:36,43,47): lift the soft stack limit to the hard one."""
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run(case: C.Case, fault=None, fvw_scalars=None, fast: bool = False, timeout: float = 3600,
        workdir: Optional[str] = None, keep_outputs: bool = True, cores: Optional[str] = None,
        rates: Optional[tuple] = None):
    """Run the reference; returns (outputs, wall seconds, loop CPU seconds).
    rates: (NL, NW, dL, dW, dtseis) for the inversion accumulators, which
    come back as outputs 'MSRX', 'MSRZ' (NL*NW x ntfd) and 'MomentRate'."""
    exe = variant(case, fast)
    if not os.access(exe, os.X_OK):
        raise FileNotFoundError(exe)
    wd = workdir or tempfile.mkdtemp(prefix="fd3dref_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    write_inputs(case, wd, fault, fvw_scalars)
    if rates is not None:
        with open(os.path.join(wd, "rates_in.txt"), "w") as f:
            f.write(" ".join(repr(x) for x in rates) + "\n")
    cmd = [exe]
    if cores is not None:
        cmd = ["taskset", "-c", cores, exe]
    t0 = time.perf_counter()
    p = subprocess.run(cmd, cwd=wd, capture_output=True, text=True, timeout=timeout, preexec_fn=_big_stack)
    wall = time.perf_counter() - t0
    if p.returncode != 0:
        raise RuntimeError(f"reference failed rc={p.returncode}\n{p.stdout[-2000:]}\n{p.stderr[-2000:]}")
    loop = None
    for line in p.stdout.splitlines():
        if "CPU TIME OF TIME LOOP" in line:
            loop = float(line.split(":")[-1])
    res = read_outputs(case, wd) if keep_outputs else {}
    if keep_outputs and rates is not None:
        dt = np.float64 if _f64(case) else np.float32
        r = np.fromfile(os.path.join(wd, "result", "rates.bin"), dtype=dt)
        nlw = int(rates[0]) * int(rates[1])
        n = case.ntfd
        res["MSRX"] = r[:nlw * n].reshape(nlw, n)
        res["MSRZ"] = r[nlw * n:2 * nlw * n].reshape(nlw, n)
        res["MomentRate"] = r[2 * nlw * n:2 * nlw * n + n]
    if workdir is None:
        subprocess.run(["rm", "-rf", wd])
    res["_stdout"] = p.stdout
    return res, wall, loop
