#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ from the reference build.

Runs oracle/_ref/fd3d_ref_* (the unmodified FD3D_TSN Fortran sources built
with flang by ``make -C oracle ref`` and driven by oracle/ref_harness.f90)
on small seeded problems and stores what the reference wrote:

* slip weakening (FP32 build): SHA-256 of every per-step output slab of
  result/sliprateX/Z.res and shearstressX/Z.res, the last slab of each,
  the station samples and the final 2D arrays;
* rate-and-state (FP64 build): the same final arrays and stations, plus the
  time histories of every output at a sub-sampled set of fault nodes.

Usage:  python tests/golden/make_golden.py          (needs oracle/_ref built)
        python tests/golden/make_golden.py tpv104   (the rate-and-state example only)
"""
import dataclasses
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from fd3d_tsn_amd import cases as C  # noqa: E402
import refrun  # noqa: E402


def spread(case):
    nx, ny, nz, nfs = case.nxt, case.nyt, case.nzt, case.nfs
    st = [(nx // 2, ny // 2, nz // 2), (nx // 2 + 1, ny - 1, nz // 2), (nx // 2 + 2, ny, nz // 2 + 1),
          (2, ny - 2, nz // 2), (nx - 1, ny - 3, nz // 2 - 1), (nx // 2, 2, nz // 2), (nx // 2, ny // 2, 2),
          (nx // 2 - 1, ny, nz - nfs), (2, 2, 2), (nx - 2, ny, nz - 1)]
    return case.replace(stations=st)


def golden_cases():
    return {
        "sw_fs_tiny": spread(C.tiny("sw", True, ntfd=120, seed=11)),
        "sw_free_tiny": spread(C.tiny("sw", False, ntfd=120, seed=12)),
        "fvw_free_tiny": spread(C.tiny("fvw", False, ntfd=150, seed=13)),
        "fvw_fs_tiny": spread(C.tiny("fvw", True, ntfd=150, seed=14)),
        "plain_sw_128x64x64": spread(C.c1_plain_sw()),
    }


OUTS = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ")
FINALS = ("ruptime", "rise", "sliptime", "slipX", "slipZ", "stressdropX", "stressdropZ")


def slab_hash(a32: np.ndarray) -> bytes:
    """SHA-256 of one slab in the byte order of the .res stream (x fastest)."""
    return hashlib.sha256(np.ascontiguousarray(a32.T, dtype=np.float32).tobytes()).digest()


def sub_nodes(nxT, nzT):
    return np.arange(1, nxT, 3), np.arange(1, nzT, 3)


DRIVER_FILES = ["sliprateX.res", "sliprateZ.res", "shearstressX.res", "shearstressZ.res", "ruptime.res",
                "risetime.res", "slipX.res", "slipZ.res", "stressdropX.res", "stressdropZ.res", "contour.res",
                "czone.res", "rvel.res", "vmodel.inp", "friction.inp"]


def driver_cases():
    """Small problems whose whole result/ directory the driver must reproduce."""
    return {
        "files_sw_fs": spread(C.tiny("sw", True, ntfd=90, seed=21)).replace(stations=[(20, 10, 12), (26, 22, 15)]),
        "files_sw_free": spread(C.tiny("sw", False, ntfd=90, seed=22)).replace(stations=[(21, 9, 20)]),
    }


def driver_fixtures(manifest):
    import shutil
    import tempfile
    for name, case in driver_cases().items():
        print(f"[golden] {name}: reference result/ files", flush=True)
        wd = tempfile.mkdtemp(prefix="fd3dgold_")
        refrun.run(case, workdir=wd)
        files = DRIVER_FILES + [f"stan{s}.txt" for s in range(len(case.stations))]
        digests = {f: hashlib.sha256(open(os.path.join(wd, "result", f), "rb").read()).hexdigest() for f in files}
        inputs = os.path.join(HERE, name)
        os.makedirs(inputs, exist_ok=True)
        for f in ("inputfd3d.dat", "crustal.dat", "fault_in.bin"):
            shutil.copy(os.path.join(wd, f), os.path.join(inputs, f))
        shutil.rmtree(wd)
        cfg = dataclasses.asdict(case)
        cfg["stations"] = [list(s) for s in case.stations]
        manifest[name] = {"case": cfg, "reference_build": os.path.basename(refrun.variant(case)),
                          "inputs": name, "result_sha256": digests}


EXAMPLE_RUNS = {
    # fixture: (example, reference variant, harness setup, steps)
    "files_tpv5_200": ("TPV5", "fd3d_ref_sw_free", "tpv5", 200),
    "files_amatrice_200": ("Amatrice2016", "fd3d_ref_sw_free", "inversion", 200),
}


def example_fixtures(manifest):
    """The reference's own example inputs (tests/golden/examples/), time loop
    cut to a few hundred steps: the result/ files the driver must reproduce."""
    import shutil
    import subprocess
    import tempfile
    for name, (example, variant, setup, steps) in EXAMPLE_RUNS.items():
        print(f"[golden] {name}: {example} ({setup}), {steps} steps", flush=True)
        wd = tempfile.mkdtemp(prefix="fd3dgold_")
        src = os.path.join(HERE, "examples", example)
        for f in os.listdir(src):
            shutil.copy(os.path.join(src, f), os.path.join(wd, f))
        lines = open(os.path.join(wd, "inputfd3d.dat")).read().split("\n")
        lines[2] = f"{steps}"
        open(os.path.join(wd, "inputfd3d.dat"), "w").write("\n".join(lines))
        open(os.path.join(wd, "harness_setup.txt"), "w").write(setup + "\n")
        os.makedirs(os.path.join(wd, "result"))
        p = subprocess.run([os.path.join(refrun.REF_DIR, variant)], cwd=wd, capture_output=True, text=True)
        assert p.returncode == 0, p.stdout[-2000:]
        nsta = int(lines[7].split()[0])
        files = DRIVER_FILES + [f"stan{s}.txt" for s in range(nsta)]
        digests = {f: hashlib.sha256(open(os.path.join(wd, "result", f), "rb").read()).hexdigest() for f in files}
        manifest[name] = {"example": example, "setup": setup, "steps": steps, "reference_build": variant,
                          "result_sha256": digests}
        shutil.rmtree(wd)


# Rate-and-state example: the reference's -fdefault-real-8 build (the -r8 of
# examples/TPV104/compile.sh:13) on examples/TPV104 with the time loop cut
# to 100 steps.  Its .res streams hold 8-byte reals; the fixture keeps each
# per-step stream (psi.res included, fd3d_theo.f90:546-553) at every 10th
# fault node in x and z, and the final slip files (text) likewise, as
# float32 (rounding 6e-8, far below the 1e-4 bar the driver is held to).
FVW_EXAMPLE = ("files_tpv104_100", "TPV104", "fd3d_ref_fvw_free", "tpv104", 100, 10)
FVW_STREAMS = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ", "psi")


def read_text_column(path):
    return np.asarray([float(x) for x in open(path).read().split()], dtype=np.float64)


def fvw_example_fixture(manifest):
    import shutil
    import subprocess
    import tempfile
    name, example, variant, setup, steps, every = FVW_EXAMPLE
    print(f"[golden] {name}: {example} ({setup}, {variant}), {steps} steps", flush=True)
    wd = tempfile.mkdtemp(prefix="fd3dgold_")
    src = os.path.join(HERE, "examples", example)
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), os.path.join(wd, f))
    lines = open(os.path.join(wd, "inputfd3d.dat")).read().split("\n")
    lines[2] = f"{steps}"
    open(os.path.join(wd, "inputfd3d.dat"), "w").write("\n".join(lines))
    open(os.path.join(wd, "harness_setup.txt"), "w").write(setup + "\n")
    os.makedirs(os.path.join(wd, "result"))
    p = subprocess.run([os.path.join(refrun.REF_DIR, variant)], cwd=wd, capture_output=True, text=True,
                       preexec_fn=refrun._big_stack)
    assert p.returncode == 0, p.stdout[-2000:]
    nxT, nyT, nzT = (int(v) for v in lines[0].split()[:3])
    data = {}
    for n in FVW_STREAMS:
        a = np.fromfile(os.path.join(wd, "result", f"{n}.res"), dtype=np.float64).reshape(steps, nzT, nxT)
        data[f"hist_{n}"] = a[:, ::every, ::every].astype(np.float32)
    for n in ("slipX", "slipZ"):
        a = read_text_column(os.path.join(wd, "result", f"{n}.res"))
        data[f"text_{n}"] = a.astype(np.float32)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **data)
    shutil.rmtree(wd)
    manifest[name] = {"example": example, "setup": setup, "steps": steps, "reference_build": variant,
                      "precision": "float64 (-fdefault-real-8)", "node_stride": every, "file": os.path.basename(path),
                      "sha256": hashlib.sha256(open(path, "rb").read()).hexdigest()}


def rates_cases():
    """Inversion accumulators (MSRX, MSRZ, MomentRate) of the reference's time
    loop, with a 4 x 2 patch grid and 0.04 s bins set through the harness."""
    c = C.tiny("sw", True, ntfd=120, seed=61)
    nxT, nzT = c.fault_shape
    return {"rates_sw_fs": (c, (4, 2, nxT * c.dh / 4, nzT * c.dh / 2, 0.04))}


def rates_fixtures(manifest):
    for name, (case, rates) in rates_cases().items():
        print(f"[golden] {name}: inversion accumulators {rates}", flush=True)
        res, _w, _l = refrun.run(case, rates=rates)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, MSRX=res["MSRX"].astype(np.float32), MSRZ=res["MSRZ"].astype(np.float32),
                            MomentRate=res["MomentRate"].astype(np.float32))
        cfg = dataclasses.asdict(case)
        cfg["stations"] = [list(s) for s in case.stations]
        manifest[name] = {"case": cfg, "rates": list(rates), "file": os.path.basename(path),
                          "reference_build": os.path.basename(refrun.variant(case)),
                          "sha256": hashlib.sha256(open(path, "rb").read()).hexdigest()}


def main():
    if sys.argv[1:] == ["tpv104"]:
        # only the rate-and-state example fixture, into the existing manifest
        path = os.path.join(HERE, "manifest.json")
        manifest = json.load(open(path))
        fvw_example_fixture(manifest)
        with open(path, "w") as f:
            json.dump(manifest, f, indent=1, sort_keys=True)
        return
    manifest = {}
    fvw_example_fixture(manifest)
    rates_fixtures(manifest)
    example_fixtures(manifest)
    driver_fixtures(manifest)
    for name, case in golden_cases().items():
        print(f"[golden] {name}: {case.nxt}x{case.nyt}x{case.nzt}, {case.ntfd} steps", flush=True)
        res, wall, loop = refrun.run(case)
        fvw = case.friction == "fvw"
        data = {}
        for n in FINALS:
            data[f"final_{n}"] = res[n].astype(np.float64 if fvw else np.float32)
        data["seis"] = res["seis"].astype(np.float64 if fvw else np.float32)
        if not fvw:
            hashes = np.zeros((case.ntfd, len(OUTS), 32), dtype=np.uint8)
            for o, n in enumerate(OUTS):
                a = res[n].astype(np.float32)
                for t in range(case.ntfd):
                    hashes[t, o] = np.frombuffer(slab_hash(a[t]), dtype=np.uint8)
                data[f"last_{n}"] = a[-1]
            data["slab_sha256"] = hashes
        else:
            xi, zk = sub_nodes(*case.fault_shape)
            for n in OUTS + ("psi",):
                data[f"hist_{n}"] = res[n][:, xi][:, :, zk]
            data["hist_x"] = xi
            data["hist_z"] = zk
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **data)
        cfg = dataclasses.asdict(case)
        cfg["stations"] = [list(s) for s in case.stations]
        manifest[name] = {
            "case": cfg,
            "reference_build": os.path.basename(refrun.variant(case)),
            "precision": "float64 (-fdefault-real-8)" if fvw else "float32",
            "flags": "flang -cpp -O2 -ffp-contract=off" + (" -DFVW -DTPV104" if fvw else "") +
                     (" -DFSPACE" if case.fspace else ""),
            "file": os.path.basename(path),
            "sha256": hashlib.sha256(open(path, "rb").read()).hexdigest(),
            "reference_loop_seconds": loop,
        }
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("[golden] wrote", ", ".join(manifest))


if __name__ == "__main__":
    main()
