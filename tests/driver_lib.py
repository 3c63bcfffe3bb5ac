"""Running the Fortran driver (fd3d_tsn_amd/lib/fd3d_tsn_driver) in a scratch directory."""
import os
import shutil
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fd3d_tsn_amd", "lib", "fd3d_tsn_driver")
EXAMPLES = os.path.join(ROOT, "tests", "golden", "examples")


def scratch():
    return tempfile.mkdtemp(prefix="fd3dtsn_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)


def copy_example(name, wd):
    for f in os.listdir(os.path.join(EXAMPLES, name)):
        shutil.copy(os.path.join(EXAMPLES, name, f), os.path.join(wd, f))


def run_driver(wd, *args, timeout=600):
    os.makedirs(os.path.join(wd, "result"), exist_ok=True)
    p = subprocess.run([DRIVER, *args], cwd=wd, capture_output=True, text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError(f"driver failed rc={p.returncode}\n{p.stdout[-3000:]}\n{p.stderr[-3000:]}")
    return p.stdout
