"""Full-size checks of the benchmark problems: the benchmarked path against
the oracle at the benchmark's size, and against the other launch forms.

The 1024x512x512 slip-weakening rupture (1064x532x552 cells with its
20-node PML) runs, as benchmarked, as one fused velocity + stress interior
sweep replayed from a captured step graph.  The separate sweeps run
kernel by kernel are bit-exact against the oracle on the small cases
(test_gpu_parity.py), and both forms keep the reference's per-point
arithmetic, so from the same random wavefield they must agree bit for
bit after a few steps, in every cell of every field and every output
slab.  The rate-and-state problem (512x256x256, free surface) is checked
the same way: its benchmarked path is the fused sweep with the top-plane
stresses after the surface's ghost velocities.

Directly against the oracle at full size (its 3D sweeps run on the host's
cores): the benchmarked 1024x512x512 path from a random wavefield, two
steps, every output slab, station and cell of the nine fields bit for bit
against the FP32 oracle; the 512x256x256 rate-and-state path from a
perturbed state, four steps, within 1e-9 relative L2 of the mixed oracle
(double transcendentals differ in the last ulp between device and host).
"""
import sys

import numpy as np
import pytest

from fd3d_tsn_amd import autotune, capi
from fd3d_tsn_amd import cases as C
from fd3d_tsn_amd import serve
import oracle_lib as O

pytestmark = pytest.mark.gpu

OUTS = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ", "psi")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_wavefield(solvers, case, seed, vel_scale, str_scale):
    """The same random wavefield (every cell non-zero) into each solver (or
    oracle), one field at a time."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (case.nxt, case.nyt, case.nzt)
    for n in capi.FIELDS:
        scale = vel_scale if n in ("u", "v", "w") else str_scale
        a = (rng.standard_normal(case.cells, dtype=np.float32) * np.float32(scale)).reshape(shape, order="F")
        for s in solvers:
            s.set_field(n, a)
        del a


def assert_same(a, b, case, steps):
    ra, rb = a.step(steps), b.step(steps)
    for n in OUTS:
        if n in ra:
            assert np.array_equal(ra[n], rb[n]), f"output {n}"
    assert np.array_equal(ra["seis"], rb["seis"]), "stations"
    for n in capi.FIELDS:
        fa, fb = a.field(n), b.field(n)
        bad = np.count_nonzero(fa != fb)
        assert bad == 0, f"field {n}: {bad} of {case.cells} cells differ"
        del fa, fb


def test_large_sw_fused_graph_equals_separate_kernel_by_kernel():
    case = C.c2_large_sw()
    say(f"[fullsize] {case.name} {case.nxt}x{case.nyt}x{case.nzt}: fused (graph) against separate (kernel by kernel)")
    with capi.Solver(case) as a, capi.Solver(case) as b:
        assert a.get_tuning()[0] & 4, "the library's default on this grid is the fused sweep"
        # 16 x 47 tiles, 94 per XCD: the default walks them in strips of 2
        assert a.get_tuning()[0] & 96 == 32, a.get_tuning()
        assert autotune.apply(a), "the size class is in the tuning cache"
        ta = a.get_tuning()
        assert ta[0] & 4, f"the fused sweep is the benchmarked path (tuning {ta})"
        b.set_tuning(1, 16, 32, 16, 32)
        b.set_profiling(True)
        random_wavefield((a, b), case, seed=41, vel_scale=1e-2, str_scale=1e6)
        assert_same(a, b, case, steps=3)


def test_fvw_fused_graph_equals_separate_kernel_by_kernel():
    case = C.c3_fvw()
    say(f"[fullsize] {case.name} {case.nxt}x{case.nyt}x{case.nzt}: fused with a free surface (graph) against "
        f"separate (kernel by kernel)")
    with capi.Solver(case) as a, capi.Solver(case) as b:
        assert autotune.apply(a), "the size class is in the tuning cache"
        ta = a.get_tuning()
        assert ta[0] & 4, f"the fused sweep is the benchmarked path (tuning {ta})"
        b.set_tuning(1, 16, 32, 4, 64)
        b.set_profiling(True)
        # small perturbations of the initial state: every cell non-zero, the
        # fault far from Newton trouble
        random_wavefield((a, b), case, seed=43, vel_scale=1e-6, str_scale=1e3)
        assert_same(a, b, case, steps=4)


def test_large_sw_graph_against_oracle():
    case = C.c2_large_sw()
    steps = 2
    say(f"[fullsize] {case.name} {case.nxt}x{case.nyt}x{case.nzt}: benchmarked path against the FP32 oracle, "
        f"{steps} steps from a random wavefield")
    orc = O.Oracle(case, O.F32)
    with capi.Solver(case) as a:
        assert autotune.apply(a) and a.get_tuning()[0] & 4, "the fused sweep is the benchmarked path"
        random_wavefield((a, orc), case, seed=45, vel_scale=1e-2, str_scale=1e6)
        say("[fullsize] oracle steps")
        ro = orc.step(steps)
        rg = a.step(steps)
        for n in OUTS[:4]:
            assert np.array_equal(rg[n].astype(np.float64), ro[n]), f"output {n}"
        assert np.array_equal(rg["seis"].astype(np.float64), ro["seis"]), "stations"
        for n in capi.FIELDS:
            fg = a.field(n)
            fo = orc.field(n).astype(np.float32)
            bad = np.count_nonzero(fg != fo)
            assert bad == 0, f"field {n}: {bad} of {case.cells} cells differ from the oracle"
            del fg, fo
            say(f"[fullsize] field {n}: bit-exact")


def test_fvw_graph_against_mixed_oracle():
    case = C.c3_fvw()
    steps = 4
    say(f"[fullsize] {case.name} {case.nxt}x{case.nyt}x{case.nzt}: benchmarked path against the mixed oracle, "
        f"{steps} steps from a perturbed state")
    fault, scal = C.fault_fvw(case)
    orc = O.Oracle(case, O.MIXED, fault=fault, fvw_scalars=scal)
    with capi.Solver(case, fault=fault, fvw_scalars=scal) as a:
        assert autotune.apply(a) and a.get_tuning()[0] & 4, "the fused sweep is the benchmarked path"
        random_wavefield((a, orc), case, seed=47, vel_scale=1e-6, str_scale=1e3)
        ro = orc.step(steps)
        rg = a.step(steps)
        errs = {n: serve.rel_l2(rg[n], ro[n].astype(np.float32)) for n in OUTS}
        errs["seis"] = serve.rel_l2(rg["seis"], ro["seis"])
        for n in capi.FIELDS:
            errs[n] = serve.rel_l2(a.field(n), orc.field(n).astype(np.float32))
        say(f"[fullsize] worst relative L2 against the mixed oracle {max(errs.values()):.2e} "
            f"({max(errs, key=errs.get)})")
        for n, e in errs.items():
            assert e <= 1e-9, f"{n}: relative L2 {e:.3e} against the mixed oracle"
