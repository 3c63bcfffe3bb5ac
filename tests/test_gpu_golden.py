"""The HIP path against the golden fixtures of the reference build, without
the oracle in between: bit-exact for slip weakening (every per-step slab of
all 500 steps of the 128x64x64 plain workload), and for rate-and-state the
stated error of the FP32-wavefield / FP64-friction path against the FP64
reference."""
import sys

import numpy as np
import pytest

from fd3d_tsn_amd import capi
from fd3d_tsn_amd import cases as C

import golden_lib as G

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["sw_fs_tiny", "sw_free_tiny", "plain_sw_128x64x64"])
def test_gpu_sw_bit_exact_against_reference(name):
    case = G.case_of(name)
    gold = G.load(name)
    say(f"[golden] {name} {case.nxt}x{case.nyt}x{case.nzt} {case.ntfd} steps")
    with capi.Solver(case) as s:
        res = s.step(case.ntfd)
        fin = s.final()
    for o, n in enumerate(G.OUTS):
        t = G.first_hash_mismatch(res[n], gold["slab_sha256"], o)
        assert t is None, f"{n}: first differing step {t}"
    assert np.array_equal(res["seis"], gold["seis"])
    for n in G.FINALS:
        assert np.array_equal(fin[n], gold[f"final_{n}"]), n


@pytest.mark.parametrize("name", ["fvw_free_tiny", "fvw_fs_tiny"])
def test_gpu_fvw_against_fp64_reference(name):
    case = G.case_of(name)
    gold = G.load(name)
    fault, scal = C.fault_fvw(case)
    say(f"[golden] {name} {case.nxt}x{case.nyt}x{case.nzt} {case.ntfd} steps")
    with capi.Solver(case, fault=fault, fvw_scalars=scal) as s:
        res = s.step(case.ntfd)
        fin = s.final()
    xi, zk = gold["hist_x"], gold["hist_z"]
    for n in G.OUTS + ("psi",):
        h = res[n][:, xi][:, :, zk].astype(np.float64)
        g = gold[f"hist_{n}"]
        err = np.sqrt(np.sum((h - g) ** 2) / np.sum(g * g))
        say(f"  {n:13s} relative L2 vs FP64 reference {err:.2e}")
        assert err <= 1e-4, (n, err)
    rt, rt_ref = fin["ruptime"].astype(np.float64), gold["final_ruptime"]
    assert np.all(np.abs(rt - rt_ref) <= case.dt * 1.01)
