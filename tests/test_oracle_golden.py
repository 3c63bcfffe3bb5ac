"""The oracle against the golden fixtures written by the reference build.

Slip weakening is compared bit for bit (SHA-256 of every per-step output
slab, station samples and final arrays); rate-and-state against the
-fdefault-real-8 reference, also exactly (same glibc transcendentals)."""
import numpy as np
import pytest

import golden_lib as G
import oracle_lib as O

SW = ["sw_fs_tiny", "sw_free_tiny"]
FVW = ["fvw_free_tiny", "fvw_fs_tiny"]


@pytest.mark.parametrize("name", SW)
def test_oracle_sw_matches_reference_fixture(name):
    case = G.case_of(name)
    gold = G.load(name)
    orc = O.Oracle(case, O.F32)
    res = orc.step(case.ntfd)
    fin = orc.final()
    for o, n in enumerate(G.OUTS):
        t = G.first_hash_mismatch(res[n], gold["slab_sha256"], o)
        assert t is None, f"{n}: first differing step {t}"
        assert np.array_equal(res[n][-1].astype(np.float32), gold[f"last_{n}"])
    assert np.array_equal(res["seis"].astype(np.float32), gold["seis"])
    for n in G.FINALS:
        assert np.array_equal(fin[n].astype(np.float32), gold[f"final_{n}"]), n


def test_oracle_plain_workload_first_100_steps():
    """The 128x64x64 plain workload: the first 100 of its 500 steps (the GPU
    test checks all 500)."""
    name = "plain_sw_128x64x64"
    case = G.case_of(name)
    gold = G.load(name)
    orc = O.Oracle(case, O.F32)
    res = orc.step(100)
    for o, n in enumerate(G.OUTS):
        t = G.first_hash_mismatch(res[n], gold["slab_sha256"][:100], o)
        assert t is None, f"{n}: first differing step {t}"
    assert np.array_equal(res["seis"].astype(np.float32), gold["seis"][:100])


@pytest.mark.parametrize("name", FVW)
def test_oracle_fvw_matches_fp64_reference_fixture(name):
    case = G.case_of(name)
    gold = G.load(name)
    orc = O.Oracle(case, O.F64)
    res = orc.step(case.ntfd)
    fin = orc.final()
    assert orc.capped == 0
    xi, zk = gold["hist_x"], gold["hist_z"]
    for n in G.OUTS + ("psi",):
        h = res[n][:, xi][:, :, zk]
        err = np.max(np.abs(h - gold[f"hist_{n}"])) / max(np.max(np.abs(gold[f"hist_{n}"])), 1e-300)
        assert err <= 1e-13, (n, err)
    for n in G.FINALS:
        assert np.allclose(fin[n], gold[f"final_{n}"], rtol=1e-13, atol=0), n
    assert np.allclose(res["seis"], gold["seis"], rtol=1e-12, atol=1e-300)
