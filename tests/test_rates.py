"""Inversion accumulators of the time loop (fd3d_theo.f90:560-578): the
mean slip rate over NL x NW fault patches per dtseis bin (MSRX, MSRZ) and
the moment rate per bin.

The reference fixture (tests/golden/rates_sw_fs.npz) comes from the
reference build with a 4 x 2 patch grid and 0.04 s bins.  A float32
restatement of the accumulation, fed with the oracle's output slabs (which
are bit-identical to the reference's), must reproduce it exactly: this pins
the reading of the bin index, patch bounds (including the reference's
one-row overhang past the fault at the last patch row) and the
float/double mix of the update.  The GPU accumulates on the device and is
compared with the same fixture: patches exactly, the moment rate (a
fixed-order double sum on the device, a node-by-node float sum in the
reference) to 1e-6 of the largest bin.
"""
import json
import os

import numpy as np
import pytest

from fd3d_tsn_amd import capi
from fd3d_tsn_amd import cases as C

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def fixture():
    man = json.load(open(os.path.join(GOLD, "manifest.json")))["rates_sw_fs"]
    case = C.Case(**{**man["case"], "stations": [tuple(s) for s in man["case"]["stations"]]})
    return case, tuple(man["rates"]), np.load(os.path.join(GOLD, man["file"]))


def restate(case, rates, srx, srz):
    """float32 restatement of fd3d_theo.f90:560-578 over per-step slabs."""
    f = np.float32
    NL, NW, dL, dW, dtseis = rates
    dL, dW, dtseis, dt, dh = f(dL), f(dW), f(dtseis), f(case.dt), f(case.dh)
    nabc, nfs, nxt, nzt = case.nabc, case.nfs, case.nxt, case.nzt
    n = srx.shape[0]
    msrx = np.zeros((NL * NW, n), np.float32)
    msrz = np.zeros((NL * NW, n), np.float32)
    mr = np.zeros(n, np.float32)
    mu = C.material_scalars(case, np.float32)[1]
    per = dtseis / dt

    def val(a, t, i, k):   # sliprateout(i, k), zero off the fault slab
        if nabc + 1 <= i <= nxt - nabc and nabc + 1 <= k <= nzt - nfs:
            return a[t, i - nabc - 1, k - nabc - 1]
        return f(0.0)

    for t in range(n):
        it = t + 1
        kb = int((f(it - 1) * dt) / dtseis) + 1
        for j in range(1, NW + 1):
            jto = max(1, int((dW / dh) * f(j))) + 1 + nabc
            jfrom = min(jto, int((dW / dh) * f(j - 1)) + 1) + 1 + nabc
            for i in range(1, NL + 1):
                ifrom = int((dL / dh) * f(i - 1)) + 1 + nabc
                ito = int((dL / dh) * f(i)) + nabc
                sx = f(0.0)
                sz = f(0.0)
                for k in range(jfrom, jto + 1):
                    for ii in range(ifrom, ito + 1):
                        sx = f(sx + val(srx, t, ii, k))
                        sz = f(sz + val(srz, t, ii, k))
                div = np.float64(f((ito - ifrom + 1) * (jto - jfrom + 1)) * per)
                p = (j - 1) * NL + i - 1
                msrx[p, kb - 1] = f(np.float64(msrx[p, kb - 1]) + np.float64(sx) / div)
                msrz[p, kb - 1] = f(np.float64(msrz[p, kb - 1]) + np.float64(sz) / div)
        acc = mr[kb - 1]
        for k in range(nabc + 1, nzt - nfs + 1):
            for i in range(nabc + 1, nxt - nabc + 1):
                sx, sz = val(srx, t, i, k), val(srz, t, i, k)
                acc = f(acc + f(f(f(f(np.sqrt(f(sx * sx + sz * sz)) * mu) * dh) * dh) / per))
        mr[kb - 1] = acc
    return msrx, msrz, mr


def test_restatement_equals_reference_fixture():
    case, rates, gold = fixture()
    res = O.Oracle(case, O.F32).step(case.ntfd)
    msrx, msrz, mr = restate(case, rates, res["sliprateX"].astype(np.float32), res["sliprateZ"].astype(np.float32))
    assert np.array_equal(msrx, gold["MSRX"])
    assert np.array_equal(msrz, gold["MSRZ"])
    assert np.array_equal(mr, gold["MomentRate"])
    assert np.count_nonzero(gold["MomentRate"]) > 10 and np.count_nonzero(gold["MSRX"]) > 10


@pytest.mark.gpu
def test_gpu_bins_equal_reference():
    case, rates, gold = fixture()
    NL, NW, dL, dW, dtseis = rates
    with capi.Solver(case) as s:
        s.set_source_bins(NL, NW, dL, dW, dtseis, case.ntfd)
        s.step(case.ntfd, outputs=False)
        got = s.source_bins()
    assert np.array_equal(got["MSRX"], gold["MSRX"])
    assert np.array_equal(got["MSRZ"], gold["MSRZ"])
    scale = np.abs(gold["MomentRate"]).max()
    assert np.max(np.abs(got["MomentRate"].astype(np.float64) - gold["MomentRate"])) <= 1e-6 * scale


@pytest.mark.gpu
def test_gpu_bins_in_a_batch_equal_alone():
    case, rates, gold = fixture()
    NL, NW, dL, dW, dtseis = rates
    other = C.tiny("sw", False, nabc=5, n=(37, 13, 19), seed=62, ntfd=case.ntfd)
    with capi.Solver(case) as s, capi.Solver(other) as o, capi.Batch(2, tz=8) as b:
        s.set_source_bins(NL, NW, dL, dW, dtseis, case.ntfd)
        b.add(s)
        b.add(o)
        for _ in range(case.ntfd // 24):
            b.step(24)
            b.sync()
        got = s.source_bins()
    with capi.Solver(case) as s:
        s.set_source_bins(NL, NW, dL, dW, dtseis, case.ntfd)
        s.step(case.ntfd, outputs=False)
        alone = s.source_bins()
    for k in ("MSRX", "MSRZ", "MomentRate"):
        assert np.array_equal(got[k], alone[k]), k
    assert np.array_equal(got["MSRX"], gold["MSRX"])
