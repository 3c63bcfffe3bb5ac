"""Host-side input logic: grid conventions, input text, material, fields."""
import numpy as np
import pytest

from fd3d_tsn_amd import cases as C
from fd3d_tsn_amd.fields import random_field


def test_grid_conventions():
    c = C.c1_plain_sw()
    assert (c.nxt, c.nyt, c.nzt) == (148, 74, 84)          # fd3d_init.f90:167-169 with FSPACE
    assert c.cells == 919968
    assert c.fault_shape == (128, 64)
    free = c.replace(fspace=False)
    assert free.nfs == 2 and free.nzt == 64 + 10 + 2
    big = C.c2_large_sw()
    assert (big.nxt, big.nyt, big.nzt) == (1064, 532, 552)
    assert abs(big.cfl() - 0.24) < 1e-6


def test_inputfd3d_text_reads_back():
    c = C.tiny("sw", True).replace(stations=[(3, 4, 5)])
    lines = c.inputfd3d_text().splitlines()
    assert lines[0].split() == [str(c.nxtT), str(c.nytT), str(c.nztT)]
    assert np.float32(lines[1]) == np.float32(c.dh) and np.float32(lines[3]) == np.float32(c.dt)
    assert lines[5].split()[0] == str(c.nabc)
    assert lines[7] == "1" and lines[8] == "3 4 5" and lines[9] == "0"


def test_material_matches_readcrustalmodel_formula():
    c = C.tiny("sw", True)
    lam, mu, d = C.material(c, np.float32)
    vp = np.float32("6.000") * np.float32(1e3)
    vs = np.float32("3.464") * np.float32(1e3)
    rho = np.float32("2.670") * np.float32(1e3)
    assert mu[0, 0, 0] == rho * (vs * vs)
    assert lam[0, 0, 0] == rho * (vp * vp - np.float32(2) * (vs * vs))
    assert d.flags.f_contiguous and d.shape == (c.nxt, c.nyt, c.nzt)


@pytest.mark.parametrize("kind", ["gaussian", "vonkarman"])
def test_random_field_statistics_and_seeding(kind):
    f1 = random_field(256, 128, 100.0, 2000.0, 3e6, seed=4, kind=kind)
    f2 = random_field(256, 128, 100.0, 2000.0, 3e6, seed=4, kind=kind)
    f3 = random_field(256, 128, 100.0, 2000.0, 3e6, seed=5, kind=kind)
    assert np.array_equal(f1, f2) and not np.array_equal(f1, f3)
    assert abs(f1.mean()) < 1e-6 * 3e6 and abs(f1.std() - 3e6) < 1e-6 * 3e6
    # correlated: neighbouring samples are strongly correlated at 2 km / 100 m
    r = np.corrcoef(f1[:-1].ravel(), f1[1:].ravel())[0, 1]
    assert r > 0.8


def test_sw_fault_setup():
    c = C.c1_plain_sw()
    f = C.fault_sw(c)
    peak, dyn = f["peak_xz"], f["dyn_xz"]
    inside = peak < 1e9
    assert np.isclose(peak[inside].max(), np.float32(120e6 * 0.677))
    assert np.all(dyn == np.float32(120e6 * 0.525))
    assert inside.sum() == 100 * 50   # 20 x 10 km on a 200 m grid, box centred between nodes
    assert np.isclose(f["striniX"].max(), 81.6e6)


def test_fvw_setup_is_steady_state_consistent():
    c = C.tiny("fvw", False)
    arr, sc = C.fault_fvw(c)
    a, psi = arr["a"], arr["psi"]
    assert a.min() >= c.a0 - 1e-15 and a.max() <= 2 * c.a0 + 1e-15
    # psi from the friction law at the initial slip rate reproduces the prestress:
    # tau = Sn * a * asinh(V/(2 V0) exp(psi/a)) with V = 2*vini
    tau = sc["Sn"] * a * np.arcsinh((2 * c.vini) / (2 * c.v0) * np.exp(psi / a))
    assert np.allclose(tau, c.tini, rtol=1e-9)


def test_batched_workload_shape():
    cs = C.c4_batched(64, seed=4)
    assert len(cs) == 64 and len({c.name for c in cs}) == 64
    for c in cs:
        assert 96 <= c.nxtT <= 320 and 96 <= c.nytT <= 320 and 96 <= c.nztT <= 320
        assert c.dh == 100.0 and c.ntfd == 1500 and c.nabc == 10 and c.fspace
        assert abs(c.cfl() - 0.24) < 1e-6
        assert 0.2 <= c.dc <= 0.8 and 0.6 <= c.mus <= 0.75
        f = C.fault_sw(c.replace(nxtT=c.nxtT, nytT=8))
        # nucleation patch above static strength, prestress below it on average
        assert f["striniX"].max() >= c.sn * c.mus
        assert np.mean(f["striniX"]) < c.sn * c.mus


def test_request_stream_shape():
    from fd3d_tsn_amd import serve
    reqs = serve.c5_requests(200, seed=5)
    fr = np.mean([r.case.friction == "fvw" for r in reqs])
    assert 0.2 < fr < 0.4
    gaps = np.diff([r.arrival for r in reqs])
    assert 0.03 < gaps.mean() < 0.07
    for r in reqs:
        c = r.case
        assert 64 <= c.nxtT <= 384 and 64 <= c.nytT <= 192 and 64 <= c.nztT <= 192
        assert 800 <= c.ntfd <= 3000 and 100 <= c.dh <= 200
        assert abs(c.cfl() - 0.24) < 1e-3
