"""HIP path against the oracle on the same seeded inputs.

Slip weakening: bit-exact (every output slab, station sample, final 2D
array and the full 3D wavefield), because the kernels keep the
reference's per-point arithmetic.  Rate-and-state: the wavefield is FP32
and the Newton solve / state update FP64, like the oracle's mixed mode;
they differ only through the device's and glibc's transcendental
functions, so the bar is relative L2 <= 1e-9 against the mixed oracle.
Against the all-FP64 oracle (the reference built -fdefault-real-8) the
FP32 wavefield gives relative L2 errors around 1e-6; the bar there is
north_star's 1e-4 and rupture time within one dt.
"""
import sys

import numpy as np
import pytest

from fd3d_tsn_amd import capi
from fd3d_tsn_amd import cases as C

import oracle_lib as O

pytestmark = pytest.mark.gpu

OUTS = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def spread_stations(case):
    """Stations in the interior, the near-fault rows, the fault plane and every PML region."""
    nx, ny, nz, nb, nfs = case.nxt, case.nyt, case.nzt, case.nabc, case.nfs
    st = [(nx // 2, ny // 2, nz // 2), (nx // 2 + 1, ny - 1, nz // 2), (nx // 2 + 2, ny, nz // 2 + 1),
          (2, ny - 2, nz // 2), (nx - 1, ny - 3, nz // 2 - 1), (nx // 2, 2, nz // 2), (nx // 2, ny // 2, 2),
          (nx // 2 - 1, ny, nz - nfs), (2, 2, 2), (nx - 2, ny, nz - 1)]
    return case.replace(stations=st)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / den) if den > 0 else float(np.sqrt(np.sum(a * a)))


SW_CASES = [
    ("sw_fs", dict(friction="sw", fspace=True)),
    ("sw_free", dict(friction="sw", fspace=False)),
    ("sw_fs_odd", dict(friction="sw", fspace=True, nabc=5, n=(37, 13, 19), seed=5)),
    ("sw_free_wide", dict(friction="sw", fspace=False, nabc=4, n=(150, 9, 14), seed=7)),
    ("sw_fs_tallz", dict(friction="sw", fspace=True, nabc=4, n=(24, 10, 150), seed=9)),
]


@pytest.mark.parametrize("name,kw", SW_CASES, ids=[c[0] for c in SW_CASES])
def test_sw_bit_exact(name, kw):
    nsteps = 80
    case = spread_stations(C.tiny(ntfd=nsteps, **kw))
    say(f"[parity] {name} grid {case.nxt}x{case.nyt}x{case.nzt} steps {nsteps}")
    orc = O.Oracle(case, O.F32)
    ro = orc.step(nsteps)
    fo = orc.final()
    with capi.Solver(case) as s:
        rg = s.step(nsteps)
        fg = s.final()
        fields = {n: s.field(n) for n in capi.FIELDS}
    for n in OUTS:
        bad = np.argwhere(rg[n].astype(np.float64) != ro[n])
        assert bad.size == 0, f"{n}: {len(bad)} mismatches, first {bad[0].tolist()}"
    assert np.array_equal(rg["seis"].astype(np.float64), ro["seis"])
    for n in capi.FINAL_NAMES:
        assert np.array_equal(fg[n].astype(np.float64), fo[n]), n
    for n in capi.FIELDS:
        ref = orc.field(n)
        bad = np.argwhere(fields[n].astype(np.float64) != ref)
        assert bad.size == 0, f"field {n}: {len(bad)} mismatches, first (0-based) {bad[0].tolist()}"


@pytest.mark.parametrize("fspace", [True, False], ids=["fs", "free"])
def test_random_wavefield_one_step_bit_exact(fspace):
    """Start both paths from the same random wavefield: every kernel in every
    region then acts on non-zero data from the first step."""
    case = C.tiny(friction="sw", fspace=fspace, nabc=5, n=(30, 12, 16), ntfd=3, seed=21)
    say(f"[parity] random wavefield fspace={fspace} grid {case.nxt}x{case.nyt}x{case.nzt}")
    rng = np.random.Generator(np.random.PCG64(99))
    init = {n: (rng.standard_normal((case.nxt, case.nyt, case.nzt)) *
                (1e6 if n in ("xx", "yy", "zz", "xy", "yz", "xz") else 1e-2)).astype(np.float32)
            for n in capi.FIELDS}
    orc = O.Oracle(case, O.F32)
    for n, a in init.items():
        orc.set_field(n, a)
    ro = orc.step(3)
    with capi.Solver(case) as s:
        for n, a in init.items():
            s.set_field(n, a)
        rg = s.step(3)
        fields = {n: s.field(n) for n in capi.FIELDS}
    for n in OUTS:
        assert np.array_equal(rg[n].astype(np.float64), ro[n]), n
    for n in capi.FIELDS:
        ref = orc.field(n)
        bad = np.argwhere(fields[n].astype(np.float64) != ref)
        assert bad.size == 0, f"field {n}: {len(bad)} mismatches, first (0-based) {bad[0].tolist()}"


FVW_CASES = [("fvw_free", dict(friction="fvw", fspace=False)),
             ("fvw_fs", dict(friction="fvw", fspace=True))]


@pytest.mark.parametrize("name,kw", FVW_CASES, ids=[c[0] for c in FVW_CASES])
def test_fvw_against_mixed_and_fp64(name, kw):
    nsteps = 150
    case = spread_stations(C.tiny(ntfd=nsteps, **kw))
    say(f"[parity] {name} grid {case.nxt}x{case.nyt}x{case.nzt} steps {nsteps}")
    fault, scal = C.fault_fvw(case)
    mixed = O.Oracle(case, O.MIXED, fault=fault, fvw_scalars=scal)
    rm = mixed.step(nsteps)
    fm = mixed.final()
    f64 = O.Oracle(case, O.F64, fault=fault, fvw_scalars=scal)
    r64 = f64.step(nsteps)
    f64f = f64.final()
    assert mixed.capped == 0 and f64.capped == 0
    with capi.Solver(case, fault=fault, fvw_scalars=scal) as s:
        rg = s.step(nsteps)
        fg = s.final()
    for n in OUTS + ("psi",):
        # the slabs are float32 streams, as the reference's default-REAL .res files
        e_mixed = rel_l2(rg[n], rm[n].astype(np.float32))
        e_64 = rel_l2(rg[n], r64[n])
        say(f"  {n:13s} relL2 vs mixed {e_mixed:.2e}  vs fp64 {e_64:.2e}")
        assert e_mixed <= 1e-9, (n, e_mixed)
        assert e_64 <= 1e-4, (n, e_64)
    dt = case.dt
    rt_g, rt_m, rt_64 = fg["ruptime"], fm["ruptime"], f64f["ruptime"]
    assert np.all(np.abs(rt_g - rt_m) <= 1.01 * dt)
    assert np.all(np.abs(rt_g - rt_64) <= 1.01 * dt)
    assert (rt_g < 1e4).sum() > 0, "no node ruptured"


TUNINGS = [(0, 16, 64, 16, 64), (1, 4, 1, 4, 3), (3, 8, 5, 16, 2), (2, 16, 64, 8, 7), (3, 4, 1000, 4, 1000),
           (9, 8, 5, 4, 3), (25, 4, 2, 8, 7), (19, 16, 3, 8, 1000)]


@pytest.mark.parametrize("tuning", TUNINGS, ids=[f"m{t[0]}_{t[1]}x{t[2]}_{t[3]}x{t[4]}" for t in TUNINGS])
def test_launch_shapes_bit_exact(tuning):
    """Every interior launch shape gives the same wavefield as the oracle
    (partial tiles in x and z, y chunks of 1 row up to the whole range)."""
    case = C.tiny(friction="sw", fspace=False, nabc=5, n=(70, 17, 21), ntfd=25, seed=31)
    say(f"[parity] launch shape {tuning} grid {case.nxt}x{case.nyt}x{case.nzt}")
    orc = O.Oracle(case, O.F32)
    ro = orc.step(case.ntfd)
    with capi.Solver(case) as s:
        s.set_tuning(*tuning)
        got = s.get_tuning()
        assert got[0] == tuning[0] and (not (tuning[0] & 1) or got == tuning)
        rg = s.step(case.ntfd)
        fields = {n: s.field(n) for n in capi.FIELDS}
    for n in OUTS:
        assert np.array_equal(rg[n].astype(np.float64), ro[n]), n
    for n in capi.FIELDS:
        assert np.array_equal(fields[n].astype(np.float64), orc.field(n)), n


WIDE = [(25, 8, 4, 8, 3), (25, 4, 1, 4, 64), (9, 8, 2, 16, 5), (17, 4, 3, 8, 1)]


@pytest.mark.parametrize("tuning", WIDE, ids=[f"m{t[0]}_{t[1]}x{t[2]}_{t[3]}x{t[4]}" for t in WIDE])
def test_wide_tiles_bit_exact(tuning):
    """128-wide interior tiles on a grid of two full tiles and a partial
    one along x, with a free surface: the oracle's wavefield bit for bit."""
    case = C.tiny(friction="sw", fspace=False, nabc=4, n=(268, 11, 23), ntfd=12, seed=37)
    say(f"[parity] wide tiles {tuning} grid {case.nxt}x{case.nyt}x{case.nzt}")
    orc = O.Oracle(case, O.F32)
    ro = orc.step(case.ntfd)
    with capi.Solver(case) as s:
        s.set_tuning(*tuning)
        assert s.get_tuning() == tuning
        rg = s.step(case.ntfd)
        fields = {n: s.field(n) for n in capi.FIELDS}
    for n in OUTS:
        assert np.array_equal(rg[n].astype(np.float64), ro[n]), n
    for n in capi.FIELDS:
        assert np.array_equal(fields[n].astype(np.float64), orc.field(n)), n


def test_wide_tiles_reject_bad_shapes():
    case = C.tiny(friction="sw", fspace=True, seed=38)
    with capi.Solver(case) as s:
        for bad in ((8, 8, 4, 8, 4), (9, 16, 4, 8, 4), (17, 8, 4, 16, 4), (32, 8, 4, 8, 4), (65, 8, 4, 8, 4),
                    (128, 8, 4, 8, 4)):
            with pytest.raises(capi.FD3DError) as e:
                s.set_tuning(*bad)
            assert e.value.code == -1, bad


FUSED = [("fs_ny3", dict(nabc=5, n=(70, 17, 21), seed=33), 3, True),
         ("fs_ny1", dict(nabc=4, n=(37, 13, 30), seed=34), 1, True),
         ("fs_ny64", dict(nabc=6, n=(130, 20, 14), seed=35), 64, True),
         ("fs_tallz", dict(nabc=4, n=(24, 10, 50), seed=36), 4, True),
         ("free_ny3", dict(nabc=5, n=(70, 17, 21), seed=37), 3, False),
         ("free_ny64", dict(nabc=6, n=(130, 20, 14), seed=38), 64, False),
         ("free_tallz", dict(nabc=4, n=(24, 10, 50), seed=39), 4, False),
         ("free_nz5", dict(nabc=4, n=(40, 12, 5), seed=40), 8, False),
         # 5 tiles across: a strip of 4 tiles and a narrower last strip
         ("fs_strips", dict(nabc=4, n=(300, 9, 25), seed=43), 8, True),
         ("free_strips", dict(nabc=4, n=(300, 9, 25), seed=44), 8, False),
         # strips 2 (2, 2, 1), 3 (3, 2) and 8 (one strip of 5) tiles wide
         ("fs_strips2", dict(nabc=4, n=(300, 9, 25), seed=45), 8, True, 5 | 32),
         ("free_strips3", dict(nabc=4, n=(300, 9, 25), seed=46), 8, False, 5 | 96),
         ("fs_strips8", dict(nabc=4, n=(300, 9, 25), seed=47), 8, True, 5 | 64),
         ("fs_tiledpml", dict(nabc=8, n=(70, 17, 21), seed=41), 8, True, 7),
         ("free_tiledpml", dict(nabc=8, n=(70, 17, 21), seed=42), 8, False, 7)]


@pytest.mark.parametrize("name,kw,ny,fspace,mode", [f if len(f) == 5 else f + (5,) for f in FUSED],
                         ids=[f[0] for f in FUSED])
def test_fused_interior_bit_exact(name, kw, ny, fspace, mode):
    """The fused velocity + stress sweep gives the oracle's wavefield and
    outputs bit for bit: partial tiles in x and z, chunks of 1 row up to the
    whole y range, the near-fault rows and PML neighbours read from the
    other kernels' new velocities; under a free surface the top two planes'
    stresses come after the ghost velocities (down to 2 interior planes).
    Starts from a random wavefield so every cell is non-zero from the first
    step."""
    case = C.tiny(friction="sw", fspace=fspace, ntfd=12, **kw)
    say(f"[parity] fused sweep {name} ny={ny} grid {case.nxt}x{case.nyt}x{case.nzt}")
    rng = np.random.Generator(np.random.PCG64(7))
    init = {n: (rng.standard_normal((case.nxt, case.nyt, case.nzt)) *
                (1e6 if n in ("xx", "yy", "zz", "xy", "yz", "xz") else 1e-2)).astype(np.float32)
            for n in capi.FIELDS}
    orc = O.Oracle(case, O.F32)
    for n, a in init.items():
        orc.set_field(n, a)
    ro = orc.step(case.ntfd)
    with capi.Solver(case) as s:
        for n, a in init.items():
            s.set_field(n, a)
        s.set_tuning(mode, 16, ny, 16, ny)     # mode 7: with the tiled PML sweeps
        assert s.get_tuning()[0] == mode
        rg = s.step(case.ntfd)
        fields = {n: s.field(n) for n in capi.FIELDS}
    for n in OUTS:
        assert np.array_equal(rg[n].astype(np.float64), ro[n]), n
    for n in capi.FIELDS:
        bad = np.argwhere(fields[n].astype(np.float64) != orc.field(n))
        assert bad.size == 0, f"field {n}: {len(bad)} mismatches, first (0-based) {bad[0].tolist()}"


@pytest.mark.parametrize("fspace", [False, True], ids=["free", "fs"])
def test_fused_equals_separate_rate_and_state(fspace):
    """Rate-and-state from a random wavefield: the fused sweep (graph
    replay) and the separate sweeps (kernel by kernel) agree bit for bit in
    every field and output, free surface and full space."""
    case = C.tiny(friction="fvw", fspace=fspace, ntfd=16, nabc=5, n=(70, 17, 24), seed=57)
    say(f"[parity] fused against separate, rate-and-state, fspace={fspace}")
    rng = np.random.Generator(np.random.PCG64(9))
    init = {n: (rng.standard_normal((case.nxt, case.nyt, case.nzt)) *
                (1e3 if n in ("xx", "yy", "zz", "xy", "yz", "xz") else 1e-6)).astype(np.float32)
            for n in capi.FIELDS}
    res = []
    for tuning, prof in (((5, 16, 4, 16, 4), False), ((1, 8, 16, 8, 16), True)):
        with capi.Solver(case) as s:
            for n, a in init.items():
                s.set_field(n, a)
            s.set_tuning(*tuning)
            s.set_profiling(prof)
            out = s.step(case.ntfd)
            res.append((out, {n: s.field(n) for n in capi.FIELDS}))
    (oa, fa), (ob, fb) = res
    for n in OUTS + ("psi",):
        assert np.array_equal(oa[n], ob[n]), n
    for n in capi.FIELDS:
        assert np.array_equal(fa[n], fb[n]), n


def layered_medium(case, layers=((5200.0, 3000.0, 2500.0), (5600.0, 3230.0, 2600.0), (6000.0, 3464.0, 2670.0))):
    """lam1, mu1, d1 of a depth-layered crust: the planes k are split into
    len(layers) bands, each filled with one (Vp, Vs, rho) as
    readcrustalmodel fills a layer (fd3d_init.f90:240-265)."""
    f = np.float32
    lam = np.empty((case.nxt, case.nyt, case.nzt), np.float32, order="F")
    mu = np.empty_like(lam)
    d = np.empty_like(lam)
    nb = len(layers)
    for k in range(case.nzt):
        vp, vs, rho = (f(x) for x in layers[min(nb - 1, k * nb // case.nzt)])
        lam[:, :, k] = rho * (vp * vp - f(2.0) * (vs * vs))
        mu[:, :, k] = rho * (vs * vs)
        d[:, :, k] = rho
    return lam, mu, d


@pytest.mark.parametrize("fspace", [True, False], ids=["fs", "free"])
def test_layered_medium_bit_exact(fspace):
    """Three layers of different Vp, Vs, rho: the per-plane medium tables
    give the reference's lam(k), mu(k), mu(k+1) averages in every sweep."""
    nsteps = 60
    case = spread_stations(C.tiny(ntfd=nsteps, friction="sw", fspace=fspace, nabc=5, n=(40, 14, 30), seed=11))
    med = layered_medium(case)
    say(f"[parity] layered medium fspace={fspace} grid {case.nxt}x{case.nyt}x{case.nzt} steps {nsteps}")
    orc = O.Oracle(case, O.F32, medium=med)
    ro = orc.step(nsteps)
    with capi.Solver(case, medium=med) as s:
        rg = s.step(nsteps)
        fields = {n: s.field(n) for n in capi.FIELDS}
    for n in OUTS:
        assert np.array_equal(rg[n].astype(np.float64), ro[n]), n
    assert np.array_equal(rg["seis"].astype(np.float64), ro["seis"])
    for n in capi.FIELDS:
        assert np.array_equal(fields[n].astype(np.float64), orc.field(n)), n


def test_medium_varying_within_a_plane_is_refused():
    case = C.tiny(ntfd=4, friction="sw", fspace=True)
    lam, mu, d = C.material(case, np.float32)
    mu[3, 2, case.nzt // 2] *= np.float32(1.01)
    say("[parity] non-layered medium refused")
    with pytest.raises(capi.FD3DError) as e:
        capi.Solver(case, medium=(lam, mu, d))
    assert e.value.code == -1 and "plane" in str(e.value)


def test_graph_replay_equals_kernel_by_kernel():
    """fd3d_step replays a captured step graph (profiling off) or launches
    kernel by kernel (profiling on); both give the same outputs and
    wavefield, also when a tuning change rebuilds the graph mid-run."""
    case = C.tiny(ntfd=30, friction="sw", fspace=True, nabc=4, n=(50, 12, 22), seed=41)
    say(f"[parity] graph replay vs kernel by kernel, grid {case.nxt}x{case.nyt}x{case.nzt}")
    with capi.Solver(case) as a, capi.Solver(case) as b:
        b.set_profiling(True)
        ra1, rb1 = a.step(10), b.step(10)
        a.set_tuning(1, 8, 3, 4, 5)
        b.set_tuning(1, 8, 3, 4, 5)
        ra2, rb2 = a.step(20), b.step(20)
        fa = {n: a.field(n) for n in capi.FIELDS}
        fb = {n: b.field(n) for n in capi.FIELDS}
    for ra, rb in ((ra1, rb1), (ra2, rb2)):
        for n in OUTS:
            assert np.array_equal(ra[n], rb[n]), n
    for n in capi.FIELDS:
        assert np.array_equal(fa[n], fb[n]), n
