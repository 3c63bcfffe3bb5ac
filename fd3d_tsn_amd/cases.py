"""Rupture problem definitions and their synthetic inputs.

A :class:`Case` holds what the reference reads from ``inputfd3d.dat``
(src/fd3d_init.f90:145-160), the one-layer ``crustal.dat`` material
(:229-240) and the fault setup.  ``material`` and ``fault_sw`` /
``fault_fvw`` produce the module arrays that the reference's setup
routines would fill (lam1/mu1/d1, striniX/Z, peak_xz, dyn_xz, Dc, coh or
the rate-and-state set), in Fortran order, as numpy arrays.

Grid conventions (src/fd3d_init.f90:161-170): nxt = nxtT + 2*nabc,
nyt = nytT + nabc, nzt = nztT + nabc + nfs with nfs = nabc in a full space
(FSPACE) and 2 with a free surface.  The fault is the plane j = nyt; its
nodes are i in [nabc+1, nxt-nabc], k in [nabc+1, nzt-nfs]; k grows upward.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from .fields import random_field


@dataclass
class Case:
    name: str
    nxtT: int
    nytT: int
    nztT: int
    dh: float
    dt: float
    ntfd: int
    nabc: int = 10
    pml_vp: float = 6000.0
    pml_fact: float = 28.0
    damp_s: float = 0.3
    fspace: bool = True
    friction: str = "sw"              # "sw" (linear slip weakening) or "fvw"
    # one-layer crustal.dat line: Vp, Vs (km/s), rho (g/cm^3) as decimal text
    vp: str = "6.000"
    vs: str = "3.464"
    rho: str = "2.670"
    stations: List[Tuple[int, int, int]] = field(default_factory=list)
    seed: int = 1
    # fault box (metres, measured from the first fault node at x = z = 0)
    fault_len: float = 20e3
    fault_wid: float = 10e3
    fault_xc: Optional[float] = None   # default: centre of the fault plane
    fault_zc: Optional[float] = None
    # slip weakening (TPV5-style, src/fd3d_init.f90:535-617)
    mus: float = 0.677
    mud: float = 0.525
    mus_out: float = 1.0e4
    dc: float = 0.4
    dc_range: Optional[Tuple[float, float]] = None
    sn: float = 120e6
    t0: float = 70e6
    t0_std: float = 3e6
    t0_corr: float = 2e3
    t0_kind: str = "gaussian"
    hurst: float = 0.75
    nuc_size: float = 3e3
    nuc_stress: float = 81.6e6
    nuc_xc: Optional[float] = None
    nuc_zc: Optional[float] = None
    # rate-and-state fast velocity weakening (TPV104-style, :393-501)
    f0: float = 0.6
    v0: float = 1e-6
    a0: float = 0.01
    b0: float = 0.014
    L: float = 0.4
    fw: float = 0.2
    vw0: float = 0.1
    vw_delta: float = 0.9
    vini: float = 0.5e-16
    tini: float = 40e6
    trans: float = 3e3
    perturb: float = 45e6
    nuc_radius: float = 3e3
    nuc_time: float = 1.0

    # ---------------------------------------------------------------- dims
    @property
    def nfs(self) -> int:
        return self.nabc if self.fspace else 2

    @property
    def nxt(self) -> int:
        return self.nxtT + 2 * self.nabc

    @property
    def nyt(self) -> int:
        return self.nytT + self.nabc

    @property
    def nzt(self) -> int:
        return self.nztT + self.nabc + self.nfs

    @property
    def cells(self) -> int:
        return self.nxt * self.nyt * self.nzt

    @property
    def fault_shape(self) -> Tuple[int, int]:
        """Per-step output slab (nxtT, nztT) (fd3d_theo.f90:548-553)."""
        return (self.nxt - 2 * self.nabc, self.nzt - self.nfs - self.nabc)

    def replace(self, **kw) -> "Case":
        return dataclasses.replace(self, **kw)

    def cfl(self) -> float:
        vp = float(np.float32(self.vp)) * 1e3
        return vp * self.dt / self.dh

    # -------------------------------------------------------- text inputs
    def inputfd3d_text(self) -> str:
        """The reference's inputfd3d.dat (src/fd3d_init.f90:145-160)."""
        lines = [f"{self.nxtT} {self.nytT} {self.nztT}", _num(self.dh), str(self.ntfd),
                 _num(self.dt), "0.", f"{self.nabc} {_num(self.pml_vp)} {_num(self.pml_fact)}",
                 _num(self.damp_s), str(len(self.stations))]
        lines += [f"{i} {j} {k}" for (i, j, k) in self.stations]
        lines.append("0")
        return "\n".join(lines) + "\n"

    def crustal_text(self) -> str:
        """One-layer crustal.dat (src/fd3d_init.f90:229-240)."""
        return ("Crustal model\nnumber of layers\n  1\nParameters of the layers\n"
                "depth of layer top(km)   Vp(km/s)    Vs(km/s)    Rho(g/cm**3)    Qp     Qs\n"
                f"42.000 {self.vp} {self.vs} {self.rho} 800. 400.\n")

    # ------------------------------------------------------------ helpers
    def _centre(self) -> Tuple[float, float]:
        nxT, nzT = self.fault_shape
        xc = self.fault_xc if self.fault_xc is not None else 0.5 * (nxT - 1) * self.dh
        zc = self.fault_zc if self.fault_zc is not None else 0.5 * (nzT - 1) * self.dh
        return xc, zc

    def _node_xz(self, stagger: float = 0.0) -> Tuple[np.ndarray, np.ndarray]:
        """Fault coordinates of the (nxt, nzt) plane, 0 at the first fault node."""
        i = np.arange(1, self.nxt + 1, dtype=np.float64)
        k = np.arange(1, self.nzt + 1, dtype=np.float64)
        x = (i - self.nabc - 1 + stagger) * self.dh
        z = (k - self.nabc - 1 + stagger) * self.dh
        return x[:, None] * np.ones((1, self.nzt)), np.ones((self.nxt, 1)) * z[None, :]


def _num(x: float) -> str:
    """Decimal text that reads back as the same float32 (and float64) value."""
    s = repr(float(x))
    if np.float32(s) != np.float32(x):
        raise ValueError(f"{x!r} does not survive text round trip")
    return s


def fortran(a: np.ndarray) -> np.ndarray:
    return np.asfortranarray(a)


# -------------------------------------------------------------------- medium
def material(case: Case, dtype=np.float32):
    """lam1, mu1, d1 as readcrustalmodel fills them for one layer
    (src/fd3d_init.f90:236-265): values from crustal.dat text scaled by 1e3,
    lam = rho*(vp**2 - 2*vs**2), mu = rho*vs**2, in the given precision."""
    t = np.dtype(dtype).type
    k = t(1.0e3)
    vp = t(case.vp) * k
    vs = t(case.vs) * k
    rho = t(case.rho) * k
    lam_v = rho * (vp * vp - t(2.0) * (vs * vs))
    mu_v = rho * (vs * vs)
    shape = (case.nxt, case.nyt, case.nzt)
    lam = np.full(shape, lam_v, dtype=dtype, order="F")
    mu = np.full(shape, mu_v, dtype=dtype, order="F")
    d = np.full(shape, rho, dtype=dtype, order="F")
    return lam, mu, d


def material_scalars(case: Case, dtype=np.float32):
    t = np.dtype(dtype).type
    k = t(1.0e3)
    vp = t(case.vp) * k
    vs = t(case.vs) * k
    rho = t(case.rho) * k
    return rho * (vp * vp - t(2.0) * (vs * vs)), rho * (vs * vs), rho


# ------------------------------------------------------ slip-weakening fault
def fault_sw(case: Case, dtype=np.float32) -> Dict[str, np.ndarray]:
    """Z-node slip-weakening inputs (striniX, striniZ, peak_xz, dyn_xz, Dc, coh).

    Inside the fault box: static friction mus*sn, dynamic mud*sn, shear
    prestress t0 + seeded random field, a square nucleation patch at
    nuc_stress.  Outside: static friction mus_out*sn (the TPV5 barrier,
    src/fd3d_init.f90:581-591).  Normal stress is uniform (no depth term).
    """
    x, z = case._node_xz()
    xc, zc = case._centre()
    inside = (np.abs(x - xc) <= 0.5 * case.fault_len) & (np.abs(z - zc) <= 0.5 * case.fault_wid)
    noise = random_field(case.nxt, case.nzt, case.dh, case.t0_corr, case.t0_std, case.seed,
                         kind=case.t0_kind, hurst=case.hurst)
    strinix = case.t0 + noise
    nx = case.nuc_xc if case.nuc_xc is not None else xc
    nz = case.nuc_zc if case.nuc_zc is not None else zc
    nuc = (np.abs(x - nx) <= 0.5 * case.nuc_size) & (np.abs(z - nz) <= 0.5 * case.nuc_size)
    strinix = np.where(nuc, case.nuc_stress, strinix)
    peak = np.where(inside, case.sn * case.mus, case.sn * case.mus_out)
    dyn = np.full_like(peak, case.sn * case.mud)
    if case.dc_range is None:
        dc = np.full_like(peak, case.dc)
    else:
        rng = np.random.Generator(np.random.PCG64(case.seed + 7919))
        dc = np.full_like(peak, rng.uniform(*case.dc_range))
    out = {"striniX": strinix, "striniZ": np.zeros_like(peak), "peak_xz": peak,
           "dyn_xz": dyn, "Dc": dc, "coh": np.zeros_like(peak)}
    return {k: fortran(v.astype(dtype)) for k, v in out.items()}


SW_FIELDS = ("striniX", "striniZ", "peak_xz", "dyn_xz", "Dc", "coh")
FVW_FIELDS = ("Dc", "striniX", "striniZ", "a", "b", "psi", "vw", "aX", "bX", "psiX", "vwX")
FVW_SCALARS = ("Sn", "uini", "wini", "f0", "v0", "fw", "hx0", "hz0", "perturb", "RR2", "TT2",
               "strinixI")


# ------------------------------------------------------- rate-and-state fault
def _taper(dist: np.ndarray, half: float, trans: float) -> np.ndarray:
    """TPV104 boxcar-with-tanh taper B(dist) (src/fd3d_init.f90:446-460)."""
    out = np.ones_like(dist)
    outer = dist >= half + trans
    ramp = (dist > half) & (dist < half + trans)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = 0.5 * (1.0 + np.tanh(trans / (dist - half - trans) + trans / (dist - half)))
    out = np.where(ramp, r, out)
    out = np.where(outer, 0.0, out)
    return out


def fault_fvw(case: Case) -> Tuple[Dict[str, np.ndarray], Dict[str, float]]:
    """Rate-and-state inputs in float64, laid out like forwardspecialTPV104
    (src/fd3d_init.f90:393-501): a and vw ramp from the fault box to the
    velocity-strengthening border through the tanh taper, psi from the
    steady state at the initial slip rate, X-node parameters evaluated at
    the staggered positions (i+1/2, k+1/2).  Scalars include the smooth
    nucleation centre, radius^2 and duration used by the time loop
    (fd3d_theo.f90:515-542)."""
    dh = case.dh
    xc, zc = case._centre()
    rng = np.random.Generator(np.random.PCG64(case.seed))
    nuc_x = case.nuc_xc if case.nuc_xc is not None else xc + rng.uniform(-0.25, 0.25) * case.fault_len
    nuc_z = case.nuc_zc if case.nuc_zc is not None else zc + rng.uniform(-0.25, 0.25) * case.fault_wid
    # the reference's nucleation uses hx = i*dh, hz = k*dh (fd3d_theo.f90:528-529)
    hx0 = nuc_x + (case.nabc + 1) * dh
    hz0 = nuc_z + (case.nabc + 1) * dh
    wfx, wfz = 0.5 * case.fault_len, 0.5 * case.fault_wid

    def params(stagger: float):
        x, z = case._node_xz(stagger)
        bx = _taper(np.abs(x - xc), wfx, case.trans)
        bz = _taper(np.abs(z - zc), wfz, case.trans)
        a = case.a0 + case.a0 * (1.0 - bx * bz)
        vw = case.vw0 + case.vw_delta * (1.0 - bx * bz)
        strinix = np.full_like(a, case.tini)
        psi = a * (np.log((2 * case.v0) / (2 * case.vini)) + np.log(np.sinh(strinix / (a * case.sn))))
        return a, vw, psi, strinix

    a, vw, psi, strinix = params(0.0)
    aX, vwX, psiX, _ = params(0.5)
    b = np.full_like(a, case.b0)
    arrays = {"Dc": np.full_like(a, case.L), "striniX": strinix, "striniZ": np.zeros_like(a),
              "a": a, "b": b, "psi": psi, "vw": vw, "aX": aX, "bX": b.copy(), "psiX": psiX,
              "vwX": vwX}
    scal = {"Sn": case.sn, "uini": case.vini, "wini": 0.0, "f0": case.f0, "v0": case.v0,
            "fw": case.fw, "hx0": hx0, "hz0": hz0, "perturb": case.perturb,
            "RR2": case.nuc_radius ** 2, "TT2": case.nuc_time, "strinixI": case.tini}
    return {k: fortran(v.astype(np.float64)) for k, v in arrays.items()}, scal


# ----------------------------------------------------------------- workloads
def c1_plain_sw() -> Case:
    """Plain method, CPU-runnable: 128x64x64, dh 200 m, dt 8 ms, 500 steps,
    10-node PML, full space, 20x10 km slip-weakening fault."""
    return Case("plain_sw_128x64x64", 128, 64, 64, dh=200.0, dt=0.008, ntfd=500, nabc=10,
                fspace=True, seed=1)


def c2_large_sw() -> Case:
    """Single large rupture: 1024x512x512, dh 50 m, dt 2 ms, 4000 steps,
    20-node PML, full space, 40x20 km fault, von Karman prestress
    (H 0.75, 5 km, 8 MPa)."""
    return Case("large_sw_1024x512x512", 1024, 512, 512, dh=50.0, dt=0.002, ntfd=4000, nabc=20,
                fspace=True, fault_len=40e3, fault_wid=20e3, t0_kind="vonkarman", t0_std=8e6,
                t0_corr=5e3, hurst=0.75, seed=2,
                stations=_seeded_stations(1024, 512, 512, 20, True, 16, seed=2))


def c3_fvw() -> Case:
    """Rate-and-state single rupture: 512x256x256, dh 100 m, dt 2 ms,
    3000 steps, 20-node PML, free surface."""
    return Case("fvw_512x256x256", 512, 256, 256, dh=100.0, dt=0.002, ntfd=3000, nabc=20,
                fspace=False, friction="fvw", fault_len=36e3, fault_wid=18e3, seed=3)


def c4_batched(n: int = 64, seed: int = 4) -> List[Case]:
    """Batched uneven domains: n independent slip-weakening ruptures with
    interior sides drawn uniformly from 96..320 per axis, dh 100 m,
    dt 4 ms (CFL 0.24), 1500 steps, 10-node PML, full space.  Per
    instance: Dc ~ U(0.2, 0.8) m, mu_s ~ U(0.6, 0.75), its own Gaussian
    prestress field (2 km correlation) and nucleation location.
    Chosen where the workload leaves it open: the fault box covers 80% of
    the fault plane; prestress 2 MPa std about sn*(mu_d + 0.6*(mu_s-mu_d));
    the 3x3 km nucleation patch sits 1.5 MPa above static strength,
    anywhere in the central half of the box."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for q in range(n):
        nx, ny, nz = (int(v) for v in rng.integers(96, 321, size=3))
        dc = float(rng.uniform(0.2, 0.8))
        mus = float(rng.uniform(0.6, 0.75))
        fx, fz = rng.uniform(-0.25, 0.25, size=2)
        dh, sn, mud = 100.0, 120e6, 0.525
        flen, fwid = 0.8 * nx * dh, 0.8 * nz * dh
        xc, zc = 0.5 * (nx - 1) * dh, 0.5 * (nz - 1) * dh
        out.append(Case(f"batch{q:02d}_{nx}x{ny}x{nz}", nx, ny, nz, dh=dh, dt=0.004, ntfd=1500, nabc=10,
                        fspace=True, seed=seed * 1000 + q, fault_len=flen, fault_wid=fwid, mus=mus, mud=mud,
                        dc=dc, sn=sn, t0=sn * (mud + 0.6 * (mus - mud)), t0_std=2e6, t0_corr=2e3,
                        nuc_size=3e3, nuc_stress=sn * mus + 1.5e6,
                        nuc_xc=xc + float(fx) * flen, nuc_zc=zc + float(fz) * fwid))
    return out


def _seeded_stations(nxT, nyT, nzT, nabc, fspace, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed + 101))
    nfs = nabc if fspace else 2
    nxt, nyt, nzt = nxT + 2 * nabc, nyT + nabc, nzT + nabc + nfs
    st = []
    for _ in range(n):
        st.append((int(rng.integers(nabc + 1, nxt - nabc + 1)), int(rng.integers(nabc + 1, nyt + 1)),
                   int(rng.integers(nabc + 1, nzt - nfs + 1))))
    return st


def tiny(friction: str = "sw", fspace: bool = True, nabc: int = 6, n=(40, 16, 20),
         ntfd: int = 60, seed: int = 11, **kw) -> Case:
    """Small problems for parity tests (the oracle runs them in well under a second)."""
    nxT, nyT, nzT = n
    dh = 200.0
    base = dict(name=f"tiny_{friction}_{'fs' if fspace else 'free'}_{nxT}x{nyT}x{nzT}",
                nxtT=nxT, nytT=nyT, nztT=nzT, dh=dh, dt=0.008, ntfd=ntfd, nabc=nabc,
                fspace=fspace, friction=friction, seed=seed,
                fault_len=0.6 * nxT * dh, fault_wid=0.6 * nzT * dh,
                nuc_size=4 * dh, t0_corr=3 * dh)
    if friction == "fvw":
        base.update(dt=0.004, nuc_radius=5 * dh, trans=3 * dh, nuc_time=0.051)
    base.update(kw)
    return Case(**base)
