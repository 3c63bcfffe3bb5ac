"""Closed-form checks of the oracle's stencils (independent of the reference build).

The 4th-order staggered operator c1(f(i)-f(i-1)) + c2(f(i+1)-f(i-2)),
c1 = 9/8, c2 = -1/24, equals df/di at i-1/2 exactly for polynomials of
degree <= 4.  Starting from polynomial stresses (velocities) and zero
velocities (stresses), one step of the oracle must give, at interior cells
whose stencils see only interior points, the analytic staggered
derivatives -- with the reference's material averages, including the
diagonal mu(i+1,j+1,k+1) of the yz update (fd3d_deriv.f90:602-604).
A wrong coefficient, offset, sign or operand fails these checks."""
import numpy as np
import pytest

from fd3d_tsn_amd import cases as C

import oracle_lib as O


def cubic(rng):
    """Random cubic P(i, j, k) and its gradient, in index units."""
    c = rng.normal(size=20)

    def P(i, j, k):
        terms = [np.ones_like(i), i, j, k, i * i, j * j, k * k, i * j, j * k, i * k,
                 i ** 3, j ** 3, k ** 3, i * i * j, i * j * j, j * j * k, j * k * k, i * i * k, i * k * k, i * j * k]
        return sum(cc * t for cc, t in zip(c, terms))

    def dP(i, j, k):
        di = c[1] + 2 * c[4] * i + c[7] * j + c[9] * k + 3 * c[10] * i * i + 2 * c[13] * i * j + c[14] * j * j + \
            2 * c[17] * i * k + c[18] * k * k + c[19] * j * k
        dj = c[2] + 2 * c[5] * j + c[7] * i + c[8] * k + 3 * c[11] * j * j + c[13] * i * i + 2 * c[14] * i * j + \
            2 * c[15] * j * k + c[16] * k * k + c[19] * i * k
        dk = c[3] + 2 * c[6] * k + c[8] * j + c[9] * i + 3 * c[12] * k * k + c[15] * j * j + 2 * c[16] * j * k + \
            c[17] * i * i + 2 * c[18] * i * k + c[19] * i * j
        return di, dj, dk

    return P, dP


def grids(case):
    i = np.arange(1, case.nxt + 1, dtype=np.float64)[:, None, None]
    j = np.arange(1, case.nyt + 1, dtype=np.float64)[None, :, None]
    k = np.arange(1, case.nzt + 1, dtype=np.float64)[None, None, :]
    return np.broadcast_arrays(i, j, k)


def deep(case):
    nb, nfs = case.nabc, case.nfs
    return (slice(nb + 2, case.nxt - nb - 2), slice(nb + 2, case.nyt - 4), slice(nb + 2, case.nzt - nfs - 2))


@pytest.fixture(scope="module")
def case():
    c = C.tiny("sw", True, nabc=4, n=(24, 16, 20), ntfd=1, seed=3)
    return c.replace(fault_len=1.0, fault_wid=1.0, mus_out=1e4, nuc_size=0.0)


def quiet_fault(case):
    f = C.fault_sw(case, np.float64)
    f["striniX"][:] = 0.0
    return f


def test_velocity_update_is_exact_on_cubics(case):
    rng = np.random.Generator(np.random.PCG64(5))
    I, J, K = grids(case)
    polys = {n: cubic(rng) for n in ("xx", "yy", "zz", "xy", "yz", "xz")}
    orc = O.Oracle(case, O.F64, fault=quiet_fault(case))
    for n, (P, _) in polys.items():
        orc.set_field(n, P(I, J, K))
    orc.step(1, want_out=False)
    lam, mu, rho = C.material_scalars(case, np.float64)
    dth = case.dt / case.dh
    s = deep(case)
    Ii, Jj, Kk = I[s], J[s], K[s]
    d = {n: dp for n, (_, dp) in polys.items()}
    u_exp = (dth / rho) * (d["xx"](Ii - .5, Jj, Kk)[0] + d["xy"](Ii, Jj - .5, Kk)[1] + d["xz"](Ii, Jj, Kk - .5)[2])
    v_exp = (dth / rho) * (d["xy"](Ii + .5, Jj, Kk)[0] + d["yy"](Ii, Jj + .5, Kk)[1] + d["yz"](Ii, Jj, Kk - .5)[2])
    w_exp = (dth / rho) * (d["xz"](Ii + .5, Jj, Kk)[0] + d["yz"](Ii, Jj - .5, Kk)[1] + d["zz"](Ii, Jj, Kk + .5)[2])
    for name, exp in (("u", u_exp), ("v", v_exp), ("w", w_exp)):
        got = orc.field(name)[s]
        assert np.allclose(got, exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max()), name


def test_stress_update_is_exact_on_cubics_with_heterogeneous_medium(case):
    rng = np.random.Generator(np.random.PCG64(8))
    I, J, K = grids(case)
    polys = {n: cubic(rng) for n in ("u", "v", "w")}
    lam0, mu0, rho0 = C.material_scalars(case, np.float64)
    lam = lam0 * (1 + 0.2 * rng.random(I.shape))
    mu = mu0 * (1 + 0.2 * rng.random(I.shape))
    rho = np.full(I.shape, rho0)
    orc = O.Oracle(case, O.F64, fault=quiet_fault(case), medium=(lam, mu, rho))
    for n, (P, _) in polys.items():
        orc.set_field(n, P(I, J, K))
    orc.step(1, want_out=False)
    dth = case.dt / case.dh
    s = deep(case)
    Ii, Jj, Kk = I[s], J[s], K[s]
    du, dv, dw = (polys[n][1] for n in ("u", "v", "w"))
    ux = du(Ii + .5, Jj, Kk)[0]
    vy = dv(Ii, Jj - .5, Kk)[1]
    wz = dw(Ii, Jj, Kk - .5)[2]
    L, M = lam[s], mu[s]
    A = L + 2 * M
    ex = {
        "xx": dth * (A * ux + L * (vy + wz)),
        "yy": dth * (A * vy + L * (ux + wz)),
        "zz": dth * (A * wz + L * (ux + vy)),
    }
    mu_j = mu[s[0], slice(s[1].start + 1, s[1].stop + 1), s[2]]
    mu_k = mu[s[0], s[1], slice(s[2].start + 1, s[2].stop + 1)]
    mu_d = mu[slice(s[0].start + 1, s[0].stop + 1), slice(s[1].start + 1, s[1].stop + 1),
              slice(s[2].start + 1, s[2].stop + 1)]
    ex["xy"] = dth * 0.5 * (M + mu_j) * (du(Ii, Jj + .5, Kk)[1] + dv(Ii - .5, Jj, Kk)[0])
    ex["xz"] = dth * 0.5 * (M + mu_k) * (du(Ii, Jj, Kk + .5)[2] + dw(Ii - .5, Jj, Kk)[0])
    ex["yz"] = dth * 0.5 * (M + mu_d) * (dv(Ii, Jj, Kk + .5)[2] + dw(Ii, Jj + .5, Kk)[1])
    for name, exp in ex.items():
        got = orc.field(name)[s]
        assert np.allclose(got, exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max()), name


def test_constant_fields_do_not_move(case):
    """Uniform stresses and velocities produce no interior update."""
    I, J, K = grids(case)
    orc = O.Oracle(case, O.F64, fault=quiet_fault(case))
    for n, val in (("xx", 3e6), ("yy", -2e6), ("xy", 1e6), ("u", 0.3), ("w", -0.1)):
        orc.set_field(n, np.full(I.shape, val))
    orc.step(1, want_out=False)
    s = deep(case)
    assert np.all(orc.field("u")[s] == 0.3) and np.all(orc.field("w")[s] == -0.1)
    assert np.all(orc.field("xx")[s] == 3e6) and np.all(orc.field("xy")[s] == 1e6)
