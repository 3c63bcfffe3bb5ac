// TEST INFRASTRUCTURE ONLY -- the oracle (see fd3d_oracle.h).
//
// A CPU restatement of one FD3D_TSN time step, written to keep
// the reference's per-point arithmetic: the same operands, the same
// left-to-right association, no fused multiply-add (built with
// -ffp-contract=off), correctly rounded IEEE division and square root.
// With that, the float mode reproduces the reference's FP32 CPU build bit
// for bit on the slip-weakening path (checked by tests/test_oracle_vs_ref.py
// against the flang build in oracle/_ref).
//
// The 3D sweeps (dvel, dstres, pml_uvw, pml_xyz) split their outer z loop
// over OpenMP threads: each cell's update reads only values the sweep does
// not write and writes only that cell, so the result does not depend on
// the thread count (OMP_NUM_THREADS=1 gives the sequential program).
//
// Routine-by-routine provenance (reference file:line) is given at each
// function.  The code is organised differently from the reference: the
// five PML regions share one generic routine driven by a region table, and
// the per-step sequence lives in step().
//
// Parity pinning:
//  * every routine below is exercised by the whole-trajectory comparison
//    against oracle/_ref (fault slabs every step, stations inside and
//    outside the PML, final 2D arrays), bit-exact in ORC_F32 for slip
//    weakening with and without FSPACE;
//  * rate-and-state routines (tasu1/tasw1 Newton, FVW friction, nucleation)
//    are compared in ORC_F64 against the -fdefault-real-8 reference build;
//  * closed-form checks in tests/test_oracle_math.py: polynomial exactness
//    of the 4th-order operators, zero update on constant fields, PML
//    profile formula, Newton root against a bracketing solver.
#include "fd3d_oracle.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

template <typename R, typename F>
struct Oracle {
  // ---- configuration (fd3d_init.f90:145-171)
  int nxt = 0, nyt = 0, nzt = 0, nabc = 0, nfs = 0, fspace = 0, fvw = 0, nucl = 0;
  int nsta = 0, newton_max = 200;
  R dh = 0, dt = 0, damp_s = 0, pml_vp = 0, pml_fact = 0, omegaM = 0;
  F Sn = 0, uini = 0, wini = 0, f0 = 0, v0 = 0, fw = 0;
  F hx0 = 0, hz0 = 0, perturb = 0, RR2 = 0, TT2 = 0, strinixI = 0;
  std::vector<int> sta;
  size_t nx = 0, nxy = 0, n3 = 0, n2 = 0;
  int it = 0;
  long newton_capped = 0;

  // ---- 3D state (displt_com, strfld_com, medium_com)
  std::vector<R> lam, mu, d, u, v, w, xx, yy, zz, xy, yz, xz;
  // ---- 2D fault-plane state (traction_com, friction_com, source_com)
  std::vector<R> tx, tz, v1t, avdx, avdz, RFx, RFz, au1, av1, aw1, uZ, wX, tabsX, tabsZ;
  std::vector<R> striniX, striniZ, peak_xz, dyn_xz, Dc, coh;
  std::vector<R> peakX, dynX, DcX, peakZ, dynZ, DcZ, T0X, T0Z;
  std::vector<F> fa, fb, fpsi, fvw_, aX, bX, psiX, vwX, aZ, bZ, psiZ, vwZ;
  std::vector<R> ruptime, rise, sliptime, slipX, slipZ, schangeX, schangeZ, distX, distZ;
  std::vector<R> srX, srZ;
  std::vector<F> psiout;

  // ---- PML regions (pml_com; allocation fd3d_deriv.f90:1324-1437)
  struct Region {
    int ib, ie, jb, je_uw, je_v, je_xz, kb, ke;   // grid ranges (1-based, inclusive)
    int ni, nj, nk;                                // local dims (u-type j extent)
    bool ztop;                                     // zero xz/yz splits at k2 = nk
    int nj_zero_xz, nj_zero_yz;                    // j2 extents of that zeroing
    std::vector<R> ox, oxS, oy, oyS, oz, ozS;      // 1-based damping profiles
    std::vector<R> s[24];                          // split fields
  };
  enum { SU1, SU2, SU3, SV1, SV2, SV3, SW1, SW2, SW3, SXX1, SXX2, SXX3, SYY1, SYY2, SYY3,
         SZZ1, SZZ2, SZZ3, SXY1, SXY2, SXZ1, SXZ2, SYZ1, SYZ2 };
  std::vector<Region> reg;

  size_t I3(int i, int j, int k) const {
    return (size_t)(i - 1) + nx * ((size_t)(j - 1) + (size_t)nyt * (size_t)(k - 1));
  }
  size_t I2(int i, int k) const { return (size_t)(i - 1) + nx * (size_t)(k - 1); }
  static size_t L(const Region &r, int i2, int j2, int k2) {
    return (size_t)(i2 - 1) + (size_t)r.ni * ((size_t)(j2 - 1) + (size_t)r.nj * (size_t)(k2 - 1));
  }
  static R sq(R x) { return x * x; }
  static F sqF(F x) { return x * x; }

  void alloc() {
    nx = (size_t)nxt;
    nxy = nx * (size_t)nyt;
    n3 = nxy * (size_t)nzt;
    n2 = nx * (size_t)nzt;
    for (auto *a : {&lam, &mu, &d, &u, &v, &w, &xx, &yy, &zz, &xy, &yz, &xz}) a->assign(n3, R(0));
    for (auto *a : {&tx, &tz, &v1t, &avdx, &avdz, &RFx, &RFz, &au1, &av1, &aw1, &uZ, &wX,
                    &tabsX, &tabsZ, &striniX, &striniZ, &peak_xz, &dyn_xz, &Dc, &coh,
                    &peakX, &dynX, &DcX, &peakZ, &dynZ, &DcZ, &T0X, &T0Z, &ruptime, &rise,
                    &sliptime, &slipX, &slipZ, &schangeX, &schangeZ, &distX, &distZ, &srX, &srZ})
      a->assign(n2, R(0));
    for (auto *a : {&fa, &fb, &fpsi, &fvw_, &aX, &bX, &psiX, &vwX, &aZ, &bZ, &psiZ, &vwZ, &psiout})
      a->assign(n2, F(0));
  }

  // ------------------------------------------------------------------
  // init_pml (fd3d_deriv.f90:1319-1493): damping profiles per region.
  // omega(i) = 0.5*dt*omegaM*(i/(nabc-1))**4, the M variant at i-0.5,
  // R variants reversed; omegaM = pml_fact*pml_vp/(2*dh*(nabc-1))
  // (fd3d_init.f90:171).  x**4 is formed ((x*x)*x)*x, which is how
  // flang -O2 lowers the integer power.
  // ------------------------------------------------------------------
  void init_pml() {
    omegaM = (pml_fact * pml_vp) / ((R(2) * dh) * R(nabc - 1));
    std::vector<R> om(nabc + 1, 0), omR(nabc + 1, 0), omM(nabc + 1, 0), omRM(nabc + 1, 0);
    const R scale = (R(0.5) * dt) * omegaM;
    for (int i = 1; i <= nabc - 1; ++i) {
      R q = R(i) / R(nabc - 1);
      om[i] = scale * (((q * q) * q) * q);
      omR[nabc - i] = om[i];
      R qm = (R(i) - R(0.5)) / R(nabc - 1);
      omM[i] = scale * (((qm * qm) * qm) * qm);
      omRM[nabc - i] = omM[i];
    }
    const int kint_b = nabc + 1, kint_e = nzt - nfs;
    reg.clear();
    auto mk = [&](int ib, int ie, int jb, int je_uw, int je_v, int je_xz, int kb, int ke,
                  bool ztop, int nzx, int nzy) {
      Region r;
      r.ib = ib; r.ie = ie; r.jb = jb; r.je_uw = je_uw; r.je_v = je_v; r.je_xz = je_xz;
      r.kb = kb; r.ke = ke;
      r.ni = ie - ib + 1; r.nj = je_uw - jb + 1; r.nk = ke - kb + 1;
      r.ztop = ztop; r.nj_zero_xz = nzx; r.nj_zero_yz = nzy;
      r.ox.assign(r.ni + 1, 0); r.oxS.assign(r.ni + 1, 0);
      r.oy.assign(r.nj + 1, 0); r.oyS.assign(r.nj + 1, 0);
      r.oz.assign(r.nk + 1, 0); r.ozS.assign(r.nk + 1, 0);
      for (auto &s : r.s) s.assign((size_t)r.ni * r.nj * r.nk, R(0));
      return r;
    };
    // region 1: x near 1 (fd3d_deriv.f90:1324-1344, profiles :1457-1458;
    // the y/z splits use omegay1/omegaz1, which stay zero)
    {
      Region r = mk(2, nabc, nabc + 1, nyt, nyt - 1, nyt, kint_b, kint_e, true,
                    nyt - nabc, nyt - nabc - 1);
      for (int i = 1; i <= nabc - 1; ++i) { r.ox[i] = omR[i]; r.oxS[i] = omRM[i]; }
      reg.push_back(std::move(r));
    }
    // region 2: x near nxt (:1346-1366, profiles :1460-1461)
    {
      Region r = mk(nxt - nabc + 1, nxt - 1, nabc + 1, nyt, nyt - 1, nyt, kint_b, kint_e, !fspace,
                    nyt - nabc, nyt - nabc - 1);
      for (int i = 1; i <= nabc - 1; ++i) { r.ox[i] = omM[i]; r.oxS[i] = om[i]; }
      reg.push_back(std::move(r));
    }
    // region 3: y near 1 (:1368-1388, profiles :1463-1468)
    {
      Region r = mk(2, nxt - 1, 2, nabc, nabc, nabc, kint_b, kint_e, !fspace, nabc - 1, nabc - 1);
      for (int i = 1; i <= nabc - 1; ++i) {
        r.oy[i] = omR[i];
        r.ox[i] = omR[i];
        r.ox[nxt - i - 1] = omM[i];
        r.oyS[i] = omRM[i];
        r.oxS[i] = omRM[i];
        r.oxS[nxt - i - 1] = om[i];
      }
      reg.push_back(std::move(r));
    }
    // region 4: z near 1 (:1390-1410, profiles :1470-1478)
    {
      Region r = mk(2, nxt - 1, 2, nyt, nyt - 1, nyt, 2, nabc, false, 0, 0);
      for (int i = 1; i <= nabc - 1; ++i) {
        r.oz[i] = omR[i];
        r.ox[i] = omR[i];
        r.ox[nxt - i - 1] = omM[i];
        r.oy[i] = omR[i];
        r.ozS[i] = omRM[i];
        r.oxS[i] = omRM[i];
        r.oxS[nxt - i - 1] = om[i];
        r.oyS[i] = omRM[i];
      }
      reg.push_back(std::move(r));
    }
    // region 5: z near nzt, full space only (:1413-1437, profiles :1480-1489)
    if (fspace) {
      Region r = mk(2, nxt - 1, 2, nyt, nyt - 1, nyt, nzt - nabc + 1, nzt - 1, false, 0, 0);
      for (int i = 1; i <= nabc - 1; ++i) {
        r.oz[i] = omM[i];
        r.ox[i] = omR[i];
        r.ox[nxt - i - 1] = omM[i];
        r.oy[i] = omR[i];
        r.ozS[i] = om[i];
        r.oxS[i] = omRM[i];
        r.oxS[nxt - i - 1] = om[i];
        r.oyS[i] = omRM[i];
      }
      reg.push_back(std::move(r));
    }
    // Regions 1-3 read the y split profile as their "yS"/"zS" profile where
    // the reference passes omegay/omegaz (pml_xyz :2398,2425,2452-2455).
    for (auto &r : reg) {
      bool r12 = (&r == &reg[0]) || (&r == &reg[1]);
      bool r3 = (&r == &reg[2]);
      if (r12) { r.oyS = r.oy; r.ozS = r.oz; }
      if (r3) { r.ozS = r.oz; }
    }
  }

  // ------------------------------------------------------------------
  // interp_fric (fd3d_deriv.f90:3150-3190): Z-node copies and X-node
  // 4-point averages.
  // ------------------------------------------------------------------
  void interp_fric() {
    std::fill(peakX.begin(), peakX.end(), R(0));
    std::fill(DcX.begin(), DcX.end(), R(0));
    std::fill(dynX.begin(), dynX.end(), R(0));
    std::fill(T0X.begin(), T0X.end(), R(0));
    std::fill(peakZ.begin(), peakZ.end(), R(0));
    std::fill(DcZ.begin(), DcZ.end(), R(0));
    std::fill(dynZ.begin(), dynZ.end(), R(0));
    std::fill(T0Z.begin(), T0Z.end(), R(0));
    for (size_t n = 0; n < n2; ++n) {
      if (fvw) {
        aZ[n] = fa[n]; bZ[n] = fb[n]; psiZ[n] = fpsi[n]; vwZ[n] = fvw_[n];
      } else {
        peakZ[n] = peak_xz[n]; dynZ[n] = dyn_xz[n]; DcZ[n] = Dc[n];
      }
      T0Z[n] = striniZ[n];
    }
    for (int k = 2; k <= nzt - 1; ++k)
      for (int i = 2; i <= nxt - 1; ++i) {
        size_t c = I2(i, k), im = I2(i - 1, k), km = I2(i, k - 1), ikm = I2(i - 1, k - 1);
        if (!fvw) {
          peakX[c] = (((peakZ[c] + peakZ[im]) + peakZ[km]) + peakZ[ikm]) / R(4);
          dynX[c] = (((dynZ[c] + dynZ[im]) + dynZ[km]) + dynZ[ikm]) / R(4);
          DcX[c] = (((DcZ[c] + DcZ[im]) + DcZ[km]) + DcZ[ikm]) / R(4);
        }
        T0X[c] = (((striniX[c] + striniX[im]) + striniX[km]) + striniX[ikm]) / R(4);
      }
  }

  // fd3d() prologue (fd3d_theo.f90:53-88)
  void init() {
    for (auto *a : {&u, &v, &w, &xx, &yy, &zz, &xy, &yz, &xz}) std::fill(a->begin(), a->end(), R(0));
    for (auto *a : {&tx, &tz, &v1t, &avdx, &avdz, &RFx, &RFz, &au1, &av1, &aw1, &uZ, &wX,
                    &tabsX, &tabsZ, &rise, &slipX, &slipZ, &schangeX, &schangeZ, &distX, &distZ})
      std::fill(a->begin(), a->end(), R(0));
    std::fill(ruptime.begin(), ruptime.end(), R(1.e4));
    std::fill(sliptime.begin(), sliptime.end(), R(1.e4));
    if (fvw) {
      std::fill(tabsX.begin(), tabsX.end(), R(1));
      std::fill(tabsZ.begin(), tabsZ.end(), R(1));
    }
    init_pml();
    interp_fric();
    it = 0;
  }

  // ------------------------------------------------------------------
  // dvel -> uxx1, vyy1, wzz1 (fd3d_deriv.f90:24-193), called with nyt-2
  // (fd3d_theo.f90:208): 4th-order interior velocities.
  // ------------------------------------------------------------------
  void dvel() {
    const R c1 = R(9) / R(8), c2 = R(-1) / R(24), dth = dt / dh;
    const int ie = nxt - nabc, je = nyt - 2, ke = nzt - nfs;
    const size_t sy = nx, sz = nxy;
#pragma omp parallel for schedule(static)
    for (int k = nabc + 1; k <= ke; ++k)
      for (int j = nabc + 1; j <= je; ++j)
        for (int i = nabc + 1; i <= ie; ++i) {
          size_t c = I3(i, j, k);
          R dd = d[c];
          u[c] = u[c] + (dth / dd) * (c1 * (xx[c] - xx[c - 1]) + c2 * (xx[c + 1] - xx[c - 2]) +
                                      c1 * (xy[c] - xy[c - sy]) + c2 * (xy[c + sy] - xy[c - 2 * sy]) +
                                      c1 * (xz[c] - xz[c - sz]) + c2 * (xz[c + sz] - xz[c - 2 * sz]));
        }
#pragma omp parallel for schedule(static)
    for (int k = nabc + 1; k <= ke; ++k)
      for (int j = nabc + 1; j <= je; ++j)
        for (int i = nabc + 1; i <= ie; ++i) {
          size_t c = I3(i, j, k);
          R dd = d[c];
          v[c] = v[c] + (dth / dd) * (c1 * (xy[c + 1] - xy[c]) + c2 * (xy[c + 2] - xy[c - 1]) +
                                      c1 * (yy[c + sy] - yy[c]) + c2 * (yy[c + 2 * sy] - yy[c - sy]) +
                                      c1 * (yz[c] - yz[c - sz]) + c2 * (yz[c + sz] - yz[c - 2 * sz]));
        }
#pragma omp parallel for schedule(static)
    for (int k = nabc + 1; k <= ke; ++k)
      for (int j = nabc + 1; j <= je; ++j)
        for (int i = nabc + 1; i <= ie; ++i) {
          size_t c = I3(i, j, k);
          R dd = d[c];
          w[c] = w[c] + (dth / dd) * (c1 * (xz[c + 1] - xz[c]) + c2 * (xz[c + 2] - xz[c - 1]) +
                                      c1 * (yz[c] - yz[c - sy]) + c2 * (yz[c + sy] - yz[c - 2 * sy]) +
                                      c1 * (zz[c + sz] - zz[c]) + c2 * (zz[c + 2 * sz] - zz[c - sz]));
        }
  }

  // ------------------------------------------------------------------
  // bnd2d -> uxx0a, vyy0a, wzz0a + v mirror (fd3d_deriv.f90:198-235,
  // 284-326, 375-417, 466-508): 2nd-order velocities at j = nyt-1 with the
  // Dalguer-Day damping state au1/av1/aw1, then v(:,nyt,:) = v(:,nyt-1,:)
  // over the whole plane (before the PML update of this half step).
  // ------------------------------------------------------------------
  void bnd2d() {
    const R dth = dt / dh;
    const int j = nyt - 1;
    const size_t sy = nx, sz = nxy;
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k), p = I2(i, k);
        R pdx = ((xx[c] - xx[c - 1]) + (xy[c] - xy[c - sy])) + (xz[c] - xz[c - sz]);
        au1[p] = pdx - au1[p];
        u[c] = u[c] + (dth / d[c]) * (pdx + damp_s * au1[p]);
        au1[p] = pdx;
      }
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k), p = I2(i, k);
        R pdy = ((xy[c + 1] - xy[c]) + (yy[c + sy] - yy[c])) + (yz[c] - yz[c - sz]);
        av1[p] = pdy - av1[p];
        v[c] = v[c] + (dth / d[c]) * (pdy + damp_s * av1[p]);
        av1[p] = pdy;
      }
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k), p = I2(i, k);
        R pdz = ((xz[c + 1] - xz[c]) + (yz[c] - yz[c - sy])) + (zz[c + sz] - zz[c]);
        aw1[p] = pdz - aw1[p];
        w[c] = w[c] + (dth / d[c]) * (pdz + damp_s * aw1[p]);
        aw1[p] = pdz;
      }
    for (int k = 1; k <= nzt; ++k)
      for (int i = 1; i <= nxt; ++i) v[I3(i, nyt, k)] = v[I3(i, nyt - 1, k)];
  }

  // ------------------------------------------------------------------
  // tasu1 / tasw1 (fd3d_deriv.f90:1109-1234): fault-plane u and w.
  // Slip weakening (:1164, :1228): explicit split-node update from the
  // previous step's traction.  Rate-and-state (:1141-1162, :1204-1226):
  // Newton solve for the slip-rate variable uw; the reference iterates
  // without bound while |F/F'| > 1e-10, here the loop stops after
  // newton_max iterations and the node is counted.  Note the reference's
  // abs() in tasu1's slip rate (:1143) and its absence in tasw1 (:1206).
  // ------------------------------------------------------------------
  void tasu1_tasw1() {
    const R dth = dt / dh;
    const int j = nyt;
    if (!fvw) {
      for (int k = nabc + 1; k <= nzt - nfs; ++k)
        for (int i = nabc + 1; i <= nxt - nabc; ++i) {
          size_t c = I3(i, j, k), p = I2(i, k);
          u[c] = u[c] + ((dth / d[c]) * R(2)) * (tx[p] + RFx[p]);
        }
      for (int k = nabc + 1; k <= nzt - nfs; ++k)
        for (int i = nabc + 1; i <= nxt - nabc; ++i) {
          size_t c = I3(i, j, k), p = I2(i, k);
          w[c] = w[c] + (dth / d[c]) * (R(2) * (tz[p] + RFz[p]));
        }
      return;
    }
    const F dthF = F(dt) / F(dh);
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k), p = I2(i, k);
        size_t pim = I2(i - 1, k), pkm = I2(i, k - 1), pikm = I2(i - 1, k - 1);
        F dd = F(d[c]);
        F uc = F(u[c]), wx = F(wX[p]);
        F sr = std::sqrt(sqF(F(2) * (std::fabs(wx) + std::fabs(wini))) +
                         sqF(F(2) * (std::fabs(uc) + std::fabs(uini))));
        F cdelta = ((F(2) * (dthF / dd)) * Sn) * aX[p];
        F uw = std::asinh((std::exp(psiX[p] / aX[p]) * sr) / (F(2) * v0));
        F rfz4 = (((F(RFz[p]) + F(RFz[pim])) + F(RFz[pkm])) + F(RFz[pikm])) / F(4);
        F vtilde = -std::sqrt(sqF(wx + (F(2) * (dthF / dd)) * (rfz4 - F(T0Z[p]))) +
                              sqF(uc + (F(2) * (dthF / dd)) * (F(RFx[p]) - F(T0X[p]))));
        F ex = std::exp(-psiX[p] / aX[p]);
        F ferr = F(10);
        int n = 0;
        while (std::fabs(ferr) > F(1e-10)) {
          if (n >= newton_max) { ++newton_capped; break; }
          ++n;
          F fwv = (vtilde + cdelta * uw) + (std::sinh(uw) * v0) * ex;
          F dfw = cdelta + (std::cosh(uw) * v0) * ex;
          ferr = fwv / dfw;
          uw = uw - ferr;
        }
        u[c] = R(((((F(tx[p]) + F(T0X[p])) * (-std::sinh(uw))) * v0) * ex) / F(tabsX[p]) + uini);
      }
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k), p = I2(i, k);
        size_t pip = I2(i + 1, k), pkp = I2(i, k + 1), pikp = I2(i + 1, k + 1);
        F dd = F(d[c]);
        F wc = F(w[c]), uz = F(uZ[p]);
        F sr = std::sqrt(sqF(F(2) * (wc - wini)) + sqF(F(2) * (uz - uini)));
        F cdelta = ((F(2) * (dthF / dd)) * Sn) * aZ[p];
        F uw = std::asinh((std::exp(psiZ[p] / aZ[p]) * sr) / (F(2) * v0));
        F rfx4 = (((F(RFx[p]) + F(RFx[pip])) + F(RFx[pkp])) + F(RFx[pikp])) / F(4);
        F t0x4 = (((F(T0X[p]) + F(T0X[pip])) + F(T0X[pkp])) + F(T0X[pikp])) / F(4);
        F vtilde = -std::sqrt(sqF(wc + (F(2) * (dthF / dd)) * (F(RFz[p]) - F(T0Z[p]))) +
                              sqF(uz + (F(2) * (dthF / dd)) * (rfx4 - t0x4)));
        F ex = std::exp(-psiZ[p] / aZ[p]);
        F ferr = F(10);
        int n = 0;
        while (std::fabs(ferr) > F(1e-10)) {
          if (n >= newton_max) { ++newton_capped; break; }
          ++n;
          F fwv = (vtilde + cdelta * uw) + (std::sinh(uw) * v0) * ex;
          F dfw = cdelta + (std::cosh(uw) * v0) * ex;
          ferr = fwv / dfw;
          uw = uw - ferr;
        }
        w[c] = R(((F(tz[p]) + F(T0Z[p])) * (-((std::sinh(uw) * v0) * ex))) / F(tabsZ[p]) + wini);
      }
  }

  // ------------------------------------------------------------------
  // pml_uvw (fd3d_deriv.f90:1927-2299): split-field velocities,
  // p <- [p(1-w) + (dt/rho)(dsigma)/dh]/(1+w), v = sum of splits.
  // u and w run to j = je_uw, v to je_v.
  // ------------------------------------------------------------------
  void pml_uvw() {
    const size_t sy = nx, sz = nxy;
    for (auto &r : reg) {
#pragma omp parallel for schedule(static)
      for (int k = r.kb; k <= r.ke; ++k)
        for (int j = r.jb; j <= r.je_uw; ++j)
          for (int i = r.ib; i <= r.ie; ++i) {
            int i2 = i - r.ib + 1, j2 = j - r.jb + 1, k2 = k - r.kb + 1;
            size_t c = I3(i, j, k), l = L(r, i2, j2, k2);
            R dd = d[c];
            R pt;
            R &u1s = r.s[SU1][l], &u2s = r.s[SU2][l], &u3s = r.s[SU3][l];
            pt = u1s * (R(1) - r.ox[i2]) + ((dt / dd) * (xx[c] - xx[c - 1])) / dh;
            u1s = pt / (R(1) + r.ox[i2]);
            pt = u2s * (R(1) - r.oy[j2]) + ((dt / dd) * (xy[c] - xy[c - sy])) / dh;
            u2s = pt / (R(1) + r.oy[j2]);
            pt = u3s * (R(1) - r.oz[k2]) + ((dt / dd) * (xz[c] - xz[c - sz])) / dh;
            u3s = pt / (R(1) + r.oz[k2]);
            u[c] = (u1s + u2s) + u3s;
            R &w1s = r.s[SW1][l], &w2s = r.s[SW2][l], &w3s = r.s[SW3][l];
            pt = w1s * (R(1) - r.oxS[i2]) + ((dt / dd) * (xz[c + 1] - xz[c])) / dh;
            w1s = pt / (R(1) + r.oxS[i2]);
            pt = w2s * (R(1) - r.oy[j2]) + ((dt / dd) * (yz[c] - yz[c - sy])) / dh;
            w2s = pt / (R(1) + r.oy[j2]);
            pt = w3s * (R(1) - r.oz[k2]) + ((dt / dd) * (zz[c + sz] - zz[c])) / dh;
            w3s = pt / (R(1) + r.oz[k2]);
            w[c] = (w1s + w2s) + w3s;
            if (j <= r.je_v) {
              R &v1s = r.s[SV1][l], &v2s = r.s[SV2][l], &v3s = r.s[SV3][l];
              pt = v1s * (R(1) - r.oxS[i2]) + ((dt / dd) * (xy[c + 1] - xy[c])) / dh;
              v1s = pt / (R(1) + r.oxS[i2]);
              pt = v2s * (R(1) - r.oy[j2]) + ((dt / dd) * (yy[c + sy] - yy[c])) / dh;
              v2s = pt / (R(1) + r.oy[j2]);
              pt = v3s * (R(1) - r.oz[k2]) + ((dt / dd) * (yz[c] - yz[c - sz])) / dh;
              v3s = pt / (R(1) + r.oz[k2]);
              v[c] = (v1s + v2s) + v3s;
            }
          }
    }
  }

  // ------------------------------------------------------------------
  // fuvw (fd3d_deriv.f90:1018-1062), free surface only: ghost velocities
  // at ks+1 = nzt-nfs+1.
  // ------------------------------------------------------------------
  void fuvw() {
    const int ks = nzt - nfs;
    for (int j = 1; j <= nyt; ++j)
      for (int i = 2; i <= nxt - 1; ++i)
        u[I3(i, j, ks + 1)] = u[I3(i, j, ks)] - (w[I3(i, j, ks)] - w[I3(i - 1, j, ks)]);
    for (int j = 2; j <= nyt - 1; ++j)
      for (int i = 2; i <= nxt - 1; ++i)
        v[I3(i, j, ks + 1)] = v[I3(i, j, ks)] - (w[I3(i, j + 1, ks)] - w[I3(i, j, ks)]);
    for (int j = 2; j <= nyt; ++j)
      for (int i = 2; i <= nxt - 2; ++i) {
        R xl = lam[I3(i, j, ks + 1)], xm = mu[I3(i, j, ks + 1)];
        R a = R(2) * xl, b = xl + R(2) * xm;
        R s = ((((((u[I3(i + 1, j, ks)] - u[I3(i, j, ks)]) + u[I3(i + 1, j, ks + 1)]) -
                  u[I3(i, j, ks + 1)]) + v[I3(i, j, ks)]) - v[I3(i, j - 1, ks)]) +
               v[I3(i, j, ks + 1)]) - v[I3(i, j - 1, ks + 1)];
        w[I3(i, j, ks + 1)] = w[I3(i, j, ks - 1)] - (a / b) * s;
      }
  }

  // ------------------------------------------------------------------
  // dstres (fd3d_deriv.f90:513-622), called with nyt-2
  // (fd3d_theo.f90:220): 4th-order interior stresses.  Note the diagonal
  // mu(i+1,j+1,k+1) in the yz average (:603).
  // ------------------------------------------------------------------
  void dstres() {
    const R c1 = R(9) / R(8), c2 = R(-1) / R(24), dth = dt / dh;
    const size_t sy = nx, sz = nxy;
#pragma omp parallel for schedule(static)
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int j = nabc + 1; j <= nyt - 2; ++j)
        for (int i = nabc + 1; i <= nxt - nabc; ++i) {
          size_t c = I3(i, j, k);
          R xl = lam[c], xm = mu[c];
          R a = xl + R(2) * xm, b = xl;
          R dux = c1 * (u[c + 1] - u[c]) + c2 * (u[c + 2] - u[c - 1]);
          R dvy1 = c1 * (v[c] - v[c - sy]), dvy2 = c2 * (v[c + sy] - v[c - 2 * sy]);
          R dwz1 = c1 * (w[c] - w[c - sz]), dwz2 = c2 * (w[c + sz] - w[c - 2 * sz]);
          R dux1 = c1 * (u[c + 1] - u[c]), dux2 = c2 * (u[c + 2] - u[c - 1]);
          xx[c] = (xx[c] + (dth * a) * dux) + (dth * b) * (((dvy1 + dvy2) + dwz1) + dwz2);
          yy[c] = (yy[c] + (dth * a) * (dvy1 + dvy2)) + (dth * b) * (((dux1 + dux2) + dwz1) + dwz2);
          zz[c] = (zz[c] + (dth * a) * (dwz1 + dwz2)) + (dth * b) * (((dux1 + dux2) + dvy1) + dvy2);
          R xmu = R(0.5) * (mu[c] + mu[c + sy]);
          xy[c] = xy[c] + (dth * xmu) * (((c1 * (u[c + sy] - u[c]) + c2 * (u[c + 2 * sy] - u[c - sy])) +
                                          c1 * (v[c] - v[c - 1])) + c2 * (v[c + 1] - v[c - 2]));
          xmu = R(0.5) * (mu[c] + mu[c + sz]);
          xz[c] = xz[c] + (dth * xmu) * (((c1 * (u[c + sz] - u[c]) + c2 * (u[c + 2 * sz] - u[c - sz])) +
                                          c1 * (w[c] - w[c - 1])) + c2 * (w[c + 1] - w[c - 2]));
          xmu = R(0.5) * (mu[c] + mu[c + 1 + sy + sz]);
          yz[c] = yz[c] + (dth * xmu) * (((c1 * (v[c + sz] - v[c]) + c2 * (v[c + 2 * sz] - v[c - sz])) +
                                          c1 * (w[c + sy] - w[c])) + c2 * (w[c + 2 * sy] - w[c - sy]));
        }
  }

  // ------------------------------------------------------------------
  // strbnd -> sxx, sxy0, sxz0, syz0 + mirror (fd3d_deriv.f90:777-1012):
  // 2nd-order stresses at j = nyt-1; the shear kernels use dth = .5*dt/dh
  // with an unaveraged mu sum (:902,947,991).  Then xy and yz at nyt copy
  // nyt-1 over the whole plane, before the PML update.
  // ------------------------------------------------------------------
  void strbnd() {
    const int j = nyt - 1;
    const size_t sy = nx, sz = nxy;
    R dth = dt / dh;
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k);
        R xl = lam[c], xm = mu[c];
        R a = xl + R(2) * xm, b = xl;
        xx[c] = (xx[c] + (dth * a) * (u[c + 1] - u[c])) +
                (dth * b) * (((v[c] - v[c - sy]) + w[c]) - w[c - sz]);
        yy[c] = (yy[c] + (dth * a) * (v[c] - v[c - sy])) +
                (dth * b) * (((u[c + 1] - u[c]) + w[c]) - w[c - sz]);
        zz[c] = (zz[c] + (dth * a) * (w[c] - w[c - sz])) +
                (dth * b) * (((u[c + 1] - u[c]) + v[c]) - v[c - sy]);
      }
    dth = (R(0.5) * dt) / dh;
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k);
        R xmu = mu[c] + mu[c + sy];
        xy[c] = xy[c] + (dth * xmu) * (((u[c + sy] - u[c]) + v[c]) - v[c - 1]);
      }
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k);
        R xmu = mu[c] + mu[c + sz];
        xz[c] = xz[c] + (dth * xmu) * (((u[c + sz] - u[c]) + w[c]) - w[c - 1]);
      }
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k);
        R xmu = mu[c] + mu[c + 1 + sy + sz];
        yz[c] = yz[c] + (dth * xmu) * (((v[c + sz] - v[c]) + w[c + sy]) - w[c]);
      }
    for (int k = 1; k <= nzt; ++k)
      for (int i = 1; i <= nxt; ++i) {
        xy[I3(i, nyt, k)] = xy[I3(i, nyt - 1, k)];
        yz[I3(i, nyt, k)] = yz[I3(i, nyt - 1, k)];
      }
  }

  // ------------------------------------------------------------------
  // tasxz -> sxz1 (fd3d_deriv.f90:1239-1258, 677-722) and tasii
  // (:1263-1314): fault-plane xz (4th order, dth = .5*dt/dh, mu sum) and
  // the normal stresses from the extrapolated fault-normal velocity v1t.
  // ------------------------------------------------------------------
  void tasxz_tasii() {
    const R c1 = R(9) / R(8), c2 = R(-1) / R(24);
    const int j = nyt;
    const size_t sy = nx, sz = nxy;
    R dth = (R(0.5) * dt) / dh;
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k);
        R xmu = mu[c] + mu[c + sz];
        xz[c] = xz[c] + (dth * xmu) * (((c1 * (u[c + sz] - u[c]) + c2 * (u[c + 2 * sz] - u[c - sz])) +
                                        c1 * (w[c] - w[c - 1])) + c2 * (w[c + 1] - w[c - 2]));
      }
    dth = dt / dh;
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k), p = I2(i, k);
        R xl = lam[c], xm = mu[c];
        R a = xl + R(2) * xm, b = xl;
        R diff1 = c1 * (u[c + 1] - u[c]) + c2 * (u[c + 2] - u[c - 1]);
        R diff3 = c1 * (w[c] - w[c - sz]) + c2 * (w[c + sz] - w[c - 2 * sz]);
        R vm = v[c - sy];
        v1t[p] = vm - (b * (diff1 + diff3)) / (R(2) * a);
        R dv = v1t[p] - vm;
        xx[c] = (xx[c] + (dth * a) * diff1) + (dth * b) * (R(2) * dv + diff3);
        yy[c] = (yy[c] + ((dth * a) * R(2)) * dv) + (dth * b) * (diff1 + diff3);
        zz[c] = (zz[c] + (dth * a) * diff3) + (dth * b) * (diff1 + R(2) * dv);
      }
  }

  // ------------------------------------------------------------------
  // pml_xyz (fd3d_deriv.f90:2304-3145): split-field stresses.  The xz/yz
  // splits of the top layer k = nzt-nfs are zeroed after the update in
  // region 1 always (:2469-2491) and in regions 2 and 3 only without
  // FSPACE (:2646-2670, :2823-2847).
  // ------------------------------------------------------------------
  void pml_xyz() {
    const size_t sy = nx, sz = nxy;
    for (auto &r : reg) {
#pragma omp parallel for schedule(static)
      for (int k = r.kb; k <= r.ke; ++k)
        for (int j = r.jb; j <= r.je_uw; ++j)
          for (int i = r.ib; i <= r.ie; ++i) {
            int i2 = i - r.ib + 1, j2 = j - r.jb + 1, k2 = k - r.kb + 1;
            size_t c = I3(i, j, k), l = L(r, i2, j2, k2);
            R pt;
            R xl = lam[c], xm = mu[c];
            R a = xl + R(2) * xm, b = xl;
            R diff1 = u[c + 1] - u[c], diff2 = v[c] - v[c - sy], diff3 = w[c] - w[c - sz];
            R ax = R(1) - r.oxS[i2], bx = R(1) + r.oxS[i2];
            R ay = R(1) - r.oy[j2], by = R(1) + r.oy[j2];
            R az = R(1) - r.oz[k2], bz = R(1) + r.oz[k2];
            R *s = nullptr;
            (void)s;
            R &x1 = r.s[SXX1][l], &x2 = r.s[SXX2][l], &x3 = r.s[SXX3][l];
            pt = x1 * ax + ((dt * a) * diff1) / dh; x1 = pt / bx;
            pt = x2 * ay + ((dt * b) * diff2) / dh; x2 = pt / by;
            pt = x3 * az + ((dt * b) * diff3) / dh; x3 = pt / bz;
            xx[c] = (x1 + x2) + x3;
            R &y1 = r.s[SYY1][l], &y2 = r.s[SYY2][l], &y3 = r.s[SYY3][l];
            pt = y1 * ax + ((dt * b) * diff1) / dh; y1 = pt / bx;
            pt = y2 * ay + ((dt * a) * diff2) / dh; y2 = pt / by;
            pt = y3 * az + ((dt * b) * diff3) / dh; y3 = pt / bz;
            yy[c] = (y1 + y2) + y3;
            R &z1 = r.s[SZZ1][l], &z2 = r.s[SZZ2][l], &z3 = r.s[SZZ3][l];
            pt = z1 * ax + ((dt * b) * diff1) / dh; z1 = pt / bx;
            pt = z2 * ay + ((dt * b) * diff2) / dh; z2 = pt / by;
            pt = z3 * az + ((dt * a) * diff3) / dh; z3 = pt / bz;
            zz[c] = (z1 + z2) + z3;
            if (j <= r.je_v) {
              R xmu = (mu[c] + mu[c + sy]) / R(2);
              R &p1 = r.s[SXY1][l], &p2 = r.s[SXY2][l];
              pt = p1 * (R(1) - r.oyS[j2]) + ((dt * xmu) * (u[c + sy] - u[c])) / dh;
              p1 = pt / (R(1) + r.oyS[j2]);
              pt = p2 * (R(1) - r.ox[i2]) + ((dt * xmu) * (v[c] - v[c - 1])) / dh;
              p2 = pt / (R(1) + r.ox[i2]);
              xy[c] = p1 + p2;
            }
            if (j <= r.je_xz) {
              R xmu = (mu[c] + mu[c + sz]) / R(2);
              R &p1 = r.s[SXZ1][l], &p2 = r.s[SXZ2][l];
              pt = p1 * (R(1) - r.ozS[k2]) + ((dt * xmu) * (u[c + sz] - u[c])) / dh;
              p1 = pt / (R(1) + r.ozS[k2]);
              pt = p2 * (R(1) - r.ox[i2]) + ((dt * xmu) * (w[c] - w[c - 1])) / dh;
              p2 = pt / (R(1) + r.ox[i2]);
              xz[c] = p1 + p2;
            }
            if (j <= r.je_v) {
              R xmu = (mu[c] + mu[c + 1 + sy + sz]) / R(2);
              R &p1 = r.s[SYZ1][l], &p2 = r.s[SYZ2][l];
              pt = p1 * (R(1) - r.ozS[k2]) + ((dt * xmu) * (v[c + sz] - v[c])) / dh;
              p1 = pt / (R(1) + r.ozS[k2]);
              pt = p2 * (R(1) - r.oyS[j2]) + ((dt * xmu) * (w[c + sy] - w[c])) / dh;
              p2 = pt / (R(1) + r.oyS[j2]);
              yz[c] = p1 + p2;
            }
          }
      if (r.ztop) {
        for (int j2 = 1; j2 <= r.nj_zero_xz; ++j2)
          for (int i2 = 1; i2 <= r.ni; ++i2) {
            size_t l = L(r, i2, j2, r.nk);
            r.s[SXZ1][l] = R(0);
            r.s[SXZ2][l] = R(0);
          }
        for (int j2 = 1; j2 <= r.nj_zero_yz; ++j2)
          for (int i2 = 1; i2 <= r.ni; ++i2) {
            size_t l = L(r, i2, j2, r.nk);
            r.s[SYZ1][l] = R(0);
            r.s[SYZ2][l] = R(0);
          }
      }
    }
  }

  // fres (fd3d_deriv.f90:1067-1104), free surface only: stress imaging.
  void fres() {
    const int ks = nzt - nfs;
    for (int j = 1; j <= nyt; ++j)
      for (int i = 1; i <= nxt; ++i) {
        zz[I3(i, j, ks + 1)] = -zz[I3(i, j, ks)];
        zz[I3(i, j, ks + 2)] = -zz[I3(i, j, ks - 1)];
        xz[I3(i, j, ks + 1)] = -xz[I3(i, j, ks - 1)];
        xz[I3(i, j, ks + 2)] = -xz[I3(i, j, ks - 2)];
        yz[I3(i, j, ks + 1)] = -yz[I3(i, j, ks - 1)];
        yz[I3(i, j, ks + 2)] = -yz[I3(i, j, ks - 2)];
        xz[I3(i, j, ks)] = R(0);
        yz[I3(i, j, ks)] = R(0);
      }
  }

  // ------------------------------------------------------------------
  // Trial traction (fd3d_theo.f90:232-309): 4th-order in-plane force,
  // Dalguer-Day damping RF = pd + damp_s*(pd - pd_prev),
  // t = -RF - 0.5*rho*(dh/dt)*v_fault.  The top row k = nzt-nfs uses the
  // same stencil with FSPACE (:260-280) and the free-surface stencil
  // otherwise (:282-308), followed by the ghost-row traction mirror.
  // ------------------------------------------------------------------
  void traction() {
    const R c1 = R(9) / R(8), c2 = R(-1) / R(24), dht = dh / dt;
    const int j = nyt;
    const size_t sy = nx, sz = nxy;
    const int ks = nzt - nfs;
    const int kend = fspace ? ks : ks - 1;
    for (int k = nabc + 1; k <= kend; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t c = I3(i, j, k), p = I2(i, k);
        R pdz = ((c1 * (xz[c + 1] - xz[c]) + c2 * (xz[c + 2] - xz[c - 1])) / R(2) +
                 (c1 * (zz[c + sz] - zz[c]) + c2 * (zz[c + 2 * sz] - zz[c - sz])) / R(2)) -
                yz[c - sy];
        avdz[p] = damp_s * (pdz - avdz[p]);
        RFz[p] = pdz + avdz[p];
        tz[p] = -RFz[p] - ((R(0.5) * d[c]) * dht) * w[c];
        avdz[p] = pdz;
        R pdx = ((c1 * (xx[c] - xx[c - 1]) + c2 * (xx[c + 1] - xx[c - 2])) / R(2) +
                 (c1 * (xz[c] - xz[c - sz]) + c2 * (xz[c + sz] - xz[c - 2 * sz])) / R(2)) -
                xy[c - sy];
        avdx[p] = damp_s * (pdx - avdx[p]);
        RFx[p] = pdx + avdx[p];
        tx[p] = -RFx[p] - ((R(0.5) * d[c]) * dht) * u[c];
        avdx[p] = pdx;
      }
    if (fspace) return;
    const int k = ks;
    for (int i = nabc + 1; i <= nxt - nabc; ++i) {
      size_t c = I3(i, j, k), p = I2(i, k);
      R pdz = ((xz[c + 1 - sz] - xz[c - sz]) / R(8) + (zz[c + sz] - zz[c]) / R(4)) - yz[c - sy - sz] / R(4);
      avdz[p] = damp_s * (pdz - avdz[p]);
      RFz[p] = pdz + avdz[p];
      tz[p] = -RFz[p] - ((((R(0.5) * d[c]) * dht) * w[c]) / R(2));
      avdz[p] = pdz;
      R pdx = ((c1 * (xx[c] - xx[c - 1]) + c2 * (xx[c + 1] - xx[c - 2])) / R(2) +
               (c1 * (xz[c] - xz[c - sz]) + c2 * (xz[c + sz] - xz[c - 2 * sz])) / R(2)) -
              xy[c - sy];
      avdx[p] = damp_s * (pdx - avdx[p]);
      RFx[p] = pdx + avdx[p];
      tx[p] = -RFx[p] - ((R(0.5) * d[c]) * dht) * u[c];
      avdx[p] = pdx;
    }
    for (int i = nabc + 1; i <= nxt - nabc; ++i) {
      tx[I2(i, nzt - 1)] = tx[I2(i, nzt - 2)];
      tz[I2(i, nzt - 1)] = -tz[I2(i, nzt - 3)];
    }
  }

  // |T| at Z- and X-nodes by 4-point interpolation (fd3d_theo.f90:313-324)
  void tabs() {
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t p = I2(i, k), pi = I2(i + 1, k), pk = I2(i, k + 1), pik = I2(i + 1, k + 1);
        R tint = (((((((tx[p] + T0X[p]) + tx[pi]) + T0X[pi]) + tx[pk]) + T0X[pk]) + tx[pik]) + T0X[pik]) / R(4);
        tabsZ[p] = std::sqrt(sq(tint) + sq(tz[p] + T0Z[p]));
        size_t qi = I2(i - 1, k), qk = I2(i, k - 1), qik = I2(i - 1, k - 1);
        tint = (((((((tz[p] + T0Z[p]) + tz[qi]) + T0Z[qi]) + tz[qk]) + T0Z[qk]) + tz[qik]) + T0Z[qik]) / R(4);
        tabsX[p] = std::sqrt(sq(tx[p] + T0X[p]) + sq(tint));
      }
  }

  // ------------------------------------------------------------------
  // Friction at Z-nodes (fd3d_theo.f90:326-373) and X-nodes (:375-410).
  // Slip weakening: strength S(d); if |T| >= S the traction is scaled
  // back to S and the slip path grows; rupture time at slip rate > 1e-3.
  // Rate-and-state: exact exponential state update and S = Sn*a*asinh.
  // The X-node slip-weakening test compares distX with Dc but divides by
  // DcX (:396-397), as the reference does.
  // ------------------------------------------------------------------
  void friction(R time) {
    const R rup_tresh = R(1.e-3);
    const size_t sy = nx;
    (void)sy;
    const int j = nyt;
    const F dtF = F(dt);
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t p = I2(i, k);
        R tab = tabsZ[p];
        R U00 = u[I3(i, j, k)], U10 = u[I3(i + 1, j, k)], U01 = u[I3(i, j, k + 1)], U11 = u[I3(i + 1, j, k + 1)];
        R W = w[I3(i, j, k)];
        uZ[p] = (((U00 + U10) + U01) + U11) / R(4);
        R u1out = -std::sqrt(sq(W) + sq(uZ[p]));
        if (fvw) {
          F uz = F(uZ[p]), wf = F(W);
          F sr = std::sqrt(sqF(F(2) * (std::fabs(wf) + std::fabs(wini))) +
                           sqF(F(2) * (std::fabs(uz) + std::fabs(uini))));
          F flv = f0 - (bZ[p] - aZ[p]) * std::log(sr / v0);
          F q = sr / vwZ[p];
          F q8 = ((((((q * q) * q) * q) * q) * q) * q) * q;
          F fss = fw + (flv - fw) / std::pow(F(1) + q8, F(1) / F(8));
          F psiss = aZ[p] * (std::log(std::sinh(fss / aZ[p])) + std::log((F(2) * v0) / sr));
          psiZ[p] = (psiZ[p] - psiss) * std::exp(-((sr * dtF) / F(Dc[p]))) + psiss;
          F fric = (Sn * aZ[p]) * std::asinh((sr * std::exp(psiZ[p] / aZ[p])) / (F(2) * v0));
          distZ[p] = distZ[p] - (R(2) * u1out) * dt;
          schangeZ[p] = R(fric);
          psiout[p] = psiZ[p];
          if (std::fabs(R(2) * u1out) > rup_tresh) {
            if (ruptime[p] != R(1.e4)) rise[p] = time;
            if (ruptime[p] == R(1.e4)) ruptime[p] = time;
          }
        } else {
          R fric;
          if (distZ[p] <= DcZ[p])
            fric = (peakZ[p] * (R(1) - distZ[p] / DcZ[p]) + (dynZ[p] * distZ[p]) / DcZ[p]) + coh[p];
          else
            fric = dynZ[p] + coh[p];
          if (tab >= fric) {
            distZ[p] = distZ[p] - (R(2) * u1out) * dt;
            tz[p] = ((tz[p] + T0Z[p]) * fric) / tab - T0Z[p];
            if (R(-2) * u1out > rup_tresh) {
              if (ruptime[p] != R(1.e4)) rise[p] = time;
              if (ruptime[p] == R(1.e4)) ruptime[p] = time;
            }
          }
        }
        slipZ[p] = slipZ[p] - (R(2) * W) * dt;
        slipX[p] = slipX[p] - (R(2) * uZ[p]) * dt;
        if ((sliptime[p] == R(1.e4)) && (distZ[p] > Dc[p])) sliptime[p] = time;
      }
    for (int k = nabc + 1; k <= nzt - nfs; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t p = I2(i, k);
        R tab = tabsX[p];
        R W00 = w[I3(i, j, k)], W10 = w[I3(i - 1, j, k)], W01 = w[I3(i, j, k - 1)], W11 = w[I3(i - 1, j, k - 1)];
        R U = u[I3(i, j, k)];
        wX[p] = (((W00 + W10) + W01) + W11) / R(4);
        R u1out = -std::sqrt(sq(wX[p]) + sq(U));
        if (fvw) {
          F wx = F(wX[p]), uf = F(U);
          F sr = std::sqrt(sqF(F(2) * (std::fabs(wx) + std::fabs(wini))) +
                           sqF(F(2) * (std::fabs(uf) + std::fabs(uini))));
          F flv = f0 - (bX[p] - aX[p]) * std::log(sr / v0);
          F q = sr / vwX[p];
          F q8 = ((((((q * q) * q) * q) * q) * q) * q) * q;
          F fss = fw + (flv - fw) / std::pow(F(1) + q8, F(1) / F(8));
          F psiss = aX[p] * (std::log(std::sinh(fss / aX[p])) + std::log((F(2) * v0) / sr));
          psiX[p] = (psiX[p] - psiss) * std::exp(-((sr * dtF) / F(Dc[p]))) + psiss;
          F fric = (Sn * aX[p]) * std::asinh((sr * std::exp(psiX[p] / aX[p])) / (F(2) * v0));
          distX[p] = distX[p] - (R(2) * u1out) * dt;
          schangeX[p] = R(fric);
        } else {
          R fric;
          if (distX[p] <= Dc[p])
            fric = (peakX[p] * (R(1) - distX[p] / DcX[p]) + (dynX[p] * distX[p]) / DcX[p]) + coh[p];
          else
            fric = dynX[p] + coh[p];
          if (tab >= fric) {
            distX[p] = distX[p] - (R(2) * u1out) * dt;
            tx[p] = ((tx[p] + T0X[p]) * fric) / tab - T0X[p];
          }
        }
      }
  }

  // ------------------------------------------------------------------
  // Per-step outputs (fd3d_theo.f90:413-453): all at Z-nodes.  Slip rate
  // is -2x the fault velocity (half-space symmetry).  The top row uses
  // 2-point forms; the rate-and-state top row divides by 4 (:442).
  // ------------------------------------------------------------------
  void outputs() {
    const int j = nyt, ks = nzt - nfs;
    for (int k = nabc + 1; k <= ks - 1; ++k)
      for (int i = nabc + 1; i <= nxt - nabc; ++i) {
        size_t p = I2(i, k);
        R U00 = u[I3(i, j, k)], U10 = u[I3(i + 1, j, k)], U01 = u[I3(i, j, k + 1)], U11 = u[I3(i + 1, j, k + 1)];
        if (!fvw) {
          size_t pi = I2(i + 1, k), pk = I2(i, k + 1), pik = I2(i + 1, k + 1);
          schangeZ[p] = tz[p] + T0Z[p];
          schangeX[p] = (((((((tx[p] + T0X[p]) + tx[pi]) + T0X[pi]) + tx[pk]) + T0X[pk]) + tx[pik]) + T0X[pik]) / R(4);
        }
        srZ[p] = -(R(2) * w[I3(i, j, k)]);
        srX[p] = ((((-(R(2) * U00)) - R(2) * U10) - R(2) * U01) - R(2) * U11) / R(4);
      }
    for (int i = nabc + 1; i <= nxt - nabc; ++i) {
      size_t p = I2(i, ks);
      R U00 = u[I3(i, j, ks)], U10 = u[I3(i + 1, j, ks)];
      if (!fvw) {
        size_t pi = I2(i + 1, ks);
        schangeZ[p] = tz[p] + T0Z[p];
        schangeX[p] = (((tx[p] + T0X[p]) + tx[pi]) + T0X[pi]) / R(2);
        srX[p] = ((-(R(2) * U00)) - R(2) * U10) / R(2);
      } else {
        srX[p] = ((-(R(2) * U00)) - R(2) * U10) / R(4);
      }
      srZ[p] = -(R(2) * w[I3(i, j, ks)]);
    }
  }

  // TPV104 smooth nucleation (fd3d_theo.f90:515-542): T0X ramps in time.
  void nucleation(R time) {
    F t = F(time) - F(dt) / F(2);
    if (!(t <= TT2)) return;
    F GT;
    if (t < TT2)
      GT = std::exp(sqF(t - TT2) / (t * (t - F(2) * TT2)));
    else
      GT = F(1);
    const F dhF = F(dh);
    for (int k = nabc - 1; k <= nzt - nfs; ++k)
      for (int i = nabc - 1; i <= nxt - nabc; ++i) {
        F hx = F(i) * dhF, hz = F(k) * dhF;
        F rr = sqF(hx - hx0) + sqF(hz - hz0);
        F FXZ;
        if (rr < RR2 - dhF)
          FXZ = std::exp(rr / ((rr - RR2) - dhF));
        else
          FXZ = F(0);
        T0X[I2(i, k)] = R(strinixI + (perturb * FXZ) * GT);
      }
  }

  // one time step in the reference order (fd3d_theo.f90:187-558)
  void step(double *out, double *seis) {
    ++it;
    R time = R(it) * dt;
    std::fill(srX.begin(), srX.end(), R(0));
    std::fill(srZ.begin(), srZ.end(), R(0));
    std::fill(schangeZ.begin(), schangeZ.end(), R(0));
    std::fill(schangeX.begin(), schangeX.end(), R(0));
    if (fvw) std::fill(psiout.begin(), psiout.end(), F(0));
    // velocity half step (:208-215)
    dvel();
    bnd2d();
    tasu1_tasw1();
    pml_uvw();
    if (!fspace) fuvw();
    // stress half step (:220-227)
    dstres();
    strbnd();
    tasxz_tasii();
    pml_xyz();
    if (!fspace) fres();
    // fault boundary condition (:232-453)
    traction();
    tabs();
    friction(time);
    outputs();
    if (seis) {
      for (int s = 0; s < nsta; ++s) {
        size_t c = I3(sta[3 * s], sta[3 * s + 1], sta[3 * s + 2]);
        seis[3 * s] = double(u[c]);
        seis[3 * s + 1] = double(v[c]);
        seis[3 * s + 2] = double(w[c]);
      }
    }
    if (fvw && nucl) nucleation(time);
    if (out) {
      const int nxT = nxt - 2 * nabc, nzT = nzt - nfs - nabc;
      const size_t slab = (size_t)nxT * nzT;
      for (int o = 0; o < ORC_NOUT; ++o) {
        double *dst = out + (size_t)o * slab;
        for (int k = nabc + 1; k <= nzt - nfs; ++k)
          for (int i = nabc + 1; i <= nxt - nabc; ++i) {
            size_t p = I2(i, k);
            double val = 0.0;
            switch (o) {
              case ORC_OUT_SRX: val = double(srX[p]); break;
              case ORC_OUT_SRZ: val = double(srZ[p]); break;
              case ORC_OUT_SCX: val = double(schangeX[p]); break;
              case ORC_OUT_SCZ: val = double(schangeZ[p]); break;
              default: val = fvw ? double(psiout[p]) : 0.0; break;
            }
            dst[(size_t)(i - nabc - 1) + (size_t)nxT * (size_t)(k - nabc - 1)] = val;
          }
      }
    }
  }

  std::vector<R> *field(int which) {
    switch (which) {
      case ORC_U: return &u; case ORC_V: return &v; case ORC_W: return &w;
      case ORC_XX: return &xx; case ORC_YY: return &yy; case ORC_ZZ: return &zz;
      case ORC_XY: return &xy; case ORC_YZ: return &yz; case ORC_XZ: return &xz;
      default: return nullptr;
    }
  }
};

struct Handle {
  int prec;
  void *p;
};

template <typename R, typename F>
Oracle<R, F> *make(const int *ip, const double *rp, const int *sta) {
  auto *o = new Oracle<R, F>();
  o->nxt = ip[ORC_NXT]; o->nyt = ip[ORC_NYT]; o->nzt = ip[ORC_NZT]; o->nabc = ip[ORC_NABC];
  o->fspace = ip[ORC_FSPACE]; o->fvw = ip[ORC_FVW]; o->nucl = ip[ORC_NUCL];
  o->nsta = ip[ORC_NSTA];
  o->newton_max = ip[ORC_NEWTON_MAX] > 0 ? ip[ORC_NEWTON_MAX] : 200;
  o->nfs = o->fspace ? o->nabc : 2;  // fd3d_init.f90:161-165
  o->dh = R(rp[ORC_DH]); o->dt = R(rp[ORC_DT]); o->damp_s = R(rp[ORC_DAMP_S]);
  o->pml_vp = R(rp[ORC_PML_VP]); o->pml_fact = R(rp[ORC_PML_FACT]);
  o->Sn = F(rp[ORC_SN]); o->uini = F(rp[ORC_UINI]); o->wini = F(rp[ORC_WINI]);
  o->f0 = F(rp[ORC_F0]); o->v0 = F(rp[ORC_V0]); o->fw = F(rp[ORC_FW]);
  o->hx0 = F(rp[ORC_HX0]); o->hz0 = F(rp[ORC_HZ0]); o->perturb = F(rp[ORC_PERTURB]);
  o->RR2 = F(rp[ORC_RR2]); o->TT2 = F(rp[ORC_TT2]); o->strinixI = F(rp[ORC_STRINIXI]);
  o->sta.assign(sta ? sta : (const int *)nullptr, sta ? sta + 3 * o->nsta : (const int *)nullptr);
  o->alloc();
  return o;
}

template <typename R, typename F, typename Fn>
auto dispatch(void *h, Fn fn) -> decltype(fn(static_cast<Oracle<float, float> *>(nullptr))) {
  Handle *hh = static_cast<Handle *>(h);
  switch (hh->prec) {
    case ORC_F32: return fn(static_cast<Oracle<float, float> *>(hh->p));
    case ORC_F64: return fn(static_cast<Oracle<double, double> *>(hh->p));
    default: return fn(static_cast<Oracle<float, double> *>(hh->p));
  }
}

template <typename V>
void load(V &dst, const double *src) {
  using T = typename V::value_type;
  for (size_t n = 0; n < dst.size(); ++n) dst[n] = T(src[n]);
}
template <typename V>
void store(const V &src, double *dst) {
  for (size_t n = 0; n < src.size(); ++n) dst[n] = double(src[n]);
}

}  // namespace

extern "C" {

void *orc_create(int prec, const int *ip, const double *rp, const int *sta) {
  if (ip[ORC_NABC] < 3 || ip[ORC_NXT] < 2 * ip[ORC_NABC] + 4) return nullptr;
  Handle *h = new Handle{prec, nullptr};
  if (prec == ORC_F32) h->p = make<float, float>(ip, rp, sta);
  else if (prec == ORC_F64) h->p = make<double, double>(ip, rp, sta);
  else h->p = make<float, double>(ip, rp, sta);
  return h;
}

void orc_destroy(void *h) {
  if (!h) return;
  dispatch<float, float>(h, [](auto *o) { delete o; return 0; });
  delete static_cast<Handle *>(h);
}

int orc_set_medium(void *h, const double *lam, const double *mu, const double *d) {
  return dispatch<float, float>(h, [&](auto *o) {
    load(o->lam, lam); load(o->mu, mu); load(o->d, d); return 0;
  });
}

int orc_set_fault_sw(void *h, const double *striniX, const double *striniZ, const double *peak_xz,
                     const double *dyn_xz, const double *Dc, const double *coh) {
  return dispatch<float, float>(h, [&](auto *o) {
    load(o->striniX, striniX); load(o->striniZ, striniZ); load(o->peak_xz, peak_xz);
    load(o->dyn_xz, dyn_xz); load(o->Dc, Dc); load(o->coh, coh); return 0;
  });
}

int orc_set_fault_fvw(void *h, const double *Dc, const double *striniX, const double *striniZ,
                      const double *a, const double *b, const double *psi, const double *vw,
                      const double *aX, const double *bX, const double *psiX, const double *vwX) {
  return dispatch<float, float>(h, [&](auto *o) {
    load(o->Dc, Dc); load(o->striniX, striniX); load(o->striniZ, striniZ);
    load(o->fa, a); load(o->fb, b); load(o->fpsi, psi); load(o->fvw_, vw);
    load(o->aX, aX); load(o->bX, bX); load(o->psiX, psiX); load(o->vwX, vwX); return 0;
  });
}

int orc_init(void *h) {
  return dispatch<float, float>(h, [&](auto *o) { o->init(); return 0; });
}

long orc_step(void *h, int nsteps, double *out, double *seis) {
  return dispatch<float, float>(h, [&](auto *o) -> long {
    const size_t slab = (size_t)(o->nxt - 2 * o->nabc) * (size_t)(o->nzt - o->nfs - o->nabc);
    long before = o->newton_capped;
    for (int s = 0; s < nsteps; ++s)
      o->step(out ? out + (size_t)s * ORC_NOUT * slab : nullptr,
              seis ? seis + (size_t)s * 3 * o->nsta : nullptr);
    return o->newton_capped - before;
  });
}

int orc_final(void *h, double *out7) {
  return dispatch<float, float>(h, [&](auto *o) {
    size_t n2 = o->n2;
    for (size_t n = 0; n < n2; ++n) {
      out7[n] = double(o->ruptime[n]);
      out7[n2 + n] = double(o->rise[n] - o->ruptime[n]);
      out7[2 * n2 + n] = double(o->sliptime[n]);
      out7[3 * n2 + n] = double(o->slipX[n]);
      out7[4 * n2 + n] = double(o->slipZ[n]);
      out7[5 * n2 + n] = double(o->schangeX[n] - o->T0X[n]);
      out7[6 * n2 + n] = double(o->schangeZ[n] - o->T0Z[n]);
    }
    return 0;
  });
}

int orc_get_field(void *h, int which, double *out) {
  return dispatch<float, float>(h, [&](auto *o) {
    auto *f = o->field(which);
    if (!f) return -1;
    store(*f, out);
    return 0;
  });
}

int orc_set_field(void *h, int which, const double *in) {
  return dispatch<float, float>(h, [&](auto *o) {
    auto *f = o->field(which);
    if (!f) return -1;
    load(*f, in);
    return 0;
  });
}

int orc_get_plane(void *h, int which, double *out) {
  return dispatch<float, float>(h, [&](auto *o) {
    switch (which) {
      case ORC_P_TX: store(o->tx, out); break;
      case ORC_P_TZ: store(o->tz, out); break;
      case ORC_P_RFX: store(o->RFx, out); break;
      case ORC_P_RFZ: store(o->RFz, out); break;
      case ORC_P_AVDX: store(o->avdx, out); break;
      case ORC_P_AVDZ: store(o->avdz, out); break;
      case ORC_P_TABSX: store(o->tabsX, out); break;
      case ORC_P_TABSZ: store(o->tabsZ, out); break;
      case ORC_P_UZ: store(o->uZ, out); break;
      case ORC_P_WX: store(o->wX, out); break;
      case ORC_P_DISTX: store(o->distX, out); break;
      case ORC_P_DISTZ: store(o->distZ, out); break;
      case ORC_P_PSIX: store(o->psiX, out); break;
      case ORC_P_PSIZ: store(o->psiZ, out); break;
      case ORC_P_T0X: store(o->T0X, out); break;
      case ORC_P_T0Z: store(o->T0Z, out); break;
      case ORC_P_V1T: store(o->v1t, out); break;
      case ORC_P_RUPTIME: store(o->ruptime, out); break;
      case ORC_P_RISE: store(o->rise, out); break;
      case ORC_P_SLIPTIME: store(o->sliptime, out); break;
      case ORC_P_SLIPX: store(o->slipX, out); break;
      case ORC_P_SLIPZ: store(o->slipZ, out); break;
      case ORC_P_SCHANGEX: store(o->schangeX, out); break;
      case ORC_P_SCHANGEZ: store(o->schangeZ, out); break;
      default: return -1;
    }
    return 0;
  });
}

}  // extern "C"
