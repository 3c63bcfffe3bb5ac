// Device bodies of the stages of one FD3D_TSN time step.
//
// Each stage is a __device__ function of explicit block and lane
// coordinates, so that the same code serves the one-kernel-per-stage
// launches of a single problem (kernels.hip) and the multi-domain batch
// kernels (batch.hip), where a workgroup takes (domain, stage, block)
// items from a table.  2D stages are templated on the number of lane rows
// BY of a 64 x BY block; results do not depend on BY.
//
// Wavefield buffers.  u, v, w and the six stresses exist twice; a step
// reads the state of the previous step from one set (fo) and writes the
// new state into the other (fn), so no stage overwrites a value another
// stage of the same step still has to read, whatever the order of blocks.
// Which set is which follows from the parity of the step counter
// Ctl::it, which only advance_body changes, once the step's wavefield is
// complete: the state after step n is in set n & 1.  The fault-plane
// stages after it (traction .. nucleation) read the step just counted
// (field_pair_done, time it * dt).  lam, mu and d are shared.
//
// Arithmetic contract: every update evaluates the reference's expression
// with the same operands in the same left-to-right association, built with
// -ffp-contract=off (no fused multiply-add), IEEE-correct float division
// and square root and denormals preserved.  The wavefield therefore
// matches the reference's FP32 CPU build bit for bit on the slip-weakening
// path.  The rate-and-state Newton solve and state update run in FP64.
#pragma once
#include "fd3d_dev.h"

namespace fd3d {

constexpr int BX = 64;   // one wavefront along x

__device__ inline double sqd(double x) { return x * x; }

// ---------------------------------------------------------------------------
// Interior velocities: dvel -> uxx1, vyy1, wzz1 (fd3d_deriv.f90:24-193),
// i in [nabc+1, nxt-nabc], j in [nabc+1, nyt-2], k in [nabc+1, nzt-nfs].
// Fused: each thread updates u, v and w of its cell (all three read only
// stresses, so the reference's three sweeps are independent).
// Bytes per cell: reads 6 stresses + rho + u, v, w; writes u, v, w.
// ---------------------------------------------------------------------------
constexpr int BYF = 4;  // rows per block of the one-thread-per-cell interior sweeps

__device__ inline void vel_interior_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, int gbx, int gby, int gbz, int lx, int ly) {
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int j = g.nabc + 1 + gby * BYF + ly;
  const int k = g.nabc + 1 + gbz;
  if (i > g.nxt - g.nabc || j > g.nyt - 2) return;
  const int q = I3(g, i, j, k), sy = g.sy, sz = g.sz;
  const float c1 = c.c1, c2 = c.c2;
  const float b = c.dth / fo.d[k];
  const float *xx = fo.xx, *yy = fo.yy, *zz = fo.zz, *xy = fo.xy, *yz = fo.yz, *xz = fo.xz;
  fn.u[q] = fo.u[q] + b * (c1 * (xx[q] - xx[q - 1]) + c2 * (xx[q + 1] - xx[q - 2]) +
                         c1 * (xy[q] - xy[q - sy]) + c2 * (xy[q + sy] - xy[q - 2 * sy]) +
                         c1 * (xz[q] - xz[q - sz]) + c2 * (xz[q + sz] - xz[q - 2 * sz]));
  fn.v[q] = fo.v[q] + b * (c1 * (xy[q + 1] - xy[q]) + c2 * (xy[q + 2] - xy[q - 1]) +
                         c1 * (yy[q + sy] - yy[q]) + c2 * (yy[q + 2 * sy] - yy[q - sy]) +
                         c1 * (yz[q] - yz[q - sz]) + c2 * (yz[q + sz] - yz[q - 2 * sz]));
  fn.w[q] = fo.w[q] + b * (c1 * (xz[q + 1] - xz[q]) + c2 * (xz[q + 2] - xz[q - 1]) +
                         c1 * (yz[q] - yz[q - sy]) + c2 * (yz[q + sy] - yz[q - 2 * sy]) +
                         c1 * (zz[q + sz] - zz[q]) + c2 * (zz[q + 2 * sz] - zz[q - sz]));
}

// Newton solve of the rate-and-state split-node velocity (fd3d_deriv.f90:
// 1148-1161): F(uw) = vtilde + cdelta*uw + v0*exp(-psi/a)*sinh(uw).
// The reference loops while |F/F'| > 1e-10 with no bound; this stops after
// newton_max iterations and counts the node.
__device__ inline double newton(double vtilde, double cdelta, double uw, double v0, double ex,
                                int nmax, Ctl *ctl) {
  double ferr = 10.0;
  int n = 0;
  while (fabs(ferr) > 1e-10) {
    if (n >= nmax) {
      atomicAdd(&ctl->newton_capped, 1);
      break;
    }
    ++n;
    double fw = (vtilde + cdelta * uw) + (sinh(uw) * v0) * ex;
    double dfw = cdelta + (cosh(uw) * v0) * ex;
    ferr = fw / dfw;
    uw = uw - ferr;
  }
  return uw;
}

// ---------------------------------------------------------------------------
// Near-fault and fault-plane velocities over the (i,k) plane:
//  * fault columns: bnd2d at j = nyt-1 (2nd order + damping state,
//    fd3d_deriv.f90:284-508), mirror v(nyt) = v(nyt-1), then tasu1/tasw1
//    at j = nyt (:1109-1234);
//  * other columns: the v mirror only, reading v(nyt-1) before pml_vel.
// ---------------------------------------------------------------------------
// FVW = false compiles the slip-weakening path only (for callers that
// know no problem uses rate-and-state friction); results are unchanged.
template <int BY, bool FVW = true>
__device__ inline void vel_fault_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, const Fault &fl, const FvwScalars &fv, Ctl *ctl, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int k = 1 + gby * BY + ly;
  if (i > g.nxt || k > g.nzt) return;
  const int sy = g.sy, sz = g.sz;
  const int qm = I3(g, i, g.nyt - 1, k);
  const bool fault = i >= g.nabc + 1 && i <= g.nxt - g.nabc && k >= g.nabc + 1 && k <= g.nzt - g.nfs;
  if (!fault) {
    fn.v[qm + sy] = fo.v[qm];
    return;
  }
  const int p = I2(g, i, k);
  // bnd2d at j = nyt-1
  {
    const int q = qm;
    const float *xx = fo.xx, *yy = fo.yy, *zz = fo.zz, *xy = fo.xy, *yz = fo.yz, *xz = fo.xz;
    const float b = c.dth / fo.d[k];
    float pdx = ((xx[q] - xx[q - 1]) + (xy[q] - xy[q - sy])) + (xz[q] - xz[q - sz]);
    float a = pdx - fl.au1[p];
    fn.u[q] = fo.u[q] + b * (pdx + c.damp_s * a);
    fl.au1[p] = pdx;
    float pdy = ((xy[q + 1] - xy[q]) + (yy[q + sy] - yy[q])) + (yz[q] - yz[q - sz]);
    a = pdy - fl.av1[p];
    const float vnew = fo.v[q] + b * (pdy + c.damp_s * a);
    fn.v[q] = vnew;
    fl.av1[p] = pdy;
    float pdz = ((xz[q + 1] - xz[q]) + (yz[q] - yz[q - sy])) + (zz[q + sz] - zz[q]);
    a = pdz - fl.aw1[p];
    fn.w[q] = fo.w[q] + b * (pdz + c.damp_s * a);
    fl.aw1[p] = pdz;
    fn.v[q + sy] = vnew;
  }
  // tasu1 / tasw1 at j = nyt
  const int q = qm + sy;
  if (!FVW || !fv.fvw) {
    const float b = c.dth / fo.d[k];
    fn.u[q] = fo.u[q] + (b * 2.0f) * (fl.tx[p] + fl.RFx[p]);
    fn.w[q] = fo.w[q] + b * (2.0f * (fl.tz[p] + fl.RFz[p]));
    return;
  }
  const double dthF = (double)c.dt / (double)c.dh;
  const double dd = (double)fo.d[k];
  const double two_dth_d = 2.0 * (dthF / dd);
  // tasu1: X-node (fd3d_deriv.f90:1141-1162)
  {
    const int pim = p - 1, pkm = p - g.nxt, pikm = p - 1 - g.nxt;
    const double uc = (double)fo.u[q], wx = (double)fl.wX[p];
    const double aX = fl.aX[p], psiX = fl.psiX[p];
    const double sr = sqrt(sqd(2.0 * (fabs(wx) + fabs(fv.wini))) + sqd(2.0 * (fabs(uc) + fabs(fv.uini))));
    const double cdelta = (two_dth_d * fv.Sn) * aX;
    double uw = asinh((exp(psiX / aX) * sr) / (2.0 * fv.v0));
    const double rfz4 = ((((double)fl.RFz[p] + (double)fl.RFz[pim]) + (double)fl.RFz[pkm]) +
                         (double)fl.RFz[pikm]) / 4.0;
    const double vtilde = -sqrt(sqd(wx + two_dth_d * (rfz4 - (double)fl.T0Z[p])) +
                                sqd(uc + two_dth_d * ((double)fl.RFx[p] - (double)fl.T0X[p])));
    const double ex = exp(-psiX / aX);
    uw = newton(vtilde, cdelta, uw, fv.v0, ex, fv.newton_max, ctl);
    // ((tx+T0X)*(-sinh))*v0*ex/tabsX + uini, association as the reference
    fn.u[q] = (float)((((((double)fl.tx[p] + (double)fl.T0X[p]) * (-sinh(uw))) * fv.v0) * ex) /
                         (double)fl.tabsX[p] + fv.uini);
  }
  // tasw1: Z-node (fd3d_deriv.f90:1204-1226)
  {
    const int pip = p + 1, pkp = p + g.nxt, pikp = p + 1 + g.nxt;
    const double wc = (double)fo.w[q], uz = (double)fl.uZ[p];
    const double aZ = fl.aZ[p], psiZ = fl.psiZ[p];
    const double sr = sqrt(sqd(2.0 * (wc - fv.wini)) + sqd(2.0 * (uz - fv.uini)));
    const double cdelta = (two_dth_d * fv.Sn) * aZ;
    double uw = asinh((exp(psiZ / aZ) * sr) / (2.0 * fv.v0));
    const double rfx4 = ((((double)fl.RFx[p] + (double)fl.RFx[pip]) + (double)fl.RFx[pkp]) +
                         (double)fl.RFx[pikp]) / 4.0;
    const double t0x4 = ((((double)fl.T0X[p] + (double)fl.T0X[pip]) + (double)fl.T0X[pkp]) +
                         (double)fl.T0X[pikp]) / 4.0;
    const double vtilde = -sqrt(sqd(wc + two_dth_d * ((double)fl.RFz[p] - (double)fl.T0Z[p])) +
                                sqd(uz + two_dth_d * (rfx4 - t0x4)));
    const double ex = exp(-psiZ / aZ);
    uw = newton(vtilde, cdelta, uw, fv.v0, ex, fv.newton_max, ctl);
    fn.w[q] = (float)((((double)fl.tz[p] + (double)fl.T0Z[p]) * (-((sinh(uw) * fv.v0) * ex))) /
                         (double)fl.tabsZ[p] + fv.wini);
  }
}

// Region lookup of the fused PML launches.
__device__ inline int pml_region(const Pml &pm, int blk) {
  int r = 0;
#pragma unroll
  for (int n = 1; n < kMaxRegions; ++n)
    if (n < pm.nreg && blk >= pm.r[n].blk0) r = n;
  return r;
}

// Cell of a fused PML launch: region, (i2, j2, k2) in the region's box and
// the 3D index.  The divisions by the box extents use host-formed
// multipliers (PmlRegion::mdi, mdj).
struct PmlCell {
  int r, local, i2, j2, k2, i, j, k, q;
};
__device__ inline bool pml_cell(const Dims &g, const Pml &pm, int gbx, int lx, PmlCell &c) {
  c.r = pml_region(pm, gbx);
  const PmlRegion &R = pm.r[c.r];
  c.local = (gbx - R.blk0) * kPmlBlock + lx;
  if (c.local >= R.ncell) return false;
  const int t = udiv(c.local, R.mdi, R.sdi);
  const int km = udiv(t, R.mdj, R.sdj);
  c.i2 = c.local - t * R.ni + 1;
  c.j2 = t - km * R.nj + 1;
  c.k2 = km + 1;
  c.i = R.ib + c.i2 - 1;
  c.j = R.jb + c.j2 - 1;
  c.k = R.kb + c.k2 - 1;
  c.q = I3(g, c.i, c.j, c.k);
  return true;
}

// ---------------------------------------------------------------------------
// Split-field PML velocities, all regions in one launch (pml_uvw,
// fd3d_deriv.f90:1927-2299): p <- [p(1-w) + ((dt/rho)*dsigma)/dh]/(1+w),
// field = sum of splits.  u and w to j = je_uw, v to je_v.  All accesses
// are 32-bit offsets from scalar bases (the region's split block, the
// wavefield arrays, the profile pool).
// ---------------------------------------------------------------------------
__device__ inline void pml_vel_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, const Pml &pm, int gbx, int lx) {
  PmlCell C;
  if (!pml_cell(g, pm, gbx, lx, C)) return;
  const PmlRegion &R = pm.r[C.r];
  const int q = C.q, sy = g.sy, sz = g.sz, i2 = C.i2, j2 = C.j2, k2 = C.k2;
  const float *pr = pm.prof;
  const int nI = R.ni + 1, nJ = R.nj + 1, nK = R.nk + 1;
  const int amx = R.prof_off, apx = amx + nI, amxS = amx + 2 * nI, apxS = amx + 3 * nI;
  const int amy = amx + 4 * nI, apy = amy + nJ;
  const int amz = amx + 4 * nI + 4 * nJ, apz = amz + nK;
  float *sp = R.sp;
  const int nc = R.ncell, l = C.local;
  const float dt = c.dt, dh = c.dh;
  const bool hasv = C.j <= R.je_v;
  // every load before the first store (the stores could alias the inputs
  // as far as the compiler knows, which would serialise the loads)
  const float d0 = fo.d[C.k];
  const float su1 = ld32(sp, l), su2 = ld32(sp, l + nc), su3 = ld32(sp, l + 2 * nc);
  const float sw1 = ld32(sp, l + 6 * nc), sw2 = ld32(sp, l + 7 * nc), sw3 = ld32(sp, l + 8 * nc);
  const float xx0 = ld32(fo.xx, q), xxm = ld32(fo.xx, q - 1), xy0 = ld32(fo.xy, q), xym = ld32(fo.xy, q - sy);
  const float xz0 = ld32(fo.xz, q), xzmz = ld32(fo.xz, q - sz), xzp = ld32(fo.xz, q + 1);
  const float yz0 = ld32(fo.yz, q), yzmy = ld32(fo.yz, q - sy), zz0 = ld32(fo.zz, q), zzp = ld32(fo.zz, q + sz);
  const float pamx = ld32(pr, amx + i2), papx = ld32(pr, apx + i2), pamxS = ld32(pr, amxS + i2),
              papxS = ld32(pr, apxS + i2);
  const float pamy = ld32(pr, amy + j2), papy = ld32(pr, apy + j2), pamz = ld32(pr, amz + k2),
              papz = ld32(pr, apz + k2);
  float sv1 = 0, sv2 = 0, sv3 = 0, xyp = 0, yy0 = 0, yyp = 0, yzmz = 0;
  if (hasv) {
    sv1 = ld32(sp, l + 3 * nc); sv2 = ld32(sp, l + 4 * nc); sv3 = ld32(sp, l + 5 * nc);
    xyp = ld32(fo.xy, q + 1); yy0 = ld32(fo.yy, q); yyp = ld32(fo.yy, q + sy); yzmz = ld32(fo.yz, q - sz);
  }
  const float dtd = dt / d0;
  float pt, s1, s2, s3;
  // u (x, y, z profiles)
  pt = su1 * pamx + (dtd * (xx0 - xxm)) / dh;
  s1 = pt / papx;
  pt = su2 * pamy + (dtd * (xy0 - xym)) / dh;
  s2 = pt / papy;
  pt = su3 * pamz + (dtd * (xz0 - xzmz)) / dh;
  s3 = pt / papz;
  st32(sp, l, s1); st32(sp, l + nc, s2); st32(sp, l + 2 * nc, s3);
  st32_nt(fn.u, q, (s1 + s2) + s3);
  // w (xS, y, z profiles)
  pt = sw1 * pamxS + (dtd * (xzp - xz0)) / dh;
  s1 = pt / papxS;
  pt = sw2 * pamy + (dtd * (yz0 - yzmy)) / dh;
  s2 = pt / papy;
  pt = sw3 * pamz + (dtd * (zzp - zz0)) / dh;
  s3 = pt / papz;
  st32(sp, l + 6 * nc, s1); st32(sp, l + 7 * nc, s2); st32(sp, l + 8 * nc, s3);
  st32_nt(fn.w, q, (s1 + s2) + s3);
  if (hasv) {
    pt = sv1 * pamxS + (dtd * (xyp - xy0)) / dh;
    s1 = pt / papxS;
    pt = sv2 * pamy + (dtd * (yyp - yy0)) / dh;
    s2 = pt / papy;
    pt = sv3 * pamz + (dtd * (yz0 - yzmz)) / dh;
    s3 = pt / papz;
    st32(sp, l + 3 * nc, s1); st32(sp, l + 4 * nc, s2); st32(sp, l + 5 * nc, s3);
    st32_nt(fn.v, q, (s1 + s2) + s3);
  }
}

// ---------------------------------------------------------------------------
// Free surface ghost velocities (fuvw, fd3d_deriv.f90:1018-1062), in two
// launches because w at the ghost row reads the new ghost u and v.
// ---------------------------------------------------------------------------
template <int BY>
__device__ inline void fuvw_uv_body(const Dims &g, const Fields &fo, const Fields &fn, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int j = 1 + gby * BY + ly;
  if (i < 2 || i > g.nxt - 1 || j > g.nyt) return;
  const int ks = g.nzt - g.nfs;
  const int q = I3(g, i, j, ks), sz = g.sz, sy = g.sy;
  fn.u[q + sz] = fn.u[q] - (fn.w[q] - fn.w[q - 1]);
  if (j >= 2 && j <= g.nyt - 1) fn.v[q + sz] = fn.v[q] - (fn.w[q + sy] - fn.w[q]);
}

template <int BY>
__device__ inline void fuvw_w_body(const Dims &g, const Fields &fo, const Fields &fn, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int j = 1 + gby * BY + ly;
  if (i < 2 || i > g.nxt - 2 || j < 2 || j > g.nyt) return;
  const int ks = g.nzt - g.nfs;
  const int q = I3(g, i, j, ks), sz = g.sz, sy = g.sy;
  const float xl = fn.lam[ks + 1], xm = fn.mu[ks + 1];
  const float a = 2.0f * xl, b = xl + 2.0f * xm;
  const float *u = fn.u, *v = fn.v;
  const float s = ((((((u[q + 1] - u[q]) + u[q + 1 + sz]) - u[q + sz]) + v[q]) - v[q - sy]) +
                   v[q + sz]) - v[q - sy + sz];
  fn.w[q + sz] = fn.w[q - sz] - (a / b) * s;
}

// ---------------------------------------------------------------------------
// Interior stresses: dstres (fd3d_deriv.f90:513-622), same range as the
// interior velocities.  Bytes per cell: reads u, v, w, lam, mu + 6 stresses;
// writes 6 stresses.
// ---------------------------------------------------------------------------
__device__ inline void str_interior_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, int gbx, int gby, int gbz, int lx, int ly) {
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int j = g.nabc + 1 + gby * BYF + ly;
  const int k = g.nabc + 1 + gbz;
  if (i > g.nxt - g.nabc || j > g.nyt - 2) return;
  const int q = I3(g, i, j, k), sy = g.sy, sz = g.sz;
  const float c1 = c.c1, c2 = c.c2, dth = c.dth;
  const float *u = fn.u, *v = fn.v, *w = fn.w, *mu = fo.mu;
  const float xl = fo.lam[k], xm = mu[k];
  const float a = xl + 2.0f * xm, b = xl;
  const float dux1 = c1 * (u[q + 1] - u[q]), dux2 = c2 * (u[q + 2] - u[q - 1]);
  const float dvy1 = c1 * (v[q] - v[q - sy]), dvy2 = c2 * (v[q + sy] - v[q - 2 * sy]);
  const float dwz1 = c1 * (w[q] - w[q - sz]), dwz2 = c2 * (w[q + sz] - w[q - 2 * sz]);
  fn.xx[q] = (fo.xx[q] + (dth * a) * (dux1 + dux2)) + (dth * b) * (((dvy1 + dvy2) + dwz1) + dwz2);
  fn.yy[q] = (fo.yy[q] + (dth * a) * (dvy1 + dvy2)) + (dth * b) * (((dux1 + dux2) + dwz1) + dwz2);
  fn.zz[q] = (fo.zz[q] + (dth * a) * (dwz1 + dwz2)) + (dth * b) * (((dux1 + dux2) + dvy1) + dvy2);
  float xmu = 0.5f * (xm + mu[k]);
  fn.xy[q] = fo.xy[q] + (dth * xmu) * (((c1 * (u[q + sy] - u[q]) + c2 * (u[q + 2 * sy] - u[q - sy])) +
                                      c1 * (v[q] - v[q - 1])) + c2 * (v[q + 1] - v[q - 2]));
  xmu = 0.5f * (xm + mu[k + 1]);
  fn.xz[q] = fo.xz[q] + (dth * xmu) * (((c1 * (u[q + sz] - u[q]) + c2 * (u[q + 2 * sz] - u[q - sz])) +
                                      c1 * (w[q] - w[q - 1])) + c2 * (w[q + 1] - w[q - 2]));
  xmu = 0.5f * (xm + mu[k + 1]);
  fn.yz[q] = fo.yz[q] + (dth * xmu) * (((c1 * (v[q + sz] - v[q]) + c2 * (v[q + 2 * sz] - v[q - sz])) +
                                      c1 * (w[q + sy] - w[q])) + c2 * (w[q + 2 * sy] - w[q - sy]));
}

// ---------------------------------------------------------------------------
// Near-fault and fault-plane stresses over the (i,k) plane:
//  * fault columns: strbnd at j = nyt-1 (sxx, sxy0, sxz0, syz0,
//    fd3d_deriv.f90:818-1012), mirror xy, yz to nyt, then sxz1 (tasxz,
//    :677-722) and tasii (:1263-1314) at j = nyt;
//  * other columns: the xy/yz mirror only, before pml_str.
// ---------------------------------------------------------------------------
template <int BY>
__device__ inline void str_fault_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, const Fault &fl, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int k = 1 + gby * BY + ly;
  if (i > g.nxt || k > g.nzt) return;
  const int sy = g.sy, sz = g.sz;
  const int q = I3(g, i, g.nyt - 1, k);
  const bool fault = i >= g.nabc + 1 && i <= g.nxt - g.nabc && k >= g.nabc + 1 && k <= g.nzt - g.nfs;
  if (!fault) {
    fn.xy[q + sy] = fo.xy[q];
    fn.yz[q + sy] = fo.yz[q];
    return;
  }
  const float *u = fn.u, *v = fn.v, *w = fn.w, *mu = fo.mu;
  {
    const float dth = c.dth;
    const float xl = fo.lam[k], xm = mu[k];
    const float a = xl + 2.0f * xm, b = xl;
    fn.xx[q] = (fo.xx[q] + (dth * a) * (u[q + 1] - u[q])) + (dth * b) * (((v[q] - v[q - sy]) + w[q]) - w[q - sz]);
    fn.yy[q] = (fo.yy[q] + (dth * a) * (v[q] - v[q - sy])) + (dth * b) * (((u[q + 1] - u[q]) + w[q]) - w[q - sz]);
    fn.zz[q] = (fo.zz[q] + (dth * a) * (w[q] - w[q - sz])) + (dth * b) * (((u[q + 1] - u[q]) + v[q]) - v[q - sy]);
    const float dthh = c.dthh;
    float xmu = xm + mu[k];
    const float xyn = fo.xy[q] + (dthh * xmu) * (((u[q + sy] - u[q]) + v[q]) - v[q - 1]);
    fn.xy[q] = xyn;
    xmu = xm + mu[k + 1];
    fn.xz[q] = fo.xz[q] + (dthh * xmu) * (((u[q + sz] - u[q]) + w[q]) - w[q - 1]);
    xmu = xm + mu[k + 1];
    const float yzn = fo.yz[q] + (dthh * xmu) * (((v[q + sz] - v[q]) + w[q + sy]) - w[q]);
    fn.yz[q] = yzn;
    fn.xy[q + sy] = xyn;
    fn.yz[q + sy] = yzn;
  }
  {
    const int qf = q + sy;
    const float c1 = c.c1, c2 = c.c2;
    float xmu = mu[k] + mu[k + 1];
    fn.xz[qf] = fo.xz[qf] + (c.dthh * xmu) *
                              (((c1 * (u[qf + sz] - u[qf]) + c2 * (u[qf + 2 * sz] - u[qf - sz])) +
                                c1 * (w[qf] - w[qf - 1])) + c2 * (w[qf + 1] - w[qf - 2]));
    const float dth = c.dth;
    const float xl = fo.lam[k], xm = mu[k];
    const float a = xl + 2.0f * xm, b = xl;
    const float diff1 = c1 * (u[qf + 1] - u[qf]) + c2 * (u[qf + 2] - u[qf - 1]);
    const float diff3 = c1 * (w[qf] - w[qf - sz]) + c2 * (w[qf + sz] - w[qf - 2 * sz]);
    const float vm = v[qf - sy];
    const int p = I2(g, i, k);
    const float v1t = vm - (b * (diff1 + diff3)) / (2.0f * a);
    fl.v1t[p] = v1t;
    const float dv = v1t - vm;
    fn.xx[qf] = (fo.xx[qf] + (dth * a) * diff1) + (dth * b) * (2.0f * dv + diff3);
    fn.yy[qf] = (fo.yy[qf] + ((dth * a) * 2.0f) * dv) + (dth * b) * (diff1 + diff3);
    fn.zz[qf] = (fo.zz[qf] + (dth * a) * diff3) + (dth * b) * (diff1 + 2.0f * dv);
  }
}

// ---------------------------------------------------------------------------
// Split-field PML stresses, all regions in one launch (pml_xyz,
// fd3d_deriv.f90:2304-3145), with the top-layer xz/yz split zeroing of
// regions 1 (always) and 2, 3 (free surface only) fused after the update.
// ---------------------------------------------------------------------------
__device__ inline void pml_str_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, const Pml &pm, int gbx, int lx) {
  PmlCell C;
  if (!pml_cell(g, pm, gbx, lx, C)) return;
  const PmlRegion &R = pm.r[C.r];
  const int q = C.q, sy = g.sy, sz = g.sz, i2 = C.i2, j2 = C.j2, k2 = C.k2, k = C.k;
  const float *pr = pm.prof;
  const int nI = R.ni + 1, nJ = R.nj + 1, nK = R.nk + 1;
  const int amx = R.prof_off, apx = amx + nI, amxS = amx + 2 * nI, apxS = amx + 3 * nI;
  const int amy = amx + 4 * nI, apy = amy + nJ, amyS = amy + 2 * nJ, apyS = amy + 3 * nJ;
  const int amz = amx + 4 * nI + 4 * nJ, apz = amz + nK, amzS = amz + 2 * nK, apzS = amz + 3 * nK;
  float *sp = R.sp;
  const int nc = R.ncell, l = C.local;
  const float dt = c.dt, dh = c.dh;
  (void)fo;
  const float *u = fn.u, *v = fn.v, *w = fn.w, *mu = fn.mu;
  const bool hxy = C.j <= R.je_v, hxz = C.j <= R.je_xz;
  // every load before the first store (see pml_vel_body)
  const float xl = fn.lam[k], xm = mu[k];
  const float u0 = ld32(u, q), up = ld32(u, q + 1), v0 = ld32(v, q), vmy = ld32(v, q - sy), w0 = ld32(w, q),
              wmz = ld32(w, q - sz);
  const float ax = ld32(pr, amxS + i2), bx = ld32(pr, apxS + i2), ay = ld32(pr, amy + j2), by = ld32(pr, apy + j2),
              az = ld32(pr, amz + k2), bz = ld32(pr, apz + k2);
  const float sxx1 = ld32(sp, l + 9 * nc), sxx2 = ld32(sp, l + 10 * nc), sxx3 = ld32(sp, l + 11 * nc);
  const float syy1 = ld32(sp, l + 12 * nc), syy2 = ld32(sp, l + 13 * nc), syy3 = ld32(sp, l + 14 * nc);
  const float szz1 = ld32(sp, l + 15 * nc), szz2 = ld32(sp, l + 16 * nc), szz3 = ld32(sp, l + 17 * nc);
  float mxy = 0, uy = 0, vmx = 0, sxy1 = 0, sxy2 = 0, pamyS = 0, papyS = 0, pamx = 0, papx = 0;
  float mxz = 0, uz = 0, wmx = 0, sxz1 = 0, sxz2 = 0, pamzS = 0, papzS = 0;
  float myz = 0, vz = 0, wy = 0, syz1 = 0, syz2 = 0;
  if (hxy || hxz) { pamx = ld32(pr, amx + i2); papx = ld32(pr, apx + i2); }
  if (hxy) {
    mxy = mu[k]; uy = ld32(u, q + sy); vmx = ld32(v, q - 1); sxy1 = ld32(sp, l + 18 * nc); sxy2 = ld32(sp, l + 19 * nc);
    pamyS = ld32(pr, amyS + j2); papyS = ld32(pr, apyS + j2);
    myz = mu[k + 1]; vz = ld32(v, q + sz); wy = ld32(w, q + sy); syz1 = ld32(sp, l + 22 * nc);
    syz2 = ld32(sp, l + 23 * nc);
  }
  if (hxz) {
    mxz = mu[k + 1]; uz = ld32(u, q + sz); wmx = ld32(w, q - 1); sxz1 = ld32(sp, l + 20 * nc);
    sxz2 = ld32(sp, l + 21 * nc);
  }
  if (hxy || hxz) { pamzS = ld32(pr, amzS + k2); papzS = ld32(pr, apzS + k2); }
  const float a = xl + 2.0f * xm, b = xl;
  const float diff1 = up - u0, diff2 = v0 - vmy, diff3 = w0 - wmz;
  float pt, s1, s2, s3;
  pt = sxx1 * ax + ((dt * a) * diff1) / dh; s1 = pt / bx;
  pt = sxx2 * ay + ((dt * b) * diff2) / dh; s2 = pt / by;
  pt = sxx3 * az + ((dt * b) * diff3) / dh; s3 = pt / bz;
  st32(sp, l + 9 * nc, s1); st32(sp, l + 10 * nc, s2); st32(sp, l + 11 * nc, s3);
  st32_nt(fn.xx, q, (s1 + s2) + s3);
  pt = syy1 * ax + ((dt * b) * diff1) / dh; s1 = pt / bx;
  pt = syy2 * ay + ((dt * a) * diff2) / dh; s2 = pt / by;
  pt = syy3 * az + ((dt * b) * diff3) / dh; s3 = pt / bz;
  st32(sp, l + 12 * nc, s1); st32(sp, l + 13 * nc, s2); st32(sp, l + 14 * nc, s3);
  st32_nt(fn.yy, q, (s1 + s2) + s3);
  pt = szz1 * ax + ((dt * b) * diff1) / dh; s1 = pt / bx;
  pt = szz2 * ay + ((dt * b) * diff2) / dh; s2 = pt / by;
  pt = szz3 * az + ((dt * a) * diff3) / dh; s3 = pt / bz;
  st32(sp, l + 15 * nc, s1); st32(sp, l + 16 * nc, s2); st32(sp, l + 17 * nc, s3);
  st32_nt(fn.zz, q, (s1 + s2) + s3);
  const bool zero_top = R.ztop && k2 == R.nk;
  if (hxy) {
    const float xmu = (xm + mxy) / 2.0f;
    pt = sxy1 * pamyS + ((dt * xmu) * (uy - u0)) / dh; s1 = pt / papyS;
    pt = sxy2 * pamx + ((dt * xmu) * (v0 - vmx)) / dh; s2 = pt / papx;
    st32(sp, l + 18 * nc, s1); st32(sp, l + 19 * nc, s2);
    st32_nt(fn.xy, q, s1 + s2);
  }
  if (hxz) {
    const float xmu = (xm + mxz) / 2.0f;
    pt = sxz1 * pamzS + ((dt * xmu) * (uz - u0)) / dh; s1 = pt / papzS;
    pt = sxz2 * pamx + ((dt * xmu) * (w0 - wmx)) / dh; s2 = pt / papx;
    st32_nt(fn.xz, q, s1 + s2);
    if (zero_top && j2 <= R.nzx) { s1 = 0.0f; s2 = 0.0f; }
    st32(sp, l + 20 * nc, s1); st32(sp, l + 21 * nc, s2);
  }
  if (hxy) {
    const float xmu = (xm + myz) / 2.0f;
    pt = syz1 * pamzS + ((dt * xmu) * (vz - v0)) / dh; s1 = pt / papzS;
    pt = syz2 * pamyS + ((dt * xmu) * (wy - w0)) / dh; s2 = pt / papyS;
    st32_nt(fn.yz, q, s1 + s2);
    if (zero_top && j2 <= R.nzy) { s1 = 0.0f; s2 = 0.0f; }
    st32(sp, l + 22 * nc, s1); st32(sp, l + 23 * nc, s2);
  }
}

// Free-surface stress imaging (fres, fd3d_deriv.f90:1067-1104).
template <int BY>
__device__ inline void fres_body(const Dims &g, const Fields &fo, const Fields &fn, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int j = 1 + gby * BY + ly;
  if (i > g.nxt || j > g.nyt) return;
  const int ks = g.nzt - g.nfs;
  const int q = I3(g, i, j, ks), sz = g.sz;
  fn.zz[q + sz] = -fn.zz[q];
  fn.zz[q + 2 * sz] = -fn.zz[q - sz];
  fn.xz[q + sz] = -fn.xz[q - sz];
  fn.xz[q + 2 * sz] = -fn.xz[q - 2 * sz];
  fn.yz[q + sz] = -fn.yz[q - sz];
  fn.yz[q + 2 * sz] = -fn.yz[q - 2 * sz];
  fn.xz[q] = 0.0f;
  fn.yz[q] = 0.0f;
}

// ---------------------------------------------------------------------------
// Trial traction (fd3d_theo.f90:232-308) on fault nodes.
// ---------------------------------------------------------------------------
template <int BY>
__device__ inline void traction_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, const Fault &fl, int gbx, int gby, int lx, int ly) {
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int k = g.nabc + 1 + gby * BY + ly;
  const int ks = g.nzt - g.nfs;
  if (i > g.nxt - g.nabc || k > ks) return;
  const int q = I3(g, i, g.nyt, k), sy = g.sy, sz = g.sz, p = I2(g, i, k);
  const float c1 = c.c1, c2 = c.c2, dht = c.dht;
  const float *xx = fn.xx, *zz = fn.zz, *xy = fn.xy, *yz = fn.yz, *xz = fn.xz;
  float pdz;
  const bool top_free = !g.fspace && k == ks;
  if (!top_free)
    pdz = ((c1 * (xz[q + 1] - xz[q]) + c2 * (xz[q + 2] - xz[q - 1])) / 2.0f +
           (c1 * (zz[q + sz] - zz[q]) + c2 * (zz[q + 2 * sz] - zz[q - sz])) / 2.0f) - yz[q - sy];
  else
    pdz = ((xz[q + 1 - sz] - xz[q - sz]) / 8.0f + (zz[q + sz] - zz[q]) / 4.0f) - yz[q - sy - sz] / 4.0f;
  float av = c.damp_s * (pdz - fl.avdz[p]);
  float rf = pdz + av;
  fl.RFz[p] = rf;
  const float dd = fn.d[k];
  if (!top_free)
    fl.tz[p] = -rf - ((0.5f * dd) * dht) * fn.w[q];
  else
    fl.tz[p] = -rf - ((((0.5f * dd) * dht) * fn.w[q]) / 2.0f);
  fl.avdz[p] = pdz;
  const float pdx = ((c1 * (xx[q] - xx[q - 1]) + c2 * (xx[q + 1] - xx[q - 2])) / 2.0f +
                     (c1 * (xz[q] - xz[q - sz]) + c2 * (xz[q + sz] - xz[q - 2 * sz])) / 2.0f) - xy[q - sy];
  av = c.damp_s * (pdx - fl.avdx[p]);
  rf = pdx + av;
  fl.RFx[p] = rf;
  fl.tx[p] = -rf - ((0.5f * dd) * dht) * fn.u[q];
  fl.avdx[p] = pdx;
}

// Ghost-row traction of the free surface (fd3d_theo.f90:302-308).
__device__ inline void traction_mirror_body(const Dims &g, const Fault &fl, int gidx) {
  const int i = g.nabc + 1 + gidx;
  if (i > g.nxt - g.nabc) return;
  fl.tx[I2(g, i, g.nzt - 1)] = fl.tx[I2(g, i, g.nzt - 2)];
  fl.tz[I2(g, i, g.nzt - 1)] = -fl.tz[I2(g, i, g.nzt - 3)];
}

// |T| at Z- and X-nodes (fd3d_theo.f90:313-324).
template <int BY>
__device__ inline void tabs_body(const Dims &g, const Fault &fl, int gbx, int gby, int lx, int ly) {
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int k = g.nabc + 1 + gby * BY + ly;
  if (i > g.nxt - g.nabc || k > g.nzt - g.nfs) return;
  const int p = I2(g, i, k), n = g.nxt;
  const float *tx = fl.tx, *tz = fl.tz, *T0X = fl.T0X, *T0Z = fl.T0Z;
  float tint = (((((((tx[p] + T0X[p]) + tx[p + 1]) + T0X[p + 1]) + tx[p + n]) + T0X[p + n]) + tx[p + 1 + n]) +
                T0X[p + 1 + n]) / 4.0f;
  float tzz = tz[p] + T0Z[p];
  fl.tabsZ[p] = sqrtf(tint * tint + tzz * tzz);
  tint = (((((((tz[p] + T0Z[p]) + tz[p - 1]) + T0Z[p - 1]) + tz[p - n]) + T0Z[p - n]) + tz[p - 1 - n]) +
          T0Z[p - 1 - n]) / 4.0f;
  const float txx = tx[p] + T0X[p];
  fl.tabsX[p] = sqrtf(txx * txx + tint * tint);
}

// Rate-and-state state update and friction value (fd3d_theo.f90:336-345,
// 385-390) in FP64; returns the strength, updates psi.
__device__ inline double fvw_strength(double sr, double a, double b, double vw, double dc, double &psi,
                                      const FvwScalars &fv, double dt) {
  const double flv = fv.f0 - (b - a) * log(sr / fv.v0);
  const double qv = sr / vw;
  const double q8 = ((((((qv * qv) * qv) * qv) * qv) * qv) * qv) * qv;
  const double fss = fv.fw + (flv - fv.fw) / pow(1.0 + q8, 1.0 / 8.0);
  const double psiss = a * (log(sinh(fss / a)) + log((2.0 * fv.v0) / sr));
  psi = (psi - psiss) * exp(-((sr * dt) / dc)) + psiss;
  return (fv.Sn * a) * asinh((sr * exp(psi / a)) / (2.0 * fv.v0));
}

// ---------------------------------------------------------------------------
// Friction at Z-nodes then X-nodes of one fault node (fd3d_theo.f90:326-410).
// The Z part changes only tz and Z-node state, the X part only tx and
// X-node state, so both loops of the reference fuse per node.
// ---------------------------------------------------------------------------
template <int BY, bool FVW = true>
__device__ inline void friction_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, const Fault &fl, const FvwScalars &fv, const Ctl *ctl, int gbx, int gby, int lx, int ly) {
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int k = g.nabc + 1 + gby * BY + ly;
  if (i > g.nxt - g.nabc || k > g.nzt - g.nfs) return;
  const int p = I2(g, i, k);
  const int q = I3(g, i, g.nyt, k), sz = g.sz;
  const float dt = c.dt;
  const float time = (float)ctl->it * dt;   // time at the end of this step (counted by advance)
  const float rup_tresh = 1.e-3f;
  const float *u = fn.u, *w = fn.w;
  // ---- Z-node
  {
    const float tab = fl.tabsZ[p];
    const float W = w[q];
    const float uz = (((u[q] + u[q + 1]) + u[q + sz]) + u[q + 1 + sz]) / 4.0f;
    fl.uZ[p] = uz;
    const float u1out = -sqrtf(W * W + uz * uz);
    float distZ = fl.distZ[p];
    if (FVW && fv.fvw) {
      const double sr = sqrt(sqd(2.0 * (fabs((double)W) + fabs(fv.wini))) +
                             sqd(2.0 * (fabs((double)uz) + fabs(fv.uini))));
      double psi = fl.psiZ[p];
      const double fric = fvw_strength(sr, fl.aZ[p], fl.bZ[p], fl.vwZ[p], (double)fl.Dc[p], psi, fv, (double)dt);
      fl.psiZ[p] = psi;
      distZ = distZ - (2.0f * u1out) * dt;
      fl.schangeZ[p] = (float)fric;
      if (fabsf(2.0f * u1out) > rup_tresh) {
        if (fl.ruptime[p] != 1.e4f) fl.rise[p] = time;
        if (fl.ruptime[p] == 1.e4f) fl.ruptime[p] = time;
      }
    } else {
      float fric;
      if (distZ <= fl.DcZ[p])
        fric = (fl.peakZ[p] * (1.0f - distZ / fl.DcZ[p]) + (fl.dynZ[p] * distZ) / fl.DcZ[p]) + fl.coh[p];
      else
        fric = fl.dynZ[p] + fl.coh[p];
      if (tab >= fric) {
        distZ = distZ - (2.0f * u1out) * dt;
        fl.tz[p] = ((fl.tz[p] + fl.T0Z[p]) * fric) / tab - fl.T0Z[p];
        if (-2.0f * u1out > rup_tresh) {
          if (fl.ruptime[p] != 1.e4f) fl.rise[p] = time;
          if (fl.ruptime[p] == 1.e4f) fl.ruptime[p] = time;
        }
      }
    }
    fl.distZ[p] = distZ;
    fl.slipZ[p] = fl.slipZ[p] - (2.0f * W) * dt;
    fl.slipX[p] = fl.slipX[p] - (2.0f * uz) * dt;
    if (fl.sliptime[p] == 1.e4f && distZ > fl.Dc[p]) fl.sliptime[p] = time;
  }
  // ---- X-node
  {
    const float tab = fl.tabsX[p];
    const float U = u[q];
    const float wx = (((w[q] + w[q - 1]) + w[q - sz]) + w[q - 1 - sz]) / 4.0f;
    fl.wX[p] = wx;
    const float u1out = -sqrtf(wx * wx + U * U);
    float distX = fl.distX[p];
    if (FVW && fv.fvw) {
      const double sr = sqrt(sqd(2.0 * (fabs((double)wx) + fabs(fv.wini))) +
                             sqd(2.0 * (fabs((double)U) + fabs(fv.uini))));
      double psi = fl.psiX[p];
      const double fric = fvw_strength(sr, fl.aX[p], fl.bX[p], fl.vwX[p], (double)fl.Dc[p], psi, fv, (double)dt);
      fl.psiX[p] = psi;
      distX = distX - (2.0f * u1out) * dt;
      fl.schangeX[p] = (float)fric;
    } else {
      float fric;
      if (distX <= fl.Dc[p])
        fric = (fl.peakX[p] * (1.0f - distX / fl.DcX[p]) + (fl.dynX[p] * distX) / fl.DcX[p]) + fl.coh[p];
      else
        fric = fl.dynX[p] + fl.coh[p];
      if (tab >= fric) {
        distX = distX - (2.0f * u1out) * dt;
        fl.tx[p] = ((fl.tx[p] + fl.T0X[p]) * fric) / tab - fl.T0X[p];
      }
    }
    fl.distX[p] = distX;
  }
}

// ---------------------------------------------------------------------------
// Per-step outputs into the device ring (fd3d_theo.f90:413-453) and the
// station samples (:456-467).
// ---------------------------------------------------------------------------
template <int BY>
__device__ inline void outputs_body(const Dims &g, const Fields &fo, const Fields &fn, const Fault &fl, const FvwScalars &fv, const Outputs &o, Ctl *ctl, int gbx, int gby, int lx, int ly) {
  const int it = ctl->it;   // this step, 1-based (counted by advance)
  const int slot = (it - 1) % o.ring_steps;
  const int ks = g.nzt - g.nfs;
  const long long slab = (long long)g.nxT * g.nzT;
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int k = g.nabc + 1 + gby * BY + ly;
  if (i > g.nxt - g.nabc || k > ks) return;
  const int p = I2(g, i, k), n = g.nxt;
  const int q = I3(g, i, g.nyt, k), sz = g.sz;
  const float *u = fn.u;
  const float *tx = fl.tx, *T0X = fl.T0X;
  float srx, scx, scz;
  if (k < ks) {
    srx = ((((-(2.0f * u[q])) - 2.0f * u[q + 1]) - 2.0f * u[q + sz]) - 2.0f * u[q + 1 + sz]) / 4.0f;
    scx = (((((((tx[p] + T0X[p]) + tx[p + 1]) + T0X[p + 1]) + tx[p + n]) + T0X[p + n]) + tx[p + 1 + n]) +
           T0X[p + 1 + n]) / 4.0f;
  } else if (!fv.fvw) {
    srx = ((-(2.0f * u[q])) - 2.0f * u[q + 1]) / 2.0f;
    scx = (((tx[p] + T0X[p]) + tx[p + 1]) + T0X[p + 1]) / 2.0f;
  } else {
    srx = ((-(2.0f * u[q])) - 2.0f * u[q + 1]) / 4.0f;
    scx = 0.0f;
  }
  const float srz = -(2.0f * fn.w[q]);
  if (fv.fvw) {
    scx = fl.schangeX[p];
    scz = fl.schangeZ[p];
  } else {
    scz = fl.tz[p] + fl.T0Z[p];
    fl.schangeX[p] = scx;
    fl.schangeZ[p] = scz;
  }
  float *ring = o.ring + (long long)slot * 5 * slab;
  const long long e = (long long)(i - g.nabc - 1) + (long long)g.nxT * (k - g.nabc - 1);
  ring[e] = srx;
  ring[slab + e] = srz;
  ring[2 * slab + e] = scx;
  ring[3 * slab + e] = scz;
  ring[4 * slab + e] = fv.fvw ? (float)fl.psiZ[p] : 0.0f;
  if (!isfinite(srx) || !isfinite(srz)) atomicOr(&ctl->nonfinite, 1);
}

// Source-time bins of this step (fd3d_theo.f90:560-578), after the output
// slabs of the step are in the ring.  Item 0 adds the step's moment rate
// to its bin: each term is the reference's float expression, the sum over
// the fault is a fixed-order tree in double (independent of the block size) (the reference accumulates in
// float node by node, so the two agree to rounding, not bit for bit).
// Items 1.. hold one patch per lane: the patch sum is sequential in the
// reference's order (x fastest), then added to the bin as the reference
// does (float sum / double divisor, rounded back to float).  Nodes outside
// the output slab count as zero, as sliprateoutX/Z are zero there.
// red: 256 doubles of LDS.
template <int NT>
__device__ inline void rates_body(const Dims &g, const Coef &c, const Fields &fn, const Outputs &o, const Rates &R,
                                  const Ctl *ctl, int item, int t, double *red) {
  const int it1 = ctl->it;   // this step, 1-based (counted by advance)
  const int kbin = (int)(((float)(it1 - 1) * c.dt) / R.dtseis) + 1;
  if (kbin < 1 || kbin > R.nbins) return;   // the reference reports CHYBA and writes out of range
  const long long slab = (long long)g.nxT * g.nzT;
  const float *srx = o.ring + (long long)((it1 - 1) % o.ring_steps) * 5 * slab;
  const float *srz = srx + slab;
  const float per = R.dtseis / c.dt;
  if (item == 0) {
    // 256 partial sums (node n goes to partial n % 256) and a fixed tree:
    // the same result whatever the block size NT (>= 256) that runs it
    static_assert(NT >= 256, "rates_body needs at least 256 lanes");
    if (t < 256) {
      double acc = 0.0;
      for (long long n = t; n < slab; n += 256) {
        const int k = g.nabc + 1 + (int)(n / g.nxT);
        const float sx = srx[n], sz = srz[n];
        const float m = fn.mu[k];
        acc += (double)((((sqrtf(sx * sx + sz * sz) * m) * c.dh) * c.dh) / per);
      }
      red[t] = acc;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    if (t == 0) R.mr[kbin - 1] = (float)((double)R.mr[kbin - 1] + red[0]);
    return;
  }
  const int p = (item - 1) * NT + t;
  if (p >= R.NL * R.NW) return;
  const int ip = p % R.NL + 1, jp = p / R.NL + 1;   // 1-based patch indices (do j=1,NW / do i=1,NL)
  const int jto = max(1, (int)((R.dW / c.dh) * (float)jp)) + 1 + g.nabc;
  const int jfrom = min(jto, (int)((R.dW / c.dh) * (float)(jp - 1)) + 1) + 1 + g.nabc;
  const int ifrom = (int)((R.dL / c.dh) * (float)(ip - 1)) + 1 + g.nabc;
  const int ito = (int)((R.dL / c.dh) * (float)ip) + g.nabc;
  float sx = 0.0f, sz = 0.0f;
  for (int kk = jfrom; kk <= jto; ++kk)
    for (int ii = ifrom; ii <= ito; ++ii) {
      float vx = 0.0f, vz = 0.0f;
      if (ii >= g.nabc + 1 && ii <= g.nxt - g.nabc && kk >= g.nabc + 1 && kk <= g.nzt - g.nfs) {
        const long long e = (long long)(ii - g.nabc - 1) + (long long)g.nxT * (kk - g.nabc - 1);
        vx = srx[e]; vz = srz[e];
      }
      sx = sx + vx;
      sz = sz + vz;
    }
  const double div = (double)((float)((ito - ifrom + 1) * (jto - jfrom + 1)) * per);
  const long long q = (long long)(p) * R.nbins + (kbin - 1);
  R.msrx[q] = (float)((double)R.msrx[q] + (double)sx / div);
  R.msrz[q] = (float)((double)R.msrz[q] + (double)sz / div);
}

// The step counter (and with it the buffer parity) advances once every
// wavefield stage of the step has finished; the fault-plane stages follow.
__device__ inline void advance_body(Ctl *ctl) { ctl->it += 1; }

// The old (fo) and new (fn) wavefield sets of the step in progress.
struct FieldPair {
  const Fields &fo, &fn;
};
__device__ inline FieldPair field_pair(const Fields &f0, const Fields &f1, const Ctl *ctl) {
  return (ctl->it & 1) ? FieldPair{f1, f0} : FieldPair{f0, f1};
}
// The same pair for the step advance_body has just counted (the stages
// after it).
__device__ inline FieldPair field_pair_done(const Fields &f0, const Fields &f1, const Ctl *ctl) {
  return ((ctl->it - 1) & 1) ? FieldPair{f1, f0} : FieldPair{f0, f1};
}

// Station samples of one step into the device ring (fd3d_theo.f90:456-467),
// one thread per station.
__device__ inline void stations_body(const Dims &g, const Fields &fo, const Fields &fn, const Outputs &o,
                                     const Ctl *ctl, int s) {
  (void)fo;
  if (s >= o.nsta) return;
  const int slot = (ctl->it - 1) % o.ring_steps;   // the step counted by advance
  const int q = I3(g, o.sta[3 * s], o.sta[3 * s + 1], o.sta[3 * s + 2]);
  float *dst = o.seis + ((long long)slot * o.nsta + s) * 3;
  dst[0] = fn.u[q];
  dst[1] = fn.v[q];
  dst[2] = fn.w[q];
}

// TPV104 smooth nucleation (fd3d_theo.f90:515-542), FP64.
template <int BY>
__device__ inline void nucleation_body(const Dims &g, const Coef &c, const Fault &fl, const FvwScalars &fv, const Ctl *ctl, int gbx, int gby, int lx, int ly) {
  const double time = (double)((float)ctl->it * c.dt);   // the step counted by advance
  const double t = time - (double)c.dt / 2.0;
  if (!(t <= fv.TT2)) return;
  const int i = g.nabc - 1 + gbx * BX + lx;
  const int k = g.nabc - 1 + gby * BY + ly;
  if (i > g.nxt - g.nabc || k > g.nzt - g.nfs) return;
  double GT;
  if (t < fv.TT2)
    GT = exp(sqd(t - fv.TT2) / (t * (t - 2.0 * fv.TT2)));
  else
    GT = 1.0;
  const double dhF = (double)c.dh;
  const double hx = (double)i * dhF, hz = (double)k * dhF;
  const double rr = sqd(hx - fv.hx0) + sqd(hz - fv.hz0);
  double FXZ;
  if (rr < fv.RR2 - dhF)
    FXZ = exp(rr / ((rr - fv.RR2) - dhF));
  else
    FXZ = 0.0;
  fl.T0X[I2(g, i, k)] = (float)(fv.strinixI + (fv.perturb * FXZ) * GT);
}


}  // namespace fd3d
