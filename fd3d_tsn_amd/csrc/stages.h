// Device bodies of the stages of one FD3D_TSN time step.
//
// Each stage is a __device__ function of explicit block and lane
// coordinates, so that the same code serves the one-kernel-per-stage
// launches of a single problem (kernels.hip) and the multi-domain batch
// kernels (batch.hip), where a workgroup takes (domain, stage, block)
// items from a table.  2D stages are templated on the number of lane rows
// BY of a 64 x BY block; results do not depend on BY.
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

__device__ inline void vel_interior_body(const Dims &g, const Coef &c, const Fields &f, int gbx, int gby, int gbz, int lx, int ly) {
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int j = g.nabc + 1 + gby * BYF + ly;
  const int k = g.nabc + 1 + gbz;
  if (i > g.nxt - g.nabc || j > g.nyt - 2) return;
  const int q = I3(g, i, j, k), sy = g.sy, sz = g.sz;
  const float c1 = c.c1, c2 = c.c2;
  const float b = c.dth / f.d[q];
  const float *xx = f.xx, *yy = f.yy, *zz = f.zz, *xy = f.xy, *yz = f.yz, *xz = f.xz;
  f.u[q] = f.u[q] + b * (c1 * (xx[q] - xx[q - 1]) + c2 * (xx[q + 1] - xx[q - 2]) +
                         c1 * (xy[q] - xy[q - sy]) + c2 * (xy[q + sy] - xy[q - 2 * sy]) +
                         c1 * (xz[q] - xz[q - sz]) + c2 * (xz[q + sz] - xz[q - 2 * sz]));
  f.v[q] = f.v[q] + b * (c1 * (xy[q + 1] - xy[q]) + c2 * (xy[q + 2] - xy[q - 1]) +
                         c1 * (yy[q + sy] - yy[q]) + c2 * (yy[q + 2 * sy] - yy[q - sy]) +
                         c1 * (yz[q] - yz[q - sz]) + c2 * (yz[q + sz] - yz[q - 2 * sz]));
  f.w[q] = f.w[q] + b * (c1 * (xz[q + 1] - xz[q]) + c2 * (xz[q + 2] - xz[q - 1]) +
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
// Also advances the step counter read by the later kernels of this step.
// ---------------------------------------------------------------------------
// FVW = false compiles the slip-weakening path only (for callers that
// know no problem uses rate-and-state friction); results are unchanged.
template <int BY, bool FVW = true>
__device__ inline void vel_fault_body(const Dims &g, const Coef &c, const Fields &f, const Fault &fl, const FvwScalars &fv, Ctl *ctl, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int k = 1 + gby * BY + ly;
  if (gbx == 0 && gby == 0 && lx == 0 && ly == 0) ctl->it += 1;
  if (i > g.nxt || k > g.nzt) return;
  const int sy = g.sy, sz = g.sz;
  const int qm = I3(g, i, g.nyt - 1, k);
  const bool fault = i >= g.nabc + 1 && i <= g.nxt - g.nabc && k >= g.nabc + 1 && k <= g.nzt - g.nfs;
  if (!fault) {
    f.v[qm + sy] = f.v[qm];
    return;
  }
  const int p = I2(g, i, k);
  // bnd2d at j = nyt-1
  {
    const int q = qm;
    const float *xx = f.xx, *yy = f.yy, *zz = f.zz, *xy = f.xy, *yz = f.yz, *xz = f.xz;
    const float b = c.dth / f.d[q];
    float pdx = ((xx[q] - xx[q - 1]) + (xy[q] - xy[q - sy])) + (xz[q] - xz[q - sz]);
    float a = pdx - fl.au1[p];
    f.u[q] = f.u[q] + b * (pdx + c.damp_s * a);
    fl.au1[p] = pdx;
    float pdy = ((xy[q + 1] - xy[q]) + (yy[q + sy] - yy[q])) + (yz[q] - yz[q - sz]);
    a = pdy - fl.av1[p];
    const float vnew = f.v[q] + b * (pdy + c.damp_s * a);
    f.v[q] = vnew;
    fl.av1[p] = pdy;
    float pdz = ((xz[q + 1] - xz[q]) + (yz[q] - yz[q - sy])) + (zz[q + sz] - zz[q]);
    a = pdz - fl.aw1[p];
    f.w[q] = f.w[q] + b * (pdz + c.damp_s * a);
    fl.aw1[p] = pdz;
    f.v[q + sy] = vnew;
  }
  // tasu1 / tasw1 at j = nyt
  const int q = qm + sy;
  if (!FVW || !fv.fvw) {
    const float b = c.dth / f.d[q];
    f.u[q] = f.u[q] + (b * 2.0f) * (fl.tx[p] + fl.RFx[p]);
    f.w[q] = f.w[q] + b * (2.0f * (fl.tz[p] + fl.RFz[p]));
    return;
  }
  const double dthF = (double)c.dt / (double)c.dh;
  const double dd = (double)f.d[q];
  const double two_dth_d = 2.0 * (dthF / dd);
  // tasu1: X-node (fd3d_deriv.f90:1141-1162)
  {
    const int pim = p - 1, pkm = p - g.nxt, pikm = p - 1 - g.nxt;
    const double uc = (double)f.u[q], wx = (double)fl.wX[p];
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
    f.u[q] = (float)((((((double)fl.tx[p] + (double)fl.T0X[p]) * (-sinh(uw))) * fv.v0) * ex) /
                         (double)fl.tabsX[p] + fv.uini);
  }
  // tasw1: Z-node (fd3d_deriv.f90:1204-1226)
  {
    const int pip = p + 1, pkp = p + g.nxt, pikp = p + 1 + g.nxt;
    const double wc = (double)f.w[q], uz = (double)fl.uZ[p];
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
    f.w[q] = (float)((((double)fl.tz[p] + (double)fl.T0Z[p]) * (-((sinh(uw) * fv.v0) * ex))) /
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

// ---------------------------------------------------------------------------
// Split-field PML velocities, all regions in one launch (pml_uvw,
// fd3d_deriv.f90:1927-2299): p <- [p(1-w) + ((dt/rho)*dsigma)/dh]/(1+w),
// field = sum of splits.  u and w to j = je_uw, v to je_v.
// ---------------------------------------------------------------------------
__device__ inline void pml_vel_body(const Dims &g, const Coef &c, const Fields &f, const Pml &pm, int gbx, int lx) {
  const int r = pml_region(pm, gbx);
  const PmlRegion &R = pm.r[r];
  const int local = (gbx - R.blk0) * kPmlBlock + lx;
  if (local >= R.ncell) return;
  const int i2 = local % R.ni + 1;
  const int t = local / R.ni;
  const int j2 = t % R.nj + 1;
  const int k2 = t / R.nj + 1;
  const int i = R.ib + i2 - 1, j = R.jb + j2 - 1, k = R.kb + k2 - 1;
  const int q = I3(g, i, j, k), sy = g.sy, sz = g.sz;
  const float *pr = pm.prof + R.prof_off;
  const int nI = R.ni + 1, nJ = R.nj + 1, nK = R.nk + 1;
  const float *amx = pr, *apx = pr + nI, *amxS = pr + 2 * nI, *apxS = pr + 3 * nI;
  const float *amy = pr + 4 * nI, *apy = amy + nJ;
  const float *amz = pr + 4 * nI + 4 * nJ, *apz = amz + nK;
  float *s = pm.splits + R.split_off + local;
  const long long nc = R.ncell;
  const float dt = c.dt, dh = c.dh;
  const float dtd = dt / f.d[q];
  float pt, s1, s2, s3;
  // u (x, y, z profiles)
  s1 = s[0 * nc]; s2 = s[1 * nc]; s3 = s[2 * nc];
  pt = s1 * amx[i2] + (dtd * (f.xx[q] - f.xx[q - 1])) / dh;
  s1 = pt / apx[i2];
  pt = s2 * amy[j2] + (dtd * (f.xy[q] - f.xy[q - sy])) / dh;
  s2 = pt / apy[j2];
  pt = s3 * amz[k2] + (dtd * (f.xz[q] - f.xz[q - sz])) / dh;
  s3 = pt / apz[k2];
  s[0 * nc] = s1; s[1 * nc] = s2; s[2 * nc] = s3;
  f.u[q] = (s1 + s2) + s3;
  // w (xS, y, z profiles)
  s1 = s[6 * nc]; s2 = s[7 * nc]; s3 = s[8 * nc];
  pt = s1 * amxS[i2] + (dtd * (f.xz[q + 1] - f.xz[q])) / dh;
  s1 = pt / apxS[i2];
  pt = s2 * amy[j2] + (dtd * (f.yz[q] - f.yz[q - sy])) / dh;
  s2 = pt / apy[j2];
  pt = s3 * amz[k2] + (dtd * (f.zz[q + sz] - f.zz[q])) / dh;
  s3 = pt / apz[k2];
  s[6 * nc] = s1; s[7 * nc] = s2; s[8 * nc] = s3;
  f.w[q] = (s1 + s2) + s3;
  if (j <= R.je_v) {
    s1 = s[3 * nc]; s2 = s[4 * nc]; s3 = s[5 * nc];
    pt = s1 * amxS[i2] + (dtd * (f.xy[q + 1] - f.xy[q])) / dh;
    s1 = pt / apxS[i2];
    pt = s2 * amy[j2] + (dtd * (f.yy[q + sy] - f.yy[q])) / dh;
    s2 = pt / apy[j2];
    pt = s3 * amz[k2] + (dtd * (f.yz[q] - f.yz[q - sz])) / dh;
    s3 = pt / apz[k2];
    s[3 * nc] = s1; s[4 * nc] = s2; s[5 * nc] = s3;
    f.v[q] = (s1 + s2) + s3;
  }
}

// ---------------------------------------------------------------------------
// Free surface ghost velocities (fuvw, fd3d_deriv.f90:1018-1062), in two
// launches because w at the ghost row reads the new ghost u and v.
// ---------------------------------------------------------------------------
template <int BY>
__device__ inline void fuvw_uv_body(const Dims &g, const Fields &f, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int j = 1 + gby * BY + ly;
  if (i < 2 || i > g.nxt - 1 || j > g.nyt) return;
  const int ks = g.nzt - g.nfs;
  const int q = I3(g, i, j, ks), sz = g.sz, sy = g.sy;
  f.u[q + sz] = f.u[q] - (f.w[q] - f.w[q - 1]);
  if (j >= 2 && j <= g.nyt - 1) f.v[q + sz] = f.v[q] - (f.w[q + sy] - f.w[q]);
}

template <int BY>
__device__ inline void fuvw_w_body(const Dims &g, const Fields &f, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int j = 1 + gby * BY + ly;
  if (i < 2 || i > g.nxt - 2 || j < 2 || j > g.nyt) return;
  const int ks = g.nzt - g.nfs;
  const int q = I3(g, i, j, ks), sz = g.sz, sy = g.sy;
  const float xl = f.lam[q + sz], xm = f.mu[q + sz];
  const float a = 2.0f * xl, b = xl + 2.0f * xm;
  const float *u = f.u, *v = f.v;
  const float s = ((((((u[q + 1] - u[q]) + u[q + 1 + sz]) - u[q + sz]) + v[q]) - v[q - sy]) +
                   v[q + sz]) - v[q - sy + sz];
  f.w[q + sz] = f.w[q - sz] - (a / b) * s;
}

// ---------------------------------------------------------------------------
// Interior stresses: dstres (fd3d_deriv.f90:513-622), same range as the
// interior velocities.  Bytes per cell: reads u, v, w, lam, mu + 6 stresses;
// writes 6 stresses.
// ---------------------------------------------------------------------------
__device__ inline void str_interior_body(const Dims &g, const Coef &c, const Fields &f, int gbx, int gby, int gbz, int lx, int ly) {
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int j = g.nabc + 1 + gby * BYF + ly;
  const int k = g.nabc + 1 + gbz;
  if (i > g.nxt - g.nabc || j > g.nyt - 2) return;
  const int q = I3(g, i, j, k), sy = g.sy, sz = g.sz;
  const float c1 = c.c1, c2 = c.c2, dth = c.dth;
  const float *u = f.u, *v = f.v, *w = f.w, *mu = f.mu;
  const float xl = f.lam[q], xm = mu[q];
  const float a = xl + 2.0f * xm, b = xl;
  const float dux1 = c1 * (u[q + 1] - u[q]), dux2 = c2 * (u[q + 2] - u[q - 1]);
  const float dvy1 = c1 * (v[q] - v[q - sy]), dvy2 = c2 * (v[q + sy] - v[q - 2 * sy]);
  const float dwz1 = c1 * (w[q] - w[q - sz]), dwz2 = c2 * (w[q + sz] - w[q - 2 * sz]);
  f.xx[q] = (f.xx[q] + (dth * a) * (dux1 + dux2)) + (dth * b) * (((dvy1 + dvy2) + dwz1) + dwz2);
  f.yy[q] = (f.yy[q] + (dth * a) * (dvy1 + dvy2)) + (dth * b) * (((dux1 + dux2) + dwz1) + dwz2);
  f.zz[q] = (f.zz[q] + (dth * a) * (dwz1 + dwz2)) + (dth * b) * (((dux1 + dux2) + dvy1) + dvy2);
  float xmu = 0.5f * (xm + mu[q + sy]);
  f.xy[q] = f.xy[q] + (dth * xmu) * (((c1 * (u[q + sy] - u[q]) + c2 * (u[q + 2 * sy] - u[q - sy])) +
                                      c1 * (v[q] - v[q - 1])) + c2 * (v[q + 1] - v[q - 2]));
  xmu = 0.5f * (xm + mu[q + sz]);
  f.xz[q] = f.xz[q] + (dth * xmu) * (((c1 * (u[q + sz] - u[q]) + c2 * (u[q + 2 * sz] - u[q - sz])) +
                                      c1 * (w[q] - w[q - 1])) + c2 * (w[q + 1] - w[q - 2]));
  xmu = 0.5f * (xm + mu[q + 1 + sy + sz]);
  f.yz[q] = f.yz[q] + (dth * xmu) * (((c1 * (v[q + sz] - v[q]) + c2 * (v[q + 2 * sz] - v[q - sz])) +
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
__device__ inline void str_fault_body(const Dims &g, const Coef &c, const Fields &f, const Fault &fl, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int k = 1 + gby * BY + ly;
  if (i > g.nxt || k > g.nzt) return;
  const int sy = g.sy, sz = g.sz;
  const int q = I3(g, i, g.nyt - 1, k);
  const bool fault = i >= g.nabc + 1 && i <= g.nxt - g.nabc && k >= g.nabc + 1 && k <= g.nzt - g.nfs;
  if (!fault) {
    f.xy[q + sy] = f.xy[q];
    f.yz[q + sy] = f.yz[q];
    return;
  }
  const float *u = f.u, *v = f.v, *w = f.w, *mu = f.mu;
  {
    const float dth = c.dth;
    const float xl = f.lam[q], xm = mu[q];
    const float a = xl + 2.0f * xm, b = xl;
    f.xx[q] = (f.xx[q] + (dth * a) * (u[q + 1] - u[q])) + (dth * b) * (((v[q] - v[q - sy]) + w[q]) - w[q - sz]);
    f.yy[q] = (f.yy[q] + (dth * a) * (v[q] - v[q - sy])) + (dth * b) * (((u[q + 1] - u[q]) + w[q]) - w[q - sz]);
    f.zz[q] = (f.zz[q] + (dth * a) * (w[q] - w[q - sz])) + (dth * b) * (((u[q + 1] - u[q]) + v[q]) - v[q - sy]);
    const float dthh = c.dthh;
    float xmu = xm + mu[q + sy];
    const float xyn = f.xy[q] + (dthh * xmu) * (((u[q + sy] - u[q]) + v[q]) - v[q - 1]);
    f.xy[q] = xyn;
    xmu = xm + mu[q + sz];
    f.xz[q] = f.xz[q] + (dthh * xmu) * (((u[q + sz] - u[q]) + w[q]) - w[q - 1]);
    xmu = xm + mu[q + 1 + sy + sz];
    const float yzn = f.yz[q] + (dthh * xmu) * (((v[q + sz] - v[q]) + w[q + sy]) - w[q]);
    f.yz[q] = yzn;
    f.xy[q + sy] = xyn;
    f.yz[q + sy] = yzn;
  }
  {
    const int qf = q + sy;
    const float c1 = c.c1, c2 = c.c2;
    float xmu = mu[qf] + mu[qf + sz];
    f.xz[qf] = f.xz[qf] + (c.dthh * xmu) *
                              (((c1 * (u[qf + sz] - u[qf]) + c2 * (u[qf + 2 * sz] - u[qf - sz])) +
                                c1 * (w[qf] - w[qf - 1])) + c2 * (w[qf + 1] - w[qf - 2]));
    const float dth = c.dth;
    const float xl = f.lam[qf], xm = mu[qf];
    const float a = xl + 2.0f * xm, b = xl;
    const float diff1 = c1 * (u[qf + 1] - u[qf]) + c2 * (u[qf + 2] - u[qf - 1]);
    const float diff3 = c1 * (w[qf] - w[qf - sz]) + c2 * (w[qf + sz] - w[qf - 2 * sz]);
    const float vm = v[qf - sy];
    const int p = I2(g, i, k);
    const float v1t = vm - (b * (diff1 + diff3)) / (2.0f * a);
    fl.v1t[p] = v1t;
    const float dv = v1t - vm;
    f.xx[qf] = (f.xx[qf] + (dth * a) * diff1) + (dth * b) * (2.0f * dv + diff3);
    f.yy[qf] = (f.yy[qf] + ((dth * a) * 2.0f) * dv) + (dth * b) * (diff1 + diff3);
    f.zz[qf] = (f.zz[qf] + (dth * a) * diff3) + (dth * b) * (diff1 + 2.0f * dv);
  }
}

// ---------------------------------------------------------------------------
// Split-field PML stresses, all regions in one launch (pml_xyz,
// fd3d_deriv.f90:2304-3145), with the top-layer xz/yz split zeroing of
// regions 1 (always) and 2, 3 (free surface only) fused after the update.
// ---------------------------------------------------------------------------
__device__ inline void pml_str_body(const Dims &g, const Coef &c, const Fields &f, const Pml &pm, int gbx, int lx) {
  const int r = pml_region(pm, gbx);
  const PmlRegion &R = pm.r[r];
  const int local = (gbx - R.blk0) * kPmlBlock + lx;
  if (local >= R.ncell) return;
  const int i2 = local % R.ni + 1;
  const int t = local / R.ni;
  const int j2 = t % R.nj + 1;
  const int k2 = t / R.nj + 1;
  const int i = R.ib + i2 - 1, j = R.jb + j2 - 1, k = R.kb + k2 - 1;
  const int q = I3(g, i, j, k), sy = g.sy, sz = g.sz;
  const float *pr = pm.prof + R.prof_off;
  const int nI = R.ni + 1, nJ = R.nj + 1, nK = R.nk + 1;
  const float *amx = pr, *apx = pr + nI, *amxS = pr + 2 * nI, *apxS = pr + 3 * nI;
  const float *amy = pr + 4 * nI, *apy = amy + nJ, *amyS = amy + 2 * nJ, *apyS = amy + 3 * nJ;
  const float *amz = pr + 4 * nI + 4 * nJ, *apz = amz + nK, *amzS = amz + 2 * nK, *apzS = amz + 3 * nK;
  float *s = pm.splits + R.split_off + local;
  const long long nc = R.ncell;
  const float dt = c.dt, dh = c.dh;
  const float *u = f.u, *v = f.v, *w = f.w, *mu = f.mu;
  const float xl = f.lam[q], xm = mu[q];
  const float a = xl + 2.0f * xm, b = xl;
  const float diff1 = u[q + 1] - u[q], diff2 = v[q] - v[q - sy], diff3 = w[q] - w[q - sz];
  const float ax = amxS[i2], bx = apxS[i2], ay = amy[j2], by = apy[j2], az = amz[k2], bz = apz[k2];
  float pt, s1, s2, s3;
  s1 = s[9 * nc]; s2 = s[10 * nc]; s3 = s[11 * nc];
  pt = s1 * ax + ((dt * a) * diff1) / dh; s1 = pt / bx;
  pt = s2 * ay + ((dt * b) * diff2) / dh; s2 = pt / by;
  pt = s3 * az + ((dt * b) * diff3) / dh; s3 = pt / bz;
  s[9 * nc] = s1; s[10 * nc] = s2; s[11 * nc] = s3;
  f.xx[q] = (s1 + s2) + s3;
  s1 = s[12 * nc]; s2 = s[13 * nc]; s3 = s[14 * nc];
  pt = s1 * ax + ((dt * b) * diff1) / dh; s1 = pt / bx;
  pt = s2 * ay + ((dt * a) * diff2) / dh; s2 = pt / by;
  pt = s3 * az + ((dt * b) * diff3) / dh; s3 = pt / bz;
  s[12 * nc] = s1; s[13 * nc] = s2; s[14 * nc] = s3;
  f.yy[q] = (s1 + s2) + s3;
  s1 = s[15 * nc]; s2 = s[16 * nc]; s3 = s[17 * nc];
  pt = s1 * ax + ((dt * b) * diff1) / dh; s1 = pt / bx;
  pt = s2 * ay + ((dt * b) * diff2) / dh; s2 = pt / by;
  pt = s3 * az + ((dt * a) * diff3) / dh; s3 = pt / bz;
  s[15 * nc] = s1; s[16 * nc] = s2; s[17 * nc] = s3;
  f.zz[q] = (s1 + s2) + s3;
  const bool zero_top = R.ztop && k2 == R.nk;
  if (j <= R.je_v) {
    const float xmu = (xm + mu[q + sy]) / 2.0f;
    s1 = s[18 * nc]; s2 = s[19 * nc];
    pt = s1 * amyS[j2] + ((dt * xmu) * (u[q + sy] - u[q])) / dh; s1 = pt / apyS[j2];
    pt = s2 * amx[i2] + ((dt * xmu) * (v[q] - v[q - 1])) / dh; s2 = pt / apx[i2];
    s[18 * nc] = s1; s[19 * nc] = s2;
    f.xy[q] = s1 + s2;
  }
  if (j <= R.je_xz) {
    const float xmu = (xm + mu[q + sz]) / 2.0f;
    s1 = s[20 * nc]; s2 = s[21 * nc];
    pt = s1 * amzS[k2] + ((dt * xmu) * (u[q + sz] - u[q])) / dh; s1 = pt / apzS[k2];
    pt = s2 * amx[i2] + ((dt * xmu) * (w[q] - w[q - 1])) / dh; s2 = pt / apx[i2];
    f.xz[q] = s1 + s2;
    if (zero_top && j2 <= R.nzx) { s1 = 0.0f; s2 = 0.0f; }
    s[20 * nc] = s1; s[21 * nc] = s2;
  }
  if (j <= R.je_v) {
    const float xmu = (xm + mu[q + 1 + sy + sz]) / 2.0f;
    s1 = s[22 * nc]; s2 = s[23 * nc];
    pt = s1 * amzS[k2] + ((dt * xmu) * (v[q + sz] - v[q])) / dh; s1 = pt / apzS[k2];
    pt = s2 * amyS[j2] + ((dt * xmu) * (w[q + sy] - w[q])) / dh; s2 = pt / apyS[j2];
    f.yz[q] = s1 + s2;
    if (zero_top && j2 <= R.nzy) { s1 = 0.0f; s2 = 0.0f; }
    s[22 * nc] = s1; s[23 * nc] = s2;
  }
}

// Free-surface stress imaging (fres, fd3d_deriv.f90:1067-1104).
template <int BY>
__device__ inline void fres_body(const Dims &g, const Fields &f, int gbx, int gby, int lx, int ly) {
  const int i = 1 + gbx * BX + lx;
  const int j = 1 + gby * BY + ly;
  if (i > g.nxt || j > g.nyt) return;
  const int ks = g.nzt - g.nfs;
  const int q = I3(g, i, j, ks), sz = g.sz;
  f.zz[q + sz] = -f.zz[q];
  f.zz[q + 2 * sz] = -f.zz[q - sz];
  f.xz[q + sz] = -f.xz[q - sz];
  f.xz[q + 2 * sz] = -f.xz[q - 2 * sz];
  f.yz[q + sz] = -f.yz[q - sz];
  f.yz[q + 2 * sz] = -f.yz[q - 2 * sz];
  f.xz[q] = 0.0f;
  f.yz[q] = 0.0f;
}

// ---------------------------------------------------------------------------
// Trial traction (fd3d_theo.f90:232-308) on fault nodes.
// ---------------------------------------------------------------------------
template <int BY>
__device__ inline void traction_body(const Dims &g, const Coef &c, const Fields &f, const Fault &fl, int gbx, int gby, int lx, int ly) {
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int k = g.nabc + 1 + gby * BY + ly;
  const int ks = g.nzt - g.nfs;
  if (i > g.nxt - g.nabc || k > ks) return;
  const int q = I3(g, i, g.nyt, k), sy = g.sy, sz = g.sz, p = I2(g, i, k);
  const float c1 = c.c1, c2 = c.c2, dht = c.dht;
  const float *xx = f.xx, *zz = f.zz, *xy = f.xy, *yz = f.yz, *xz = f.xz;
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
  const float dd = f.d[q];
  if (!top_free)
    fl.tz[p] = -rf - ((0.5f * dd) * dht) * f.w[q];
  else
    fl.tz[p] = -rf - ((((0.5f * dd) * dht) * f.w[q]) / 2.0f);
  fl.avdz[p] = pdz;
  const float pdx = ((c1 * (xx[q] - xx[q - 1]) + c2 * (xx[q + 1] - xx[q - 2])) / 2.0f +
                     (c1 * (xz[q] - xz[q - sz]) + c2 * (xz[q + sz] - xz[q - 2 * sz])) / 2.0f) - xy[q - sy];
  av = c.damp_s * (pdx - fl.avdx[p]);
  rf = pdx + av;
  fl.RFx[p] = rf;
  fl.tx[p] = -rf - ((0.5f * dd) * dht) * f.u[q];
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
__device__ inline void friction_body(const Dims &g, const Coef &c, const Fields &f, const Fault &fl, const FvwScalars &fv, const Ctl *ctl, int gbx, int gby, int lx, int ly) {
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int k = g.nabc + 1 + gby * BY + ly;
  if (i > g.nxt - g.nabc || k > g.nzt - g.nfs) return;
  const int p = I2(g, i, k);
  const int q = I3(g, i, g.nyt, k), sz = g.sz;
  const float dt = c.dt;
  const float time = (float)ctl->it * dt;
  const float rup_tresh = 1.e-3f;
  const float *u = f.u, *w = f.w;
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
__device__ inline void outputs_body(const Dims &g, const Fields &f, const Fault &fl, const FvwScalars &fv, const Outputs &o, Ctl *ctl, int gbx, int gby, int lx, int ly) {
  const int it = ctl->it;
  const int slot = (it - 1) % o.ring_steps;
  const int ks = g.nzt - g.nfs;
  const long long slab = (long long)g.nxT * g.nzT;
  const int i = g.nabc + 1 + gbx * BX + lx;
  const int k = g.nabc + 1 + gby * BY + ly;
  if (i > g.nxt - g.nabc || k > ks) return;
  const int p = I2(g, i, k), n = g.nxt;
  const int q = I3(g, i, g.nyt, k), sz = g.sz;
  const float *u = f.u;
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
  const float srz = -(2.0f * f.w[q]);
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

// Station samples of one step into the device ring (fd3d_theo.f90:456-467),
// one thread per station.
__device__ inline void stations_body(const Dims &g, const Fields &f, const Outputs &o, const Ctl *ctl, int s) {
  if (s >= o.nsta) return;
  const int slot = (ctl->it - 1) % o.ring_steps;
  const int q = I3(g, o.sta[3 * s], o.sta[3 * s + 1], o.sta[3 * s + 2]);
  float *dst = o.seis + ((long long)slot * o.nsta + s) * 3;
  dst[0] = f.u[q];
  dst[1] = f.v[q];
  dst[2] = f.w[q];
}

// TPV104 smooth nucleation (fd3d_theo.f90:515-542), FP64.
template <int BY>
__device__ inline void nucleation_body(const Dims &g, const Coef &c, const Fault &fl, const FvwScalars &fv, const Ctl *ctl, int gbx, int gby, int lx, int ly) {
  const double time = (double)((float)ctl->it * c.dt);
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
