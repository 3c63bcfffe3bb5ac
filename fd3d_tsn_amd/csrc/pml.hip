// Split-field PML sweeps (pml_uvw, fd3d_deriv.f90:1927-2299; pml_xyz,
// :2304-3145) as z-marches.
//
// A thread owns one (i, j) column of a region and walks a chunk of
// kPmlZChunk planes along z; the z-derivative operands (xz, yz, zz for the
// velocities; u, v, w, mu for the stresses) stay in a two-plane register
// window, so each plane of them is read once instead of twice.  Threads of
// a block are consecutive (i, j) columns of one region, x fastest, so the
// split fields (dense (ni, nj, nk) boxes, i fastest) and the wavefield rows
// are read and written in contiguous runs.  All five regions share one
// launch; a block finds its region from a prefix table.
// The per-point arithmetic is that of k_pml_vel / k_pml_str in kernels.hip.
#include "stages.h"

namespace fd3d {
namespace {

struct RegionCell {
  int r, i2, j2, kc;
};

__device__ inline bool pml_zcell(const Pml &pm, RegionCell &c) {
  int r = 0;
#pragma unroll
  for (int n = 1; n < kMaxRegions; ++n)
    if (n < pm.nreg && (int)blockIdx.x >= pm.r[n].blkz0) r = n;
  const PmlRegion &R = pm.r[r];
  const int lb = blockIdx.x - R.blkz0;
  const int kc = lb / R.nxyb, xyb = lb - kc * R.nxyb;
  const int l = xyb * kPmlBlock + threadIdx.x;
  if (l >= R.ni * R.nj) return false;
  c.r = r;
  c.i2 = l % R.ni + 1;
  c.j2 = l / R.ni + 1;
  c.kc = kc;
  return true;
}

__global__ void __launch_bounds__(kPmlBlock) k_pml_vel_zm(Dims g, Coef c, Fields f0, Fields f1, Pml pm,
                                                                const Ctl *ctl) {
  const FieldPair FP = field_pair(f0, f1, ctl);
  const Fields &fo = FP.fo, &fn = FP.fn;
  RegionCell rc;
  if (!pml_zcell(pm, rc)) return;
  const PmlRegion &R = pm.r[rc.r];
  const int i2 = rc.i2, j2 = rc.j2;
  const int k2b = rc.kc * kPmlZChunk + 1, k2e = min(k2b + kPmlZChunk - 1, R.nk);
  const int i = R.ib + i2 - 1, j = R.jb + j2 - 1;
  const bool has_v = j <= R.je_v;
  const int sy = g.sy, sz = g.sz;
  const float *pr = pm.prof + R.prof_off;
  const int nI = R.ni + 1, nJ = R.nj + 1;
  const float amx = pr[i2], apx = pr[nI + i2], amxS = pr[2 * nI + i2], apxS = pr[3 * nI + i2];
  const float amy = pr[4 * nI + j2], apy = pr[4 * nI + nJ + j2];
  const float *amz = pr + 4 * nI + 4 * nJ, *apz = amz + (R.nk + 1);
  const long long nc = R.ncell, nxy = (long long)R.ni * R.nj;
  float *s = pm.splits + R.split_off + (i2 - 1) + (long long)R.ni * (j2 - 1) + nxy * (k2b - 1);
  const float dt = c.dt, dh = c.dh;
  int q = I3(g, i, j, R.kb + k2b - 1);
  float xzm1 = fo.xz[q - sz], xz0 = fo.xz[q];
  float zz0 = fo.zz[q], zzp1 = fo.zz[q + sz];
  float yzm1 = fo.yz[q - sz], yz0 = fo.yz[q];
  for (int k2 = k2b;; ++k2) {
    const float az = amz[k2], bz = apz[k2];
    const float dtd = dt / fo.d[R.kb + k2 - 1];
    float pt, s1, s2, s3;
    s1 = s[0]; s2 = s[nc]; s3 = s[2 * nc];
    pt = s1 * amx + (dtd * (fo.xx[q] - fo.xx[q - 1])) / dh; s1 = pt / apx;
    pt = s2 * amy + (dtd * (fo.xy[q] - fo.xy[q - sy])) / dh; s2 = pt / apy;
    pt = s3 * az + (dtd * (xz0 - xzm1)) / dh; s3 = pt / bz;
    s[0] = s1; s[nc] = s2; s[2 * nc] = s3;
    fn.u[q] = (s1 + s2) + s3;
    s1 = s[6 * nc]; s2 = s[7 * nc]; s3 = s[8 * nc];
    pt = s1 * amxS + (dtd * (fo.xz[q + 1] - xz0)) / dh; s1 = pt / apxS;
    pt = s2 * amy + (dtd * (yz0 - fo.yz[q - sy])) / dh; s2 = pt / apy;
    pt = s3 * az + (dtd * (zzp1 - zz0)) / dh; s3 = pt / bz;
    s[6 * nc] = s1; s[7 * nc] = s2; s[8 * nc] = s3;
    fn.w[q] = (s1 + s2) + s3;
    if (has_v) {
      s1 = s[3 * nc]; s2 = s[4 * nc]; s3 = s[5 * nc];
      pt = s1 * amxS + (dtd * (fo.xy[q + 1] - fo.xy[q])) / dh; s1 = pt / apxS;
      pt = s2 * amy + (dtd * (fo.yy[q + sy] - fo.yy[q])) / dh; s2 = pt / apy;
      pt = s3 * az + (dtd * (yz0 - yzm1)) / dh; s3 = pt / bz;
      s[3 * nc] = s1; s[4 * nc] = s2; s[5 * nc] = s3;
      fn.v[q] = (s1 + s2) + s3;
    }
    if (k2 == k2e) break;
    q += sz;
    s += nxy;
    xzm1 = xz0; xz0 = fo.xz[q];
    zz0 = zzp1; zzp1 = fo.zz[q + sz];
    yzm1 = yz0; yz0 = fo.yz[q];
  }
}

__global__ void __launch_bounds__(kPmlBlock) k_pml_str_zm(Dims g, Coef c, Fields f0, Fields f1, Pml pm,
                                                                const Ctl *ctl) {
  const FieldPair FP = field_pair(f0, f1, ctl);
  const Fields &fo = FP.fo, &fn = FP.fn;
  RegionCell rc;
  if (!pml_zcell(pm, rc)) return;
  const PmlRegion &R = pm.r[rc.r];
  const int i2 = rc.i2, j2 = rc.j2;
  const int k2b = rc.kc * kPmlZChunk + 1, k2e = min(k2b + kPmlZChunk - 1, R.nk);
  const int i = R.ib + i2 - 1, j = R.jb + j2 - 1;
  const bool has_v = j <= R.je_v, has_xz = j <= R.je_xz;
  const int sy = g.sy, sz = g.sz;
  const float *pr = pm.prof + R.prof_off;
  const int nI = R.ni + 1, nJ = R.nj + 1, nK = R.nk + 1;
  const float amx = pr[i2], apx = pr[nI + i2], ax = pr[2 * nI + i2], bx = pr[3 * nI + i2];
  const float ay = pr[4 * nI + j2], by = pr[4 * nI + nJ + j2];
  const float amyS = pr[4 * nI + 2 * nJ + j2], apyS = pr[4 * nI + 3 * nJ + j2];
  const float *amz = pr + 4 * nI + 4 * nJ, *apz = amz + nK, *amzS = amz + 2 * nK, *apzS = amz + 3 * nK;
  const long long nc = R.ncell, nxy = (long long)R.ni * R.nj;
  float *s = pm.splits + R.split_off + (i2 - 1) + (long long)R.ni * (j2 - 1) + nxy * (k2b - 1);
  const float dt = c.dt, dh = c.dh;
  const float *u = fn.u, *v = fn.v, *w = fn.w;
  int q = I3(g, i, j, R.kb + k2b - 1);
  float u0 = u[q], up1 = u[q + sz], v0 = v[q], vp1 = v[q + sz];
  float wm1 = w[q - sz], w0 = w[q];
  for (int k2 = k2b;; ++k2) {
    const float az = amz[k2], bz = apz[k2];
    // depth-layered medium: lam, mu of this plane and mu of the next
    const int k = R.kb + k2 - 1;
    const float xl = fo.lam[k], xm = fo.mu[k], mp1 = fo.mu[k + 1];
    const float a = xl + 2.0f * xm, b = xl;
    const float diff1 = u[q + 1] - u0, diff2 = v0 - v[q - sy], diff3 = w0 - wm1;
    float pt, s1, s2, s3;
    s1 = s[9 * nc]; s2 = s[10 * nc]; s3 = s[11 * nc];
    pt = s1 * ax + ((dt * a) * diff1) / dh; s1 = pt / bx;
    pt = s2 * ay + ((dt * b) * diff2) / dh; s2 = pt / by;
    pt = s3 * az + ((dt * b) * diff3) / dh; s3 = pt / bz;
    s[9 * nc] = s1; s[10 * nc] = s2; s[11 * nc] = s3;
    fn.xx[q] = (s1 + s2) + s3;
    s1 = s[12 * nc]; s2 = s[13 * nc]; s3 = s[14 * nc];
    pt = s1 * ax + ((dt * b) * diff1) / dh; s1 = pt / bx;
    pt = s2 * ay + ((dt * a) * diff2) / dh; s2 = pt / by;
    pt = s3 * az + ((dt * b) * diff3) / dh; s3 = pt / bz;
    s[12 * nc] = s1; s[13 * nc] = s2; s[14 * nc] = s3;
    fn.yy[q] = (s1 + s2) + s3;
    s1 = s[15 * nc]; s2 = s[16 * nc]; s3 = s[17 * nc];
    pt = s1 * ax + ((dt * b) * diff1) / dh; s1 = pt / bx;
    pt = s2 * ay + ((dt * b) * diff2) / dh; s2 = pt / by;
    pt = s3 * az + ((dt * a) * diff3) / dh; s3 = pt / bz;
    s[15 * nc] = s1; s[16 * nc] = s2; s[17 * nc] = s3;
    fn.zz[q] = (s1 + s2) + s3;
    const bool zero_top = R.ztop && k2 == R.nk;
    if (has_v) {
      const float xmu = (xm + xm) / 2.0f;
      s1 = s[18 * nc]; s2 = s[19 * nc];
      pt = s1 * amyS + ((dt * xmu) * (u[q + sy] - u0)) / dh; s1 = pt / apyS;
      pt = s2 * amx + ((dt * xmu) * (v0 - v[q - 1])) / dh; s2 = pt / apx;
      s[18 * nc] = s1; s[19 * nc] = s2;
      fn.xy[q] = s1 + s2;
    }
    if (has_xz) {
      const float xmu = (xm + mp1) / 2.0f;
      s1 = s[20 * nc]; s2 = s[21 * nc];
      pt = s1 * amzS[k2] + ((dt * xmu) * (up1 - u0)) / dh; s1 = pt / apzS[k2];
      pt = s2 * amx + ((dt * xmu) * (w0 - w[q - 1])) / dh; s2 = pt / apx;
      fn.xz[q] = s1 + s2;
      if (zero_top && j2 <= R.nzx) { s1 = 0.0f; s2 = 0.0f; }
      s[20 * nc] = s1; s[21 * nc] = s2;
    }
    if (has_v) {
      const float xmu = (xm + mp1) / 2.0f;
      s1 = s[22 * nc]; s2 = s[23 * nc];
      pt = s1 * amzS[k2] + ((dt * xmu) * (vp1 - v0)) / dh; s1 = pt / apzS[k2];
      pt = s2 * amyS + ((dt * xmu) * (w[q + sy] - w0)) / dh; s2 = pt / apyS;
      fn.yz[q] = s1 + s2;
      if (zero_top && j2 <= R.nzy) { s1 = 0.0f; s2 = 0.0f; }
      s[22 * nc] = s1; s[23 * nc] = s2;
    }
    if (k2 == k2e) break;
    q += sz;
    s += nxy;
    u0 = up1; up1 = u[q + sz];
    v0 = vp1; vp1 = v[q + sz];
    wm1 = w0; w0 = w[q];
  }
}

}  // namespace

hipError_t launch_pml_zm(bool stress, const Problem &p, hipStream_t s) {
  if (p.pml.nblkz == 0) return hipSuccess;
  if (stress)
    hipLaunchKernelGGL(k_pml_str_zm, dim3(p.pml.nblkz), dim3(kPmlBlock), 0, s, p.g, p.c, p.f[0], p.f[1], p.pml,
                       p.ctl);
  else
    hipLaunchKernelGGL(k_pml_vel_zm, dim3(p.pml.nblkz), dim3(kPmlBlock), 0, s, p.g, p.c, p.f[0], p.f[1], p.pml,
                       p.ctl);
  return hipGetLastError();
}

}  // namespace fd3d
