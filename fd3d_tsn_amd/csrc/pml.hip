// Split-field PML sweeps (pml_uvw, fd3d_deriv.f90:1927-2299; pml_xyz,
// :2304-3145) as tiled y-marches (tuning mode bit 1).
//
// The PML cells are covered by 64 x kPmlTZ x-z tiles on the grid of the
// interior tiles (origins i = nabc+1 + 64 bx, k = nabc+1 + kPmlTZ bz), so
// tile rows start on the cache-line boundaries of the interior.  A block
// takes one (bx, bz, j0, j1) entry of the tile list built on the host
// (build_pml_tiles): tiles that meet an x- or z-PML box march every row
// of it, the others only the y-PML rows 2..nabc.  As in the interior
// sweeps, the x-z plane of the current row is staged in LDS (double
// buffered, one barrier per row) with the one-cell halos the second-order
// PML differences read, and the y neighbours come from register windows,
// so no neighbour value is fetched twice.  Each lane finds its cell's
// region (the reference's five disjoint boxes) row by row; the update of a
// cell is that of pml_vel_body / pml_str_body (stages.h), operand for
// operand, so the result is bit-identical to the flat sweeps.
#include "stages.h"

namespace fd3d {
namespace {

constexpr int TX = 64, TZ = kPmlTZ, NT = TX * TZ;
// one plane: x in [-1, TX], z in [-1, TZ] relative to the tile origin
constexpr int PW = TX + 2, PH = TZ + 2, PN = PW * PH;
__device__ inline int pat(int x, int z) { return (z + 1) * PW + (x + 1); }

// Tile of this block and the lane's cell column.
struct PTile {
  int i, k, j0, j1, i0, k0;
  bool in_arr;          // (i, k) inside the array
  unsigned boxes;       // bit r: (i, k) inside region r's x-z box
};
__device__ inline PTile ptile(const Dims &g, const Pml &pm, int tx, int tz) {
  const int4 e = pm.tiles[blockIdx.x];
  PTile T;
  T.i0 = g.nabc + 1 + TX * e.x;
  T.k0 = g.nabc + 1 + TZ * e.y;
  T.i = T.i0 + tx;
  T.k = T.k0 + tz;
  T.j0 = e.z;
  T.j1 = e.w;
  T.in_arr = T.i >= 1 && T.i <= g.nxt && T.k >= 1 && T.k <= g.nzt;
  T.boxes = 0;
#pragma unroll
  for (int r = 0; r < kMaxRegions; ++r)
    if (r < pm.nreg && T.i >= pm.r[r].ib && T.i <= pm.r[r].ie && T.k >= pm.r[r].kb && T.k <= pm.r[r].ke)
      T.boxes |= 1u << r;
  return T;
}
// Region of the lane's cell at row j, -1 if none.
__device__ inline int pregion(const Pml &pm, unsigned boxes, int j) {
  int reg = -1;
#pragma unroll
  for (int r = kMaxRegions - 1; r >= 0; --r)
    if ((boxes >> r) & 1u)
      if (j >= pm.r[r].jb && j <= pm.r[r].je_uw) reg = r;
  return reg;
}

// ---------------------------------------------------------------------------
// Velocities.  Planes of row j: xx (x-1), xy (x+1), xz (x+1, z-1), yz
// (z-1), zz (z+1); windows: xy(j-1), yz(j-1), yy(j), yy(j+1).
// ---------------------------------------------------------------------------
__device__ inline void pml_vel_cell(const Dims &g, const Coef &c, const Fields &fn, const PmlRegion &R, const float *pr,
                                    const float (*L)[PN], int a, int i, int j, int k, float dtd, float xym, float yzmy,
                                    float yy0, float yyp) {
  const int i2 = i - R.ib + 1, j2 = j - R.jb + 1, k2 = k - R.kb + 1;
  const int l = (i2 - 1) + R.ni * ((j2 - 1) + R.nj * (k2 - 1));
  const int q = I3(g, i, j, k), nc = R.ncell;
  float *sp = R.sp;
  const int nI = R.ni + 1, nJ = R.nj + 1, nK = R.nk + 1;
  const int amx = R.prof_off, apx = amx + nI, amxS = amx + 2 * nI, apxS = amx + 3 * nI;
  const int amy = amx + 4 * nI, apy = amy + nJ;
  const int amz = amx + 4 * nI + 4 * nJ, apz = amz + nK;
  const float dh = c.dh;
  const bool hasv = j <= R.je_v;
  const float su1 = ld32(sp, l), su2 = ld32(sp, l + nc), su3 = ld32(sp, l + 2 * nc);
  const float sw1 = ld32(sp, l + 6 * nc), sw2 = ld32(sp, l + 7 * nc), sw3 = ld32(sp, l + 8 * nc);
  float sv1 = 0, sv2 = 0, sv3 = 0;
  if (hasv) { sv1 = ld32(sp, l + 3 * nc); sv2 = ld32(sp, l + 4 * nc); sv3 = ld32(sp, l + 5 * nc); }
  const float pamx = ld32(pr, amx + i2), papx = ld32(pr, apx + i2), pamxS = ld32(pr, amxS + i2),
              papxS = ld32(pr, apxS + i2);
  const float pamy = ld32(pr, amy + j2), papy = ld32(pr, apy + j2), pamz = ld32(pr, amz + k2),
              papz = ld32(pr, apz + k2);
  const float *Lxx = L[0], *Lxy = L[1], *Lxz = L[2], *Lyz = L[3], *Lzz = L[4];
  const float xx0 = Lxx[a], xxm = Lxx[a - 1], xy0 = Lxy[a], xz0 = Lxz[a], xzmz = Lxz[a - PW], xzp = Lxz[a + 1];
  const float yz0 = Lyz[a], zz0 = Lzz[a], zzp = Lzz[a + PW];
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
    const float xyp = Lxy[a + 1], yzmz = Lyz[a - PW];
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

__global__ void __launch_bounds__(NT) k_pml_vel_t(Dims g, Coef c, Fields f0, Fields f1, Pml pm, const Ctl *ctl) {
  __shared__ float lds[2][5][PN];
  const FieldPair FP = field_pair(f0, f1, ctl);
  const Fields &fo = FP.fo, &fn = FP.fn;
  const int tx = threadIdx.x, tz = threadIdx.y;
  const int wv = __builtin_amdgcn_readfirstlane(tz);
  const PTile T = ptile(g, pm, tx, tz);
  const int sy = g.sy;
  const int a = pat(tx, tz);
  const int qc = T.in_arr ? I3(g, T.i, 1, T.k) : 0;
  const float dtd = c.dt / fo.d[T.in_arr ? T.k : 1];
  // halo load of this wavefront (wv < 6): xx column -1, xy column TX,
  // xz column TX (lanes e < TZ, row e), xz row -1, yz row -1, zz row TZ
  // (lane = column)
  // hf >= 0: this lane holds a halo cell (LDS slot hl); hon: the cell is
  // inside the array (else the slot gets 0)
  int hf = -1, hx = 0, hz = 0;
  if (wv < 3 && tx < TZ) { hf = wv; hx = wv == 0 ? -1 : TX; hz = tx; }
  else if (wv >= 3 && wv < 6) { hf = wv - 1; hx = tx; hz = wv == 5 ? TZ : -1; }
  const int hi = T.i0 + hx, hk = T.k0 + hz;
  const bool hon = hf >= 0 && hi >= 1 && hi <= g.nxt && hk >= 1 && hk <= g.nzt;
  const int hq = hon ? I3(g, hi, 1, hk) : 0;
  const int hl = hf >= 0 ? pat(hx, hz) : 0;
  const float *hsrc = hf == 0 ? fo.xx : hf == 1 ? fo.xy : hf == 2 ? fo.xz : hf == 3 ? fo.yz : fo.zz;
  // windows at the first row
  const int j0 = T.j0;
  float xym = 0, yzmy = 0, yy0 = 0;
  if (T.in_arr && j0 - 1 >= 1) { xym = ld32(fo.xy, qc + (j0 - 2) * sy); yzmy = ld32(fo.yz, qc + (j0 - 2) * sy); }
  if (T.in_arr) yy0 = ld32(fo.yy, qc + (j0 - 1) * sy);
  int buf = 0;
  for (int j = j0; j <= T.j1; ++j) {
    const int jo = (j - 1) * sy, q = qc + jo;
    float(*L)[PN] = lds[buf];
    float cxx = 0, cxy = 0, cxz = 0, cyz = 0, czz = 0, yyp = 0, hv = 0;
    if (T.in_arr) {
      cxx = ld32(fo.xx, q); cxy = ld32(fo.xy, q); cxz = ld32(fo.xz, q); cyz = ld32(fo.yz, q); czz = ld32(fo.zz, q);
      if (j + 1 <= g.nyt) yyp = ld32(fo.yy, q + sy);
    }
    if (hon) hv = ld32(hsrc, hq + jo);
    L[0][a] = cxx; L[1][a] = cxy; L[2][a] = cxz; L[3][a] = cyz; L[4][a] = czz;
    if (hf >= 0) L[hf][hl] = hv;
    __syncthreads();
    // one pass per region among the wavefront's lanes (almost always one),
    // the region chosen wave-uniform so its table entry is a scalar load
    for (int reg = pregion(pm, T.boxes, j);;) {
      const unsigned long long m = __ballot(reg >= 0);
      if (m == 0) break;
      const int rw = __builtin_amdgcn_readlane(reg, __ffsll((long long)m) - 1);
      if (reg == rw) {
        pml_vel_cell(g, c, fn, pm.rdev[rw], pm.prof, L, a, T.i, j, T.k, dtd, xym, yzmy, yy0, yyp);
        reg = -1;
      }
    }
    xym = cxy; yzmy = cyz; yy0 = yyp;
    buf ^= 1;
  }
}

// ---------------------------------------------------------------------------
// Stresses.  Planes of row j: u (x+1, z+1), v (x-1, z+1), w (x-1, z-1);
// windows: v(j-1), u(j+1), w(j+1).
// ---------------------------------------------------------------------------
__device__ inline void pml_str_cell(const Dims &g, const Coef &c, const Fields &fn, const PmlRegion &R, const float *pr,
                                    const float (*L)[PN], int a, int i, int j, int k, float xl, float xm, float mk1,
                                    float vmy, float uy, float wy) {
  const int i2 = i - R.ib + 1, j2 = j - R.jb + 1, k2 = k - R.kb + 1;
  const int l = (i2 - 1) + R.ni * ((j2 - 1) + R.nj * (k2 - 1));
  const int q = I3(g, i, j, k), nc = R.ncell;
  float *sp = R.sp;
  const int nI = R.ni + 1, nJ = R.nj + 1, nK = R.nk + 1;
  const int amx = R.prof_off, apx = amx + nI, amxS = amx + 2 * nI, apxS = amx + 3 * nI;
  const int amy = amx + 4 * nI, apy = amy + nJ, amyS = amy + 2 * nJ, apyS = amy + 3 * nJ;
  const int amz = amx + 4 * nI + 4 * nJ, apz = amz + nK, amzS = amz + 2 * nK, apzS = amz + 3 * nK;
  const float dt = c.dt, dh = c.dh;
  const bool hxy = j <= R.je_v, hxz = j <= R.je_xz;
  const float *Lu = L[0], *Lv = L[1], *Lw = L[2];
  const float u0 = Lu[a], up = Lu[a + 1], v0 = Lv[a], w0 = Lw[a], wmz = Lw[a - PW];
  const float ax = ld32(pr, amxS + i2), bx = ld32(pr, apxS + i2), ay = ld32(pr, amy + j2), by = ld32(pr, apy + j2),
              az = ld32(pr, amz + k2), bz = ld32(pr, apz + k2);
  const float sxx1 = ld32(sp, l + 9 * nc), sxx2 = ld32(sp, l + 10 * nc), sxx3 = ld32(sp, l + 11 * nc);
  const float syy1 = ld32(sp, l + 12 * nc), syy2 = ld32(sp, l + 13 * nc), syy3 = ld32(sp, l + 14 * nc);
  const float szz1 = ld32(sp, l + 15 * nc), szz2 = ld32(sp, l + 16 * nc), szz3 = ld32(sp, l + 17 * nc);
  float mxy = 0, vmx = 0, sxy1 = 0, sxy2 = 0, pamyS = 0, papyS = 0, pamx = 0, papx = 0;
  float mxz = 0, uz = 0, wmx = 0, sxz1 = 0, sxz2 = 0, pamzS = 0, papzS = 0;
  float myz = 0, vz = 0, syz1 = 0, syz2 = 0;
  if (hxy || hxz) { pamx = ld32(pr, amx + i2); papx = ld32(pr, apx + i2); }
  if (hxy) {
    mxy = xm; vmx = Lv[a - 1]; sxy1 = ld32(sp, l + 18 * nc); sxy2 = ld32(sp, l + 19 * nc);
    pamyS = ld32(pr, amyS + j2); papyS = ld32(pr, apyS + j2);
    myz = mk1; vz = Lv[a + PW]; syz1 = ld32(sp, l + 22 * nc); syz2 = ld32(sp, l + 23 * nc);
  }
  if (hxz) { mxz = mk1; uz = Lu[a + PW]; wmx = Lw[a - 1]; sxz1 = ld32(sp, l + 20 * nc); sxz2 = ld32(sp, l + 21 * nc); }
  if (hxy || hxz) { pamzS = ld32(pr, amzS + k2); papzS = ld32(pr, apzS + k2); }
  const float aa = xl + 2.0f * xm, b = xl;
  const float diff1 = up - u0, diff2 = v0 - vmy, diff3 = w0 - wmz;
  float pt, s1, s2, s3;
  pt = sxx1 * ax + ((dt * aa) * diff1) / dh; s1 = pt / bx;
  pt = sxx2 * ay + ((dt * b) * diff2) / dh; s2 = pt / by;
  pt = sxx3 * az + ((dt * b) * diff3) / dh; s3 = pt / bz;
  st32(sp, l + 9 * nc, s1); st32(sp, l + 10 * nc, s2); st32(sp, l + 11 * nc, s3);
  st32_nt(fn.xx, q, (s1 + s2) + s3);
  pt = syy1 * ax + ((dt * b) * diff1) / dh; s1 = pt / bx;
  pt = syy2 * ay + ((dt * aa) * diff2) / dh; s2 = pt / by;
  pt = syy3 * az + ((dt * b) * diff3) / dh; s3 = pt / bz;
  st32(sp, l + 12 * nc, s1); st32(sp, l + 13 * nc, s2); st32(sp, l + 14 * nc, s3);
  st32_nt(fn.yy, q, (s1 + s2) + s3);
  pt = szz1 * ax + ((dt * b) * diff1) / dh; s1 = pt / bx;
  pt = szz2 * ay + ((dt * b) * diff2) / dh; s2 = pt / by;
  pt = szz3 * az + ((dt * aa) * diff3) / dh; s3 = pt / bz;
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

__global__ void __launch_bounds__(NT) k_pml_str_t(Dims g, Coef c, Fields f0, Fields f1, Pml pm, const Ctl *ctl) {
  __shared__ float lds[2][3][PN];
  const FieldPair FP = field_pair(f0, f1, ctl);
  const Fields &fn = FP.fn;
  const int tx = threadIdx.x, tz = threadIdx.y;
  const int wv = __builtin_amdgcn_readfirstlane(tz);
  const PTile T = ptile(g, pm, tx, tz);
  const int sy = g.sy;
  const int a = pat(tx, tz);
  const int qc = T.in_arr ? I3(g, T.i, 1, T.k) : 0;
  const float *u = fn.u, *v = fn.v, *w = fn.w;
  // depth-layered medium: lam, mu of the lane's plane and mu of the next
  // (PML cells lie below plane nzt)
  const int km = (T.in_arr && T.k < g.nzt) ? T.k : 1;
  const float xl = fn.lam[km], xm = fn.mu[km], mk1 = fn.mu[km + 1];
  // halo load of this wavefront (wv < 6): u column TX, v column -1,
  // w column -1 (lanes e < TZ, row e), u row TZ, v row TZ, w row -1
  // (lane = column)
  int hf = -1, hx = 0, hz = 0;
  if (wv < 3 && tx < TZ) { hf = wv; hx = wv == 0 ? TX : -1; hz = tx; }
  else if (wv >= 3 && wv < 6) { hf = wv - 3; hx = tx; hz = wv == 5 ? -1 : TZ; }
  const int hi = T.i0 + hx, hk = T.k0 + hz;
  const bool hon = hf >= 0 && hi >= 1 && hi <= g.nxt && hk >= 1 && hk <= g.nzt;
  const int hq = hon ? I3(g, hi, 1, hk) : 0;
  const int hl = hf >= 0 ? pat(hx, hz) : 0;
  const float *hsrc = hf == 0 ? u : hf == 1 ? v : w;
  // windows at the first row: v(j0-1), u(j0), w(j0)
  const int j0 = T.j0;
  float vmy = 0, uc = 0, wc = 0;
  if (T.in_arr) {
    if (j0 - 1 >= 1) vmy = ld32(v, qc + (j0 - 2) * sy);
    uc = ld32(u, qc + (j0 - 1) * sy);
    wc = ld32(w, qc + (j0 - 1) * sy);
  }
  int buf = 0;
  for (int j = j0; j <= T.j1; ++j) {
    const int jo = (j - 1) * sy, q = qc + jo;
    float(*L)[PN] = lds[buf];
    float cv = 0, uy = 0, wy = 0, hv = 0;
    if (T.in_arr) {
      cv = ld32(v, q);
      if (j + 1 <= g.nyt) { uy = ld32(u, q + sy); wy = ld32(w, q + sy); }
    }
    if (hon) hv = ld32(hsrc, hq + jo);
    L[0][a] = uc; L[1][a] = cv; L[2][a] = wc;
    if (hf >= 0) L[hf][hl] = hv;
    __syncthreads();
    for (int reg = pregion(pm, T.boxes, j);;) {
      const unsigned long long m = __ballot(reg >= 0);
      if (m == 0) break;
      const int rw = __builtin_amdgcn_readlane(reg, __ffsll((long long)m) - 1);
      if (reg == rw) {
        pml_str_cell(g, c, fn, pm.rdev[rw], pm.prof, L, a, T.i, j, T.k, xl, xm, mk1, vmy, uy, wy);
        reg = -1;
      }
    }
    vmy = cv; uc = uy; wc = wy;
    buf ^= 1;
  }
}

}  // namespace

hipError_t launch_pml_tiled(bool stress, const Problem &p, hipStream_t s) {
  if (p.pml.ntiles == 0) return hipSuccess;
  if (stress)
    hipLaunchKernelGGL(k_pml_str_t, dim3(p.pml.ntiles), dim3(TX, TZ), 0, s, p.g, p.c, p.f[0], p.f[1], p.pml, p.ctl);
  else
    hipLaunchKernelGGL(k_pml_vel_t, dim3(p.pml.ntiles), dim3(TX, TZ), 0, s, p.g, p.c, p.f[0], p.f[1], p.pml, p.ctl);
  return hipGetLastError();
}

}  // namespace fd3d
