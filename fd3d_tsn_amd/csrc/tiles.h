// Tiled interior sweeps: the 4th-order velocity (dvel -> uxx1/vyy1/wzz1,
// fd3d_deriv.f90:24-193) and stress (dstres, :513-622) updates over
// i in [nabc+1, nxt-nabc], j in [nabc+1, nyt-2], k in [nabc+1, nzt-nfs].
//
// Layout of the work: a block owns a TX x TZ tile of the x-z plane (TX = 64
// or 128 lanes along x, TZ rows along z) and marches along y over NY rows.
//  * y neighbours come from per-thread register windows (each value is
//    loaded once per block);
//  * x and z neighbours come from the x-z plane of the current row staged
//    in LDS, double-buffered so one barrier per row suffices; centre cells
//    are written from the registers that already hold them, and the block
//    loads exactly the halo cells its stencils read (three columns or three
//    rows per field, on the sides that field's differences reach), so no
//    corner cell and no unused row of a neighbour tile is fetched.  Wider
//    tiles fetch fewer neighbour cache lines per cell for the x halos.
//  * the halo loads of a row are whole-wavefront loads of one field each
//    (the x-halo columns of a field, padded to 64 lanes, or 64 cells of a
//    z-halo row), so every access is a 32-bit offset from a wave-uniform
//    base; each lane's offsets and LDS slots are formed once per block.
// The bodies take the LDS buffer (Tile<TZ>::LDS_FLOATS floats) and the
// tile coordinates, so that single-problem launches (interior.hip) and the
// multi-domain batch kernels (batch.hip) share them.  A caller reusing the
// buffer for another tile must place a barrier between the two calls.
// Per-point arithmetic is the reference's, operand for operand (see
// stages.h), so results are bit-identical to the untiled form.
#pragma once
#include "fd3d_dev.h"

namespace fd3d {

template <int TX, int TZ>
struct Tile {
  static_assert((TX == 64 || TX == 128) && TX * TZ <= 1024 && 3 * TZ <= 64, "tile shape");
  static constexpr int W = TX + 4, H = TZ + 4, N = W * H;
  static constexpr int LDS_FLOATS = 2 * 5 * N;   // double-buffered planes, up to 5 fields
  __device__ static int at(int x, int z) { return (z + 2) * W + (x + 2); }
  // halo wavefront loads of one row: 3 x-halo loads (one per field), then
  // 3 rows x XP loads per field of the z halos
  static constexpr int XP = TX / 64;
  static constexpr int NM = 3 + 9 * XP;
  static constexpr int NWV = TX * TZ / 64;              // wavefronts per block
  static constexpr int NS = (NM + NWV - 1) / NWV;       // halo loads per wavefront
};

// Halo wavefront load m (wave-uniform) and lane e: field slot f (0..2
// among the three fields of that halo side), position (hx, hz) and whether
// the lane has a cell.  An x-halo field reads columns -2, -1, TX ("left")
// or -1, TX, TX+1; a z-halo field rows -2, -1, TZ ("low") or -1, TZ, TZ+1.
// xfirst_left: the first x-halo field is "left" and the other two are not
// (or the reverse); zfirst_low: the first two z-halo fields are "low" and
// the third is not (or the reverse).
template <int TX, int TZ>
__device__ inline bool halo_cell(int m, int e, bool xfirst_left, bool zfirst_low, int &f, int &hx, int &hz) {
  using T = Tile<TX, TZ>;
  if (m < 3) {
    // x halo of field slot m: 3 columns x TZ rows
    f = m;
    const int r = e / 3, ci = e - 3 * r;
    hz = r;
    const bool left = (m == 0) == xfirst_left;
    hx = left ? (ci == 2 ? TX : ci - 2) : (ci == 0 ? -1 : TX + ci - 1);
    return e < 3 * TZ;
  }
  const int mz = m - 3, fz = mz / (3 * T::XP), w = mz - fz * 3 * T::XP, rr = w / T::XP, part = w - rr * T::XP;
  f = fz;
  hx = part * 64 + e;
  const bool low = (fz < 2) == zfirst_low;
  hz = low ? (rr == 2 ? TZ : rr - 2) : (rr == 0 ? -1 : TZ + rr - 1);
  return true;
}

// ---------------------------------------------------------------------------
// Velocity.  Register windows along y: xy (j-2..j+1), yy (j-1..j+2),
// yz (j-2..j+1).  LDS x-z plane at row j: xx (x halo), xy (x), xz (x, z),
// yz (z), zz (z).
// ---------------------------------------------------------------------------
template <int TX, int TZ>
__device__ inline void vel_tiled_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, float *lds,
                                      int bx, int bz,
                                      int by, int tx, int tz, int NY) {
  using T = Tile<TX, TZ>;
  const int t = tz * TX + tx;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
  const int i0 = g.nabc + 1 + bx * TX, kk0 = g.nabc + 1 + bz * TZ;
  const int i = i0 + tx, k = kk0 + tz;
  const int j0 = g.nabc + 1 + by * NY;
  const int je = min(j0 + NY - 1, g.nyt - 2);
  const bool active = i <= g.nxt - g.nabc && k <= g.nzt - g.nfs;
  const bool inside = i <= g.nxt && k <= g.nzt;
  const int sy = g.sy;
  const float c1 = c.c1, c2 = c.c2;
  const float *xx = fo.xx, *yy = fo.yy, *zz = fo.zz, *xy = fo.xy, *yz = fo.yz, *xz = fo.xz;
  float xym2 = 0, xym1 = 0, xy0 = 0, xyp1 = 0, yym1 = 0, yy0 = 0, yyp1 = 0, yyp2 = 0;
  float yzm2 = 0, yzm1 = 0, yz0 = 0, yzp1 = 0;
  // the medium is depth-layered: b = dt/(dh*rho) is one value per z row
  const float b = c.dth / fo.d[active ? k : 0];
  const int qc = inside ? I3(g, i, 1, k) : 0;   // the lane's cell at row 1; row j adds (j-1)*sy
  if (inside) {
    const int q = qc + (j0 - 1) * sy;
    xym2 = ld32(xy, q - 2 * sy); xym1 = ld32(xy, q - sy); xy0 = ld32(xy, q); xyp1 = ld32(xy, q + sy);
    yym1 = ld32(yy, q - sy); yy0 = ld32(yy, q); yyp1 = ld32(yy, q + sy); yyp2 = ld32(yy, q + 2 * sy);
    yzm2 = ld32(yz, q - 2 * sy); yzm1 = ld32(yz, q - sy); yz0 = ld32(yz, q); yzp1 = ld32(yz, q + sy);
  }
  // halo loads: x halos of xx (columns -2, -1, TX), xy and xz (-1, TX,
  // TX+1); z halos of xz, yz (rows -2, -1, TZ) and zz (-1, TZ, TZ+1)
  int hoff[T::NS], hlds[T::NS];
#pragma unroll
  for (int n = 0; n < T::NS; ++n) {
    const int m = wv + n * T::NWV;
    hoff[n] = -1;
    hlds[n] = -1;
    int f, hx, hz;
    if (m < T::NM && halo_cell<TX, TZ>(m, lane, true, true, f, hx, hz)) {
      const int fld = m < 3 ? f : f + 2;
      hlds[n] = fld * T::N + T::at(hx, hz);
      const int gi = i0 + hx, gk = kk0 + hz;
      if (gi >= 1 && gi <= g.nxt && gk >= 1 && gk <= g.nzt) hoff[n] = I3(g, gi, 1, gk);
    }
  }
  int buf = 0;
  for (int j = j0; j <= je; ++j) {
    const int jo = (j - 1) * sy;
    float(*L)[T::N] = reinterpret_cast<float(*)[T::N]>(lds + buf * 5 * T::N);
    const int a0 = T::at(tx, tz);
    float xx0 = 0, xz0 = 0, zz0 = 0;
    const int q = qc + jo;
    if (inside) { xx0 = ld32(xx, q); xz0 = ld32(xz, q); zz0 = ld32(zz, q); }
    // every load of this row is issued before the stores below: the halos,
    // the next rows of the y windows and the cell's own old velocities
    float hv[T::NS];
#pragma unroll
    for (int n = 0; n < T::NS; ++n) {
      const int m = wv + n * T::NWV;
      hv[n] = 0.0f;
      if (m < T::NM) {
        const int fld = m < 3 ? m : 2 + (m - 3) / (3 * T::XP);   // wave-uniform
        const float *s = fld == 0 ? xx : fld == 1 ? xy : fld == 2 ? xz : fld == 3 ? yz : zz;
        if (hoff[n] >= 0) hv[n] = ld32(s, hoff[n] + jo);
      }
    }
    const bool adv = j < je && inside;
    float nxy = 0, nyy = 0, nyz = 0;
    if (adv) { nxy = ld32(xy, q + 2 * sy); nyy = ld32_nt(yy, q + 3 * sy); nyz = ld32(yz, q + 2 * sy); }
    float uo = 0, vo = 0, wo = 0;
    if (active) { uo = ld32_nt(fo.u, q); vo = ld32_nt(fo.v, q); wo = ld32_nt(fo.w, q); }
    L[0][a0] = xx0; L[1][a0] = xy0; L[2][a0] = xz0; L[3][a0] = yz0; L[4][a0] = zz0;
#pragma unroll
    for (int n = 0; n < T::NS; ++n)
      if (hlds[n] >= 0) L[0][hlds[n]] = hv[n];
    __syncthreads();
    if (active) {
      const float *Lxx = L[0], *Lxy = L[1], *Lxz = L[2], *Lyz = L[3], *Lzz = L[4];
      const int W = T::W;
      st32_nt(fn.u, q, uo + b * (c1 * (xx0 - Lxx[a0 - 1]) + c2 * (Lxx[a0 + 1] - Lxx[a0 - 2]) +
                                 c1 * (xy0 - xym1) + c2 * (xyp1 - xym2) +
                                 c1 * (xz0 - Lxz[a0 - W]) + c2 * (Lxz[a0 + W] - Lxz[a0 - 2 * W])));
      st32_nt(fn.v, q, vo + b * (c1 * (Lxy[a0 + 1] - xy0) + c2 * (Lxy[a0 + 2] - Lxy[a0 - 1]) +
                                 c1 * (yyp1 - yy0) + c2 * (yyp2 - yym1) +
                                 c1 * (yz0 - Lyz[a0 - W]) + c2 * (Lyz[a0 + W] - Lyz[a0 - 2 * W])));
      st32_nt(fn.w, q, wo + b * (c1 * (Lxz[a0 + 1] - xz0) + c2 * (Lxz[a0 + 2] - Lxz[a0 - 1]) +
                                 c1 * (yz0 - yzm1) + c2 * (yzp1 - yzm2) +
                                 c1 * (Lzz[a0 + W] - zz0) + c2 * (Lzz[a0 + 2 * W] - Lzz[a0 - W])));
    }
    if (adv) {
      xym2 = xym1; xym1 = xy0; xy0 = xyp1; xyp1 = nxy;
      yym1 = yy0; yy0 = yyp1; yyp1 = yyp2; yyp2 = nyy;
      yzm2 = yzm1; yzm1 = yz0; yz0 = yzp1; yzp1 = nyz;
    }
    buf ^= 1;
  }
}

// ---------------------------------------------------------------------------
// Stress.  Register windows along y: v (j-2..j+1), u (j-1..j+2),
// w (j-1..j+2).  LDS x-z plane at row j: u, v, w.  The medium is
// depth-layered, so lam and mu of a lane's row k (and mu of row k+1, for
// the xz and yz averages) are constants of the march.
// ---------------------------------------------------------------------------
template <int TX, int TZ>
__device__ inline void str_tiled_body(const Dims &g, const Coef &c, const Fields &fo, const Fields &fn, float *lds,
                                      int bx, int bz,
                                      int by, int tx, int tz, int NY) {
  using T = Tile<TX, TZ>;
  const int t = tz * TX + tx;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
  const int i0 = g.nabc + 1 + bx * TX, kk0 = g.nabc + 1 + bz * TZ;
  const int i = i0 + tx, k = kk0 + tz;
  const int j0 = g.nabc + 1 + by * NY;
  const int je = min(j0 + NY - 1, g.nyt - 2);
  const bool active = i <= g.nxt - g.nabc && k <= g.nzt - g.nfs;
  const bool inside = i <= g.nxt && k <= g.nzt;
  const int sy = g.sy;
  const float c1 = c.c1, c2 = c.c2, dth = c.dth;
  const float *u = fn.u, *v = fn.v, *w = fn.w;
  // the reference's factors, formed once: (dth*(lam+2mu)), (dth*lam) and
  // dth * 0.5*(mu + mu') for xy (mu' = mu(j+1) = mu), xz (mu(k+1)) and yz
  // (mu(i+1,j+1,k+1) = mu(k+1))
  const int km = active ? k : 0;
  const float xl = fo.lam[km], xm = fo.mu[km], mz = fo.mu[km + 1];
  const float da = dth * (xl + 2.0f * xm), db = dth * xl;
  const float dxy = dth * (0.5f * (xm + xm)), dxz = dth * (0.5f * (xm + mz));
  float vm2 = 0, vm1 = 0, v0 = 0, vp1 = 0, um1 = 0, u0 = 0, up1 = 0, up2 = 0;
  float wm1 = 0, w0 = 0, wp1 = 0, wp2 = 0;
  const int qc = inside ? I3(g, i, 1, k) : 0;
  if (inside) {
    const int q = qc + (j0 - 1) * sy;
    vm2 = ld32(v, q - 2 * sy); vm1 = ld32(v, q - sy); v0 = ld32(v, q); vp1 = ld32(v, q + sy);
    um1 = ld32(u, q - sy); u0 = ld32(u, q); up1 = ld32(u, q + sy); up2 = ld32(u, q + 2 * sy);
    wm1 = ld32(w, q - sy); w0 = ld32(w, q); wp1 = ld32(w, q + sy); wp2 = ld32(w, q + 2 * sy);
  }
  // halo loads: x halos of u (columns -1, TX, TX+1), v and w (-2, -1, TX);
  // z halos of u, v (rows -1, TZ, TZ+1) and w (-2, -1, TZ)
  int hoff[T::NS], hlds[T::NS];
#pragma unroll
  for (int n = 0; n < T::NS; ++n) {
    const int m = wv + n * T::NWV;
    hoff[n] = -1;
    hlds[n] = -1;
    int f, hx, hz;
    if (m < T::NM && halo_cell<TX, TZ>(m, lane, false, false, f, hx, hz)) {
      hlds[n] = f * T::N + T::at(hx, hz);
      const int gi = i0 + hx, gk = kk0 + hz;
      if (gi >= 1 && gi <= g.nxt && gk >= 1 && gk <= g.nzt) hoff[n] = I3(g, gi, 1, gk);
    }
  }
  int buf = 0;
  for (int j = j0; j <= je; ++j) {
    const int jo = (j - 1) * sy;
    float(*L)[T::N] = reinterpret_cast<float(*)[T::N]>(lds + buf * 3 * T::N);
    const int a0 = T::at(tx, tz);
    const int q = qc + jo;
    // every load of this row is issued before the stores below: the halos,
    // the next rows of the y windows and the cell's old stresses
    float hv[T::NS];
#pragma unroll
    for (int n = 0; n < T::NS; ++n) {
      const int m = wv + n * T::NWV;
      hv[n] = 0.0f;
      if (m < T::NM) {
        const int fld = m < 3 ? m : (m - 3) / (3 * T::XP);   // wave-uniform
        const float *s = fld == 0 ? u : fld == 1 ? v : w;
        if (hoff[n] >= 0) hv[n] = ld32(s, hoff[n] + jo);
      }
    }
    const bool adv = j < je && inside;
    float nv = 0, nu = 0, nw = 0;
    if (adv) { nv = ld32(v, q + 2 * sy); nu = ld32(u, q + 3 * sy); nw = ld32(w, q + 3 * sy); }
    float oxx = 0, oyy = 0, ozz = 0, oxy = 0, oxz = 0, oyz = 0;
    if (active) {
      oxx = ld32_nt(fo.xx, q); oyy = ld32_nt(fo.yy, q); ozz = ld32_nt(fo.zz, q);
      oxy = ld32_nt(fo.xy, q); oxz = ld32_nt(fo.xz, q); oyz = ld32_nt(fo.yz, q);
    }
    L[0][a0] = u0; L[1][a0] = v0; L[2][a0] = w0;
#pragma unroll
    for (int n = 0; n < T::NS; ++n)
      if (hlds[n] >= 0) L[0][hlds[n]] = hv[n];
    __syncthreads();
    if (active) {
      const float *Lu = L[0], *Lv = L[1], *Lw = L[2];
      const int W = T::W;
      const float dux1 = c1 * (Lu[a0 + 1] - u0), dux2 = c2 * (Lu[a0 + 2] - Lu[a0 - 1]);
      const float dvy1 = c1 * (v0 - vm1), dvy2 = c2 * (vp1 - vm2);
      const float dwz1 = c1 * (w0 - Lw[a0 - W]), dwz2 = c2 * (Lw[a0 + W] - Lw[a0 - 2 * W]);
      st32_nt(fn.xx, q, (oxx + da * (dux1 + dux2)) + db * (((dvy1 + dvy2) + dwz1) + dwz2));
      st32_nt(fn.yy, q, (oyy + da * (dvy1 + dvy2)) + db * (((dux1 + dux2) + dwz1) + dwz2));
      st32_nt(fn.zz, q, (ozz + da * (dwz1 + dwz2)) + db * (((dux1 + dux2) + dvy1) + dvy2));
      st32_nt(fn.xy, q, oxy + dxy * (((c1 * (up1 - u0) + c2 * (up2 - um1)) + c1 * (v0 - Lv[a0 - 1])) +
                                  c2 * (Lv[a0 + 1] - Lv[a0 - 2])));
      st32_nt(fn.xz, q, oxz + dxz * (((c1 * (Lu[a0 + W] - u0) + c2 * (Lu[a0 + 2 * W] - Lu[a0 - W])) +
                                   c1 * (w0 - Lw[a0 - 1])) + c2 * (Lw[a0 + 1] - Lw[a0 - 2])));
      st32_nt(fn.yz, q, oyz + dxz * (((c1 * (Lv[a0 + W] - v0) + c2 * (Lv[a0 + 2 * W] - Lv[a0 - W])) +
                                   c1 * (wp1 - w0)) + c2 * (wp2 - wm1)));
    }
    if (adv) {
      vm2 = vm1; vm1 = v0; v0 = vp1; vp1 = nv;
      um1 = u0; u0 = up1; up1 = up2; up2 = nu;
      wm1 = w0; w0 = wp1; wp1 = wp2; wp2 = nw;
    }
    buf ^= 1;
  }
}

}  // namespace fd3d
