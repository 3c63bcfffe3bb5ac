// Fused interior sweep: the velocity update (dvel, fd3d_deriv.f90:24-193)
// and the stress update (dstres, :513-622) of the interior cells in one
// pass over HBM.
//
// Separately, a step reads the six stresses twice (once for the
// velocities, once to update them) and the velocities twice (update, then
// stress stencils): 30 words per cell.  Fused, each of sigma_old, v_old,
// rho, lam, mu is read once and v_new, sigma_new written once: 21 words.
// The two wavefield sets make this race-free: every block reads the old
// set and writes the new one.
//
// Work of a block: a 64 x FZC tile of the x-z plane (the core) marching
// along y over the rows of its chunk.  Each row r:
//   1. the old stresses of row r are staged in LDS over the core plus a
//      4-cell ring (72 x FZC+8);
//   2. new velocities of row r are computed over the core plus a 2-cell
//      ring (68 x FZC+4), lanes on the ring recomputing what neighbouring
//      tiles own; positions outside the interior (PML, near-fault rows,
//      outer faces) take the values pml_vel / vel_fault already wrote;
//      the plane goes into a 3-row LDS ring, core values also to HBM;
//   3. new stresses of row r-2 are computed on the core from the LDS ring
//      (x, z derivatives) and from per-lane register windows of the core
//      column (y derivatives: v of rows r-4..r-1, u and w of r-3..r).
// The march starts 2 rows before the chunk and ends 2 rows after it, so
// each chunk recomputes 4 rows of velocities it does not own.
//
// Per-point arithmetic is that of tiles.h (the same operands, the same
// association), so the result is bit-identical to the separate sweeps.
// Full-space problems only (with a free surface the ghost velocities of
// fuvw must be formed between the two updates).
#include "stages.h"
#include "tiles.h"

namespace fd3d {
namespace {

constexpr int FZE = 16;            // z rows of lanes (core + 2-cell ring)
constexpr int FZC = FZE - 4;       // core z rows
constexpr int SW = 64 + 8;         // stress plane: x in [-4, 67] relative to the tile
constexpr int SH = FZC + 8;        //               z in [-4, FZC+3]
constexpr int VW = 64 + 4;         // velocity plane: x in [-2, 65]
constexpr int VH = FZE;            //                 z in [-2, FZC+1]
constexpr int NSP = SW * SH;       // 1440
constexpr int NVP = VW * VH;       // 1088

__device__ inline int sat(int x, int z) { return (z + 4) * SW + (x + 4); }
__device__ inline int vat(int x, int z) { return (z + 2) * VW + (x + 2); }

struct FusedLds {
  float s[5][NSP];        // xx, xy, xz, yz, zz of row r
  float v[3][3][NVP];     // ring of 3 rows x (u, v, w)
};

// Velocity of one position (x, z relative to the tile) at row r: the
// interior update where the position is an interior cell, else the value
// already in the new set (0 outside the array).
struct VelPos {
  int i, k;                   // global indices
  bool inside, interior_xz;   // in the array / interior in x and z
  // y windows of the old stresses at this position
  float xym2, xym1, xy0, xyp1, yym1, yy0, yyp1, yyp2, yzm2, yzm1, yz0, yzp1;
};

__device__ inline void vp_init(VelPos &P, const Dims &g, int i, int k) {
  P.i = i; P.k = k;
  P.inside = i >= 1 && i <= g.nxt && k >= 1 && k <= g.nzt;
  P.interior_xz = i >= g.nabc + 1 && i <= g.nxt - g.nabc && k >= g.nabc + 1 && k <= g.nzt - g.nfs;
  P.xym2 = P.xym1 = P.xy0 = P.xyp1 = 0.0f;
  P.yym1 = P.yy0 = P.yyp1 = P.yyp2 = 0.0f;
  P.yzm2 = P.yzm1 = P.yz0 = P.yzp1 = 0.0f;
}

// Old stress of field f at row j (0 outside the array).
__device__ inline float ld_row(const float *f, const Dims &g, const VelPos &P, int j) {
  return (P.inside && j >= 1 && j <= g.nyt) ? f[I3(g, P.i, j, P.k)] : 0.0f;
}

// Fill the y windows for row r: xy rows r-2..r+1, yy rows r-1..r+2, yz rows r-2..r+1.
__device__ inline void vp_fill(VelPos &P, const Dims &g, const Fields &fo, int r) {
  P.xym2 = ld_row(fo.xy, g, P, r - 2); P.xym1 = ld_row(fo.xy, g, P, r - 1);
  P.xy0 = ld_row(fo.xy, g, P, r); P.xyp1 = ld_row(fo.xy, g, P, r + 1);
  P.yym1 = ld_row(fo.yy, g, P, r - 1); P.yy0 = ld_row(fo.yy, g, P, r);
  P.yyp1 = ld_row(fo.yy, g, P, r + 1); P.yyp2 = ld_row(fo.yy, g, P, r + 2);
  P.yzm2 = ld_row(fo.yz, g, P, r - 2); P.yzm1 = ld_row(fo.yz, g, P, r - 1);
  P.yz0 = ld_row(fo.yz, g, P, r); P.yzp1 = ld_row(fo.yz, g, P, r + 1);
}

// Advance the windows from row r to row r+1 (nxy, nyy, nyz: xy(r+2), yy(r+3), yz(r+2)).
__device__ inline void vp_shift(VelPos &P, float nxy, float nyy, float nyz) {
  P.xym2 = P.xym1; P.xym1 = P.xy0; P.xy0 = P.xyp1; P.xyp1 = nxy;
  P.yym1 = P.yy0; P.yy0 = P.yyp1; P.yyp1 = P.yyp2; P.yyp2 = nyy;
  P.yzm2 = P.yzm1; P.yzm1 = P.yz0; P.yz0 = P.yzp1; P.yzp1 = nyz;
}

// New (u, v, w) at position (x, z) of the tile, row r.  S is the LDS stress
// plane of row r.  Arithmetic as vel_tiled_body.
__device__ inline void vp_update(const VelPos &P, const Dims &g, const Coef &c, const Fields &fo, const Fields &fn,
                                 const float (*S)[NSP], int x, int z, int r, float &un, float &vn, float &wn) {
  un = vn = wn = 0.0f;
  if (!P.inside || r < 1 || r > g.nyt) return;
  const int q = I3(g, P.i, r, P.k);
  if (!(P.interior_xz && r >= g.nabc + 1 && r <= g.nyt - 2)) {
    un = fn.u[q]; vn = fn.v[q]; wn = fn.w[q];
    return;
  }
  const float uo = fo.u[q], vo = fo.v[q], wo = fo.w[q], dq = fo.d[P.k];
  const float c1 = c.c1, c2 = c.c2;
  const float *Lxx = S[0], *Lxy = S[1], *Lxz = S[2], *Lyz = S[3], *Lzz = S[4];
  const int a0 = sat(x, z);
  const float xx0 = Lxx[a0], xz0 = Lxz[a0], zz0 = Lzz[a0];
  const float xy0 = P.xy0, yz0 = P.yz0;
  const float b = c.dth / dq;
  un = uo + b * (c1 * (xx0 - Lxx[a0 - 1]) + c2 * (Lxx[a0 + 1] - Lxx[a0 - 2]) +
                 c1 * (xy0 - P.xym1) + c2 * (P.xyp1 - P.xym2) +
                 c1 * (xz0 - Lxz[a0 - SW]) + c2 * (Lxz[a0 + SW] - Lxz[a0 - 2 * SW]));
  vn = vo + b * (c1 * (Lxy[a0 + 1] - xy0) + c2 * (Lxy[a0 + 2] - Lxy[a0 - 1]) +
                 c1 * (P.yyp1 - P.yy0) + c2 * (P.yyp2 - P.yym1) +
                 c1 * (yz0 - Lyz[a0 - SW]) + c2 * (Lyz[a0 + SW] - Lyz[a0 - 2 * SW]));
  wn = wo + b * (c1 * (Lxz[a0 + 1] - xz0) + c2 * (Lxz[a0 + 2] - Lxz[a0 - 1]) +
                 c1 * (yz0 - P.yzm1) + c2 * (P.yzp1 - P.yzm2) +
                 c1 * (Lzz[a0 + SW] - zz0) + c2 * (Lzz[a0 + 2 * SW] - Lzz[a0 - SW]));
}

__global__ void __launch_bounds__(64 * FZE) k_vs_fused(Dims g, Coef c, Fields f0, Fields f1, const Ctl *ctl,
                                                       int gx, int gz, int gy, int NY) {
  __shared__ FusedLds L;
  const FieldPair FP = field_pair(f0, f1, ctl);
  const Fields &fo = FP.fo, &fn = FP.fn;
  const int lid = xcd_remap(blockIdx.x, gx * gz * gy);
  const int bx = lid % gx, bz = (lid / gx) % gz, by = lid / (gx * gz);
  const int tx = threadIdx.x, ty = threadIdx.y, t = ty * 64 + tx;
  const int i0 = g.nabc + 1 + bx * 64, k0 = g.nabc + 1 + bz * FZC;
  const int j0 = g.nabc + 1 + by * NY;
  const int je = min(j0 + NY - 1, g.nyt - 2);
  const float c1 = c.c1, c2 = c.c2, dth = c.dth;
  const int sy = g.sy, sz = g.sz;

  // lane position in the velocity plane: x = tx, z = ty - 2
  const int xz_z = ty - 2;
  VelPos P;
  vp_init(P, g, i0 + tx, k0 + xz_z);
  const bool core = xz_z >= 0 && xz_z < FZC;
  const bool own = core && P.interior_xz;            // this lane's cell gets new stresses
  // wave 0 also covers the x ring: columns -2, -1, 64, 65, all 16 z rows
  const bool ring = ty == 0;
  const int hx = tx < 16 ? -2 : tx < 32 ? -1 : tx < 48 ? 64 : 65;
  const int hz = (tx & 15) - 2;
  VelPos H;
  vp_init(H, g, i0 + hx, k0 + hz);

  const int r0 = j0 - 2, r1 = je + 2;
  vp_fill(P, g, fo, r0);
  if (ring) vp_fill(H, g, fo, r0);
  // register windows of the new velocities of the lane's own column
  float vwm4 = 0, vwm3 = 0, vwm2 = 0, vwm1 = 0;   // v rows r-4..r-1
  float uwm3 = 0, uwm2 = 0, uwm1 = 0;             // u rows r-3..r-1 (u(r) is new this row)
  float wwm3 = 0, wwm2 = 0, wwm1 = 0;             // w rows r-3..r-1

  for (int r = r0; r <= r1; ++r) {
    // ---- 1. old stresses of row r over the 72 x (FZC+8) plane
    __syncthreads();   // the previous row's readers of L.s are done
    for (int h = t; h < 5 * NSP; h += 64 * FZE) {
      const int fld = h / NSP, e = h - fld * NSP;
      const int x = e % SW - 4, z = e / SW - 4;
      const int gi = i0 + x, gk = k0 + z;
      const float *src = fld == 0 ? fo.xx : fld == 1 ? fo.xy : fld == 2 ? fo.xz : fld == 3 ? fo.yz : fo.zz;
      float val = 0.0f;
      if (gi >= 1 && gi <= g.nxt && gk >= 1 && gk <= g.nzt && r >= 1 && r <= g.nyt)
        val = src[I3(g, gi, r, gk)];
      L.s[fld][e] = val;
    }
    // next rows of the y windows, issued with the plane loads
    const float pxy = ld_row(fo.xy, g, P, r + 2), pyy = ld_row(fo.yy, g, P, r + 3),
                pyz = ld_row(fo.yz, g, P, r + 2);
    float hxy = 0, hyy = 0, hyz = 0;
    if (ring) { hxy = ld_row(fo.xy, g, H, r + 2); hyy = ld_row(fo.yy, g, H, r + 3); hyz = ld_row(fo.yz, g, H, r + 2); }
    __syncthreads();

    // ---- 2. new velocities of row r over the 68 x FZE plane
    const int slot = (r - r0) % 3;
    float un, vn, wn;
    vp_update(P, g, c, fo, fn, L.s, tx, xz_z, r, un, vn, wn);
    L.v[slot][0][vat(tx, xz_z)] = un;
    L.v[slot][1][vat(tx, xz_z)] = vn;
    L.v[slot][2][vat(tx, xz_z)] = wn;
    if (own && r >= j0 && r <= je) {
      const int q = I3(g, P.i, r, P.k);
      fn.u[q] = un; fn.v[q] = vn; fn.w[q] = wn;
    }
    if (ring) {
      float hu, hv, hw;
      vp_update(H, g, c, fo, fn, L.s, hx, hz, r, hu, hv, hw);
      L.v[slot][0][vat(hx, hz)] = hu;
      L.v[slot][1][vat(hx, hz)] = hv;
      L.v[slot][2][vat(hx, hz)] = hw;
    }
    __syncthreads();

    // ---- 3. new stresses of row s = r - 2 on the core
    const int s = r - 2;
    if (own && s >= j0 && s <= je) {
      const int ps = (s - r0) % 3;
      const float *Lu = L.v[ps][0], *Lv = L.v[ps][1], *Lw = L.v[ps][2];
      const int a0 = vat(tx, xz_z);
      const int q = I3(g, P.i, s, P.k);
      const float oxx = fo.xx[q], oyy = fo.yy[q], ozz = fo.zz[q], oxy = fo.xy[q], oxz = fo.xz[q], oyz = fo.yz[q];
      // depth-layered medium: mu(j+1) = mu, mu(k+1) = mu(i+1, j+1, k+1)
      const float xl = fo.lam[P.k], xm = fo.mu[P.k], mp1 = xm, mz1 = fo.mu[P.k + 1], mdiag = mz1;
      // own column: v(s-2..s+1) = vwm2.. (rows r-4..r-1); u, w (s-1..s+2) = rows r-3..r
      const float u0 = Lu[a0], v0 = Lv[a0], w0 = Lw[a0];
      const float vm2 = vwm4, vm1 = vwm3, vp1 = vwm1;
      const float um1 = uwm3, up1 = uwm1, up2 = un;
      const float wm1 = wwm3, wp1 = wwm1, wp2 = wn;
      const float a = xl + 2.0f * xm, b = xl;
      const int W = VW;
      const float dux1 = c1 * (Lu[a0 + 1] - u0), dux2 = c2 * (Lu[a0 + 2] - Lu[a0 - 1]);
      const float dvy1 = c1 * (v0 - vm1), dvy2 = c2 * (vp1 - vm2);
      const float dwz1 = c1 * (w0 - Lw[a0 - W]), dwz2 = c2 * (Lw[a0 + W] - Lw[a0 - 2 * W]);
      fn.xx[q] = (oxx + (dth * a) * (dux1 + dux2)) + (dth * b) * (((dvy1 + dvy2) + dwz1) + dwz2);
      fn.yy[q] = (oyy + (dth * a) * (dvy1 + dvy2)) + (dth * b) * (((dux1 + dux2) + dwz1) + dwz2);
      fn.zz[q] = (ozz + (dth * a) * (dwz1 + dwz2)) + (dth * b) * (((dux1 + dux2) + dvy1) + dvy2);
      float xmu = 0.5f * (xm + mp1);
      fn.xy[q] = oxy + (dth * xmu) * (((c1 * (up1 - u0) + c2 * (up2 - um1)) + c1 * (v0 - Lv[a0 - 1])) +
                                      c2 * (Lv[a0 + 1] - Lv[a0 - 2]));
      xmu = 0.5f * (xm + mz1);
      fn.xz[q] = oxz + (dth * xmu) * (((c1 * (Lu[a0 + W] - u0) + c2 * (Lu[a0 + 2 * W] - Lu[a0 - W])) +
                                       c1 * (w0 - Lw[a0 - 1])) + c2 * (Lw[a0 + 1] - Lw[a0 - 2]));
      xmu = 0.5f * (xm + mdiag);
      fn.yz[q] = oyz + (dth * xmu) * (((c1 * (Lv[a0 + W] - v0) + c2 * (Lv[a0 + 2 * W] - Lv[a0 - W])) +
                                       c1 * (wp1 - w0)) + c2 * (wp2 - wm1));
    }
    // ---- advance the windows
    vwm4 = vwm3; vwm3 = vwm2; vwm2 = vwm1; vwm1 = vn;
    uwm3 = uwm2; uwm2 = uwm1; uwm1 = un;
    wwm3 = wwm2; wwm2 = wwm1; wwm1 = wn;
    vp_shift(P, pxy, pyy, pyz);
    if (ring) vp_shift(H, hxy, hyy, hyz);
  }
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

hipError_t launch_fused(const Problem &p, hipStream_t s) {
  const Dims &g = p.g;
  if (g.nyt - 2 - g.nabc <= 0) return hipSuccess;
  const int ny = p.tune.ny_vel;
  if (ny < 1) return hipErrorInvalidValue;
  const int gx = cdiv(g.nxt - 2 * g.nabc, 64), gz = cdiv(g.nzt - g.nfs - g.nabc, FZC);
  const int gy = cdiv(g.nyt - 2 - g.nabc, ny);
  hipLaunchKernelGGL(k_vs_fused, dim3(gx * gz * gy), dim3(64, FZE), 0, s, g, p.c, p.f[0], p.f[1], p.ctl, gx, gz, gy,
                     ny);
  return hipGetLastError();
}

}  // namespace fd3d
