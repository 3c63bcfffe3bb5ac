// Fused interior sweep: the velocity update (dvel, fd3d_deriv.f90:24-193)
// and the stress update (dstres, :513-622) of the interior cells in one
// pass over HBM.  Device body only: fused.hip launches it for one problem,
// batch.hip as a stage of a batch.
//
// Separately, the two sweeps move 27 words per interior cell (velocity: 6
// old stresses and 3 old velocities read, 3 new velocities written;
// stress: the 3 new velocities and 6 old stresses read, 6 written).  Fused,
// the old stresses and velocities are read once and the new values written
// once: 18 words.  The two wavefield sets make this race-free: every block
// reads the old set and writes the new one.
//
// Work of a block (16 wavefronts of 64 lanes).  The core, the cells that
// get new stresses, is x in [0, 63], z in [0, 10] of an x-z tile: its x
// extent is one 256-byte row segment of every array, so every load and
// store of a core row moves whole cache lines.  Wavefront w < 15 owns the
// positions x = lane, z = w - 2 (the core rows and a 2-row ring above and
// below); wavefront 15 owns the x-ring, x in {-2, -1, 64, 65} for every
// z in [-2, 12].  The block marches along y over the rows [j0, je] of its
// chunk, from j0 - 2 to je + 2.  Iteration r:
//   1. the old-stress planes of row r (loaded into registers during
//      iteration r - 2) go into one of two LDS buffers, and every lane puts
//      its new velocity of row r - 2 (from its register window) into one of
//      two velocity planes; the loads of row r + 2 are issued (they are in
//      flight across the barrier); one barrier;
//   2. every lane forms the new velocity of its position at row r from the
//      stress planes (x and z neighbours) and its register windows of xy,
//      yy, yz (y neighbours); core lanes form the new stresses of row r - 2
//      from the velocity plane (x and z neighbours) and their velocity
//      windows (y neighbours).
// Lanes on the ring around the core recompute velocities that the
// neighbouring tiles own, so no block waits on another.  Positions outside
// the interior (PML cells, the near-fault rows j >= nyt-1, the outer faces)
// take the values pml_vel and vel_fault already wrote into the new set: the
// kernel runs after them and before str_fault and pml_str, which read its
// new velocities.
//
// Per-point arithmetic is that of tiles.h (the same operands, the same
// association), so the result is bit-identical to the separate sweeps.
// Under a free surface the stresses of the top two interior planes need
// the ghost velocities that fuvw forms from the new velocities: the sweep
// stores stresses up to kmax_s = ks - 2 only (ks = nzt - nfs), and
// str_top does the last two planes after fuvw.
#pragma once
#include "stages.h"

namespace fd3d {
namespace fused {


constexpr int LX = 64;             // core x extent: one wavefront, one 256-byte row segment
constexpr int NWV = 16;            // wavefronts per block
constexpr int ZC = NWV - 5;        // core rows along z (11)
constexpr int NZP = ZC + 4;        // position rows, z in [-2, ZC + 1]
constexpr int NT = LX * NWV;
constexpr int PW = 72;             // LDS plane row: x in [-4, 67]
constexpr int PX0 = 4;             // its column of x = 0

// Old-stress planes of one row.  Plane f covers the rows z in [z0, z0 +
// rows) that the velocity stencils of the positions reach, at LDS rows
// base .. base + rows - 1 (PW floats each).  The stencils reach these
// columns of each plane:
//   xx: z in [-2, 12], x in [-4, 66]      xy: z in [-2, 12], x in [-3, 67]
//   xz: z in [-4, 13], x in [-3, 67]      yz: z in [-4, 13], x in [-2, 65]
//   zz: z in [-3, 14], x in [-2, 65]
// A plane row is 18 aligned float4 chunks (x = -4 + 4q, q = 0..17; the
// core origin starts a 256-byte boundary), loaded one chunk per lane with
// 16-byte loads, field by field: a field of R rows takes ceil(18 R / 64)
// wavefront loads.
constexpr int PROWS_XY = NZP, PROWS_Z = ZC + 7;
constexpr int PBASE_XY = NZP, PBASE_XZ = 2 * NZP, PBASE_YZ = PBASE_XZ + PROWS_Z, PBASE_ZZ = PBASE_YZ + PROWS_Z;
constexpr int NPR = PBASE_ZZ + PROWS_Z;        // plane rows (84)
constexpr int PCH = PW / 4;                     // chunks per plane row
constexpr int NLXY = (PROWS_XY * PCH + 63) / 64, NLZ = (PROWS_Z * PCH + 63) / 64;
constexpr int M_XY = NLXY, M_XZ = 2 * NLXY, M_YZ = M_XZ + NLZ, M_ZZ = M_YZ + NLZ;
constexpr int NWL = M_ZZ + NLZ;                 // wavefront loads per row (28)
constexpr int NPF = (NWL + NWV - 1) / NWV;      // per wavefront (2)
static_assert(4 * NZP <= 64, "the x-ring fits one wavefront");

// Tile order within a y chunk: strips kStrip tiles wide in x, each walked
// in z, so the 32 blocks an XCD runs at once (xcd_remap gives each XCD a
// contiguous range) cover 8 z-tiles of 4 x-tiles instead of 2 z-rows of
// 16 x-tiles.  Measured on the 1024x512x512 sweep (5 processes each): 4.60
// ms against 4.79 ms for the x-fastest order, though the L2 then fetches
// slightly more (1.42 against 1.36 times the algorithmic reads); equal on
// 512x256x256.  In the batch engine's fused stage the step of 64 uneven
// domains (96-320 cells a side) took 25.5 against 26.2 ms and 24.0 against
// 24.5 ms (two pairs on one GPU).  z-fastest order (1 tile wide): 5.93
// ms.  The width of a single problem's strips is a tuning choice (mode
// bits 5-6, 4 by default), since how the tiles divide over the 8 XCDs
// decides it: 1024x512x512 (16 x 47 tiles, 94 per XCD, one strip of 2
// each) runs 4.61 ms in strips of 2 against 4.70 ms in strips of 4 and
// 5.12 ms in strips of 3; 512x256x256 (8 x 24 tiles in 4 y chunks, 96 per
// XCD) 0.656 ms in strips of 4 against 0.716 ms in strips of 2.  The batch
// engine's fused stage keeps strips of 4.
constexpr int kStrip = 4;
// Tile (bx, bz) of number r in [0, gx * gz) in that order, strips `strip`
// tiles wide.
__device__ __forceinline__ void strip_tile(int r, int gx, int gz, int strip, int &bx, int &bz) {
  const int st = r / (strip * gz), w = min(strip, gx - st * strip), q = r - st * strip * gz;
  bx = st * strip + q % w;
  bz = q / w;
}
// Strip width of tuning mode bits 5-6.
__host__ __device__ inline int strip_of_mode(int mode) {
  const int code = (mode >> 5) & 3;
  return code == 0 ? kStrip : code == 1 ? 2 : code == 2 ? 8 : 3;
}

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const f4 gf4;

struct PlaneDef { int z0, rows, base; };
__device__ inline PlaneDef plane_def(int f) {
  switch (f) {
    case 0: return {-2, PROWS_XY, 0};
    case 1: return {-2, PROWS_XY, PBASE_XY};
    case 2: return {-4, PROWS_Z, PBASE_XZ};
    case 3: return {-4, PROWS_Z, PBASE_YZ};
    default: return {-3, PROWS_Z, PBASE_ZZ};
  }
}
// LDS slot of plane f at (x, z)
template <int F>
__device__ inline int at(int x, int z) {
  constexpr int z0 = F == 0 || F == 1 ? -2 : F == 4 ? -3 : -4;
  constexpr int base = F == 0 ? 0 : F == 1 ? PBASE_XY : F == 2 ? PBASE_XZ : F == 3 ? PBASE_YZ : PBASE_ZZ;
  return (base + z - z0) * PW + PX0 + x;
}

// Wavefront load m (wave-uniform) of lane l: its field (0..4 = xx, xy, xz,
// yz, zz), the LDS slot and position (x, z) of its chunk's first cell;
// false if the lane has no chunk.
__device__ inline bool plane_chunk(int m, int l, int &f, int &slot, int &x, int &z) {
  f = m < M_XY ? 0 : m < M_XZ ? 1 : m < M_YZ ? 2 : m < M_ZZ ? 3 : 4;
  const int m0 = f == 0 ? 0 : f == 1 ? M_XY : f == 2 ? M_XZ : f == 3 ? M_YZ : M_ZZ;
  const PlaneDef d = plane_def(f);
  const int e = (m - m0) * 64 + l, row = e / PCH, q = e - row * PCH;
  x = 4 * q - PX0;
  z = d.z0 + row;
  slot = (d.base + row) * PW + 4 * q;
  return row < d.rows;
}

// LDS of one block: old-stress planes of row r and new velocities of row
// r - 2, two buffers each.
struct FusedLds {
  float S[2][NPR * PW] __attribute__((aligned(16)));
  float V[2][3][NZP * PW];
};

// One block's work: the tile (bx, bz) of the x-z tile grid and the y chunk
// by of NY rows.  Lanes are threadIdx.x (0..63) of wavefront threadIdx.y
// (0..15) of a 64 x 16 block.
__device__ __forceinline__ void fused_body(const Dims &g, const Coef &c, const Fields &f0, const Fields &f1,
                                           const Ctl *ctl, int bx, int bz, int by, int NY, int kmax_s,
                                           FusedLds &L) {
  float(*Sb)[NPR * PW] = L.S;
  float(*Vb)[3][NZP * PW] = L.V;
  // the old set (fo) is set it & 1, the new one the other (see field_pair);
  // only the pointers this kernel uses are selected
  const bool odd = ctl->it & 1;
  const int tx = threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.y);   // the wavefront
  // the lane's position; lanes 60..63 of the x-ring wavefront have none
  // (they take position (0, 0) for their reads and write nothing)
  int x = tx, z = wv - 2;
  bool has = true;
  if (wv >= NZP) {
    const int col = tx & 3;
    x = col < 2 ? col - 2 : LX + col - 2;
    z = (tx >> 2) - 2;
    has = tx < 4 * NZP;
    if (!has) x = 0, z = 0;
  }
  const int i0 = g.nabc + 1 + bx * LX, k0 = g.nabc + 1 + bz * ZC;   // core origin
  const int j0 = g.nabc + 1 + by * NY, je = min(j0 + NY - 1, g.nyt - 2);
  const int i = i0 + x, k = k0 + z;
  const bool in_xz = has && i >= 1 && i <= g.nxt && k >= 1 && k <= g.nzt;
  const bool int_xz = has && i >= g.nabc + 1 && i <= g.nxt - g.nabc && k >= g.nabc + 1 && k <= g.nzt - g.nfs;
  const bool core = int_xz && wv >= 2 && wv < 2 + ZC;
  const bool core_s = core && k <= kmax_s;   // stores new stresses
  // In a tile the interior ends in (the last x or z tile of a grid whose
  // interior is not a multiple of the tile), the stencils of the last
  // interior column and plane reach 2 cells past it: stress planes and
  // velocities beyond are never read, so they are not loaded.
  const int xlim = g.nxt - g.nabc - i0 + 2, zlim = g.nzt - g.nfs - k0 + 2;
  const bool need_v = in_xz && x <= xlim && z <= zlim;   // velocity source read
  const int t = (z + 2) * PW + PX0 + x;   // the lane's slot in a velocity plane
  const int sy = g.sy;
  const float c1 = c.c1, c2 = c.c2, dth = c.dth;
  // depth-layered medium: the factors of the lane's plane k (interior lanes)
  const int km = int_xz ? k : g.nabc + 1;
  const float b = dth / f0.d[km];
  const float xl = f0.lam[km], xm = f0.mu[km], mz = f0.mu[km + 1];
  const float da = dth * (xl + 2.0f * xm), db = dth * xl;
  const float dxy = dth * (0.5f * (xm + xm)), dxz = dth * (0.5f * (xm + mz));
  const int qb = in_xz ? I3(g, i, 1, k) : 0;   // (i, 1, k); row j adds (j-1)*sy
  const int r0 = j0 - 2, r1 = je + 2;
  const int jlo = g.nabc + 1, jhi = g.nyt - 2;  // interior rows

#define FUSED_SEL(a) const float *o##a = odd ? f1.a : f0.a; float *n##a = odd ? f0.a : f1.a
  FUSED_SEL(xx); FUSED_SEL(yy); FUSED_SEL(zz); FUSED_SEL(xy); FUSED_SEL(yz); FUSED_SEL(xz); FUSED_SEL(u); FUSED_SEL(v); FUSED_SEL(w);
#undef FUSED_SEL
  float *nu_ = nu, *nv_ = nv, *nw_ = nw;
#define FUSED_COL(f, j) ((in_xz && (j) >= 1 && (j) <= g.nyt) ? ld32(f, qb + ((j) - 1) * sy) : 0.0f)
// velocity source of row j: the old set where this kernel updates the
// position, the new set where another kernel already did, 0 outside
#define FUSED_VSRC(j, a, bb, cc)                                                      \
  do {                                                                              \
    a = bb = cc = 0.0f;                                                             \
    if (need_v && (j) >= 1 && (j) <= g.nyt) {                                       \
      const int q_ = qb + ((j) - 1) * sy;                                           \
      if (int_xz && (j) >= jlo && (j) <= jhi) { a = ld32(ou, q_); bb = ld32(ov, q_); cc = ld32(ow, q_); } \
      else { a = ld32(nu_, q_); bb = ld32(nv_, q_); cc = ld32(nw_, q_); }            \
    }                                                                               \
  } while (0)
// plane chunks of row j into registers: wavefront wv takes wavefront
// loads m = wv, wv + NWV; each is one field (a uniform base pointer).  The
// row-independent part of each chunk's offset is formed once (-1: the
// chunk is outside the arrays and reads as 0), and so is its LDS slot
// (-1: no chunk).  A chunk lies in one padded row when its last cell does.
  int poff[NPF], pslot[NPF];
  const float *psrc[NPF];   // wave-uniform field of each load
#pragma unroll
  for (int n = 0; n < NPF; ++n) {
    const int m = wv + n * NWV;
    poff[n] = -1;
    pslot[n] = -1;
    psrc[n] = oxx;
    if (m < NWL) {
      int px, pz, f, sl;
      if (plane_chunk(m, tx, f, sl, px, pz)) {
        pslot[n] = sl;
        const int gi = i0 + px, gk = k0 + pz;
        if (gk >= 1 && gk <= g.nzt && g.off0 + gi + 3 <= g.px && px <= xlim && pz <= zlim)
          poff[n] = I3(g, gi, 1, gk);
      }
      psrc[n] = f == 0 ? oxx : f == 1 ? oxy : f == 2 ? oxz : f == 3 ? oyz : ozz;
    }
  }
#define FUSED_PLANES(pf, j)                                                            \
  do {                                                                              \
    const bool jok_ = (j) >= 1 && (j) <= g.nyt;                                     \
    const int jo_ = ((j) - 1) * sy;                                                 \
    _Pragma("unroll") for (int n = 0; n < NPF; ++n)                                 \
      pf[n] = (jok_ && poff[n] >= 0) ? *(gf4 *)((const __attribute__((address_space(1))) char *)psrc[n] + ((unsigned)(poff[n] + jo_) << 2)) : f4{0.0f, 0.0f, 0.0f, 0.0f}; \
  } while (0)

  // y windows of the old stresses: xy rows r-2..r, yy rows r-2..r+1, yz
  // rows r-2..r before iteration r (its prefetch set adds xy(r+1),
  // yy(r+2), yz(r+1))
  float xym2 = 0, xym1 = 0, xy0 = 0, yym2 = 0, yym1 = 0, yy0 = 0, yyp1 = 0, yzm2 = 0, yzm1 = 0, yz0 = 0;
  if (int_xz) {
    xym2 = FUSED_COL(oxy, r0 - 2); xym1 = FUSED_COL(oxy, r0 - 1); xy0 = FUSED_COL(oxy, r0);
    yym2 = FUSED_COL(oyy, r0 - 2); yym1 = FUSED_COL(oyy, r0 - 1); yy0 = FUSED_COL(oyy, r0); yyp1 = FUSED_COL(oyy, r0 + 1);
    yzm2 = FUSED_COL(oyz, r0 - 2); yzm1 = FUSED_COL(oyz, r0 - 1); yz0 = FUSED_COL(oyz, r0);
  }
  // new velocities of the lane's column: v rows r-4..r-1 (vA..vD), u and w
  // rows r-3..r-1 (uA..uC, wA..wC)
  float vA = 0, vB = 0, vC = 0, vD = 0, uA = 0, uB = 0, uC = 0, wA = 0, wB = 0, wC = 0;
  // old xx, xz, zz of the lane's position at rows r-1 and r-2
  float xx1 = 0, xx2 = 0, xz1 = 0, xz2 = 0, zz1 = 0, zz2 = 0;
  // Two prefetch sets, loaded two rows ahead and used alternately: the set
  // of row R holds its plane cells, xy(R+1), yy(R+2), yz(R+1) of the lane's
  // column and the velocity source of row R.
  f4 pfA[NPF], pfB[NPF];
  float qxyA = 0, qyyA = 0, qyzA = 0, quA, qvA, qwA, qxyB = 0, qyyB = 0, qyzB = 0, quB, qvB, qwB;
#define FUSED_FETCH(P, R)                                                              \
  do {                                                                              \
    FUSED_PLANES(pf##P, R);                                                            \
    if (int_xz) { qxy##P = FUSED_COL(oxy, (R) + 1); qyy##P = FUSED_COL(oyy, (R) + 2); qyz##P = FUSED_COL(oyz, (R) + 1); } \
    FUSED_VSRC(R, qu##P, qv##P, qw##P);                                                \
  } while (0)
// Iteration r with prefetch set P.
#define FUSED_ITER(P)                                                                  \
  do {                                                                              \
    const int buf = r & 1;                                                          \
    float *S = Sb[buf];                                                             \
    float(*V)[NZP * PW] = Vb[buf];                                                  \
    /* 1. planes of row r into LDS; new velocity of row r-2 into the velocity plane */ \
    _Pragma("unroll") for (int n = 0; n < NPF; ++n) {                               \
      const int m = wv + n * NWV;                                                   \
      if (m < NWL && pslot[n] >= 0) *(f4 *)&S[pslot[n]] = pf##P[n];                 \
    }                                                                               \
    if (has) { V[0][t] = uB; V[1][t] = vC; V[2][t] = wB; }                          \
    const float xyp1 = qxy##P, yyp2 = qyy##P, yzp1 = qyz##P;                        \
    const float uo = qu##P, vo = qv##P, wo = qw##P;                                 \
    /* 2. loads of row r+2 into this set, in flight across the barrier */           \
    if (r + 2 <= r1) FUSED_FETCH(P, r + 2);                                            \
    __syncthreads();                                                                \
    FUSED_COMPUTE();                                                                   \
  } while (0)
// New velocity of the lane's position at row r, new stresses of row r-2,
// then the windows advance to row r+1.
#define FUSED_COMPUTE()                                                                \
  do {                                                                              \
    float un = uo, vn = vo, wn = wo;                                                \
    const float xx0 = S[at<0>(x, z)], xz0 = S[at<2>(x, z)], zz0 = S[at<4>(x, z)];   \
    if (int_xz && r >= jlo && r <= jhi) {                                           \
      const int ax = at<0>(x, z), ay = at<1>(x, z), az = at<2>(x, z);               \
      const int ayz = at<3>(x, z), azz = at<4>(x, z);                               \
      constexpr int WXZ = PW, WYZ = PW, WZZ = PW;                                   \
      un = uo + b * (c1 * (xx0 - S[ax - 1]) + c2 * (S[ax + 1] - S[ax - 2]) +         \
                     c1 * (xy0 - xym1) + c2 * (xyp1 - xym2) +                       \
                     c1 * (xz0 - S[az - WXZ]) + c2 * (S[az + WXZ] - S[az - 2 * WXZ])); \
      vn = vo + b * (c1 * (S[ay + 1] - xy0) + c2 * (S[ay + 2] - S[ay - 1]) +         \
                     c1 * (yyp1 - yy0) + c2 * (yyp2 - yym1) +                       \
                     c1 * (yz0 - S[ayz - WYZ]) + c2 * (S[ayz + WYZ] - S[ayz - 2 * WYZ])); \
      wn = wo + b * (c1 * (S[az + 1] - xz0) + c2 * (S[az + 2] - S[az - 1]) +         \
                     c1 * (yz0 - yzm1) + c2 * (yzp1 - yzm2) +                       \
                     c1 * (S[azz + WZZ] - zz0) + c2 * (S[azz + 2 * WZZ] - S[azz - WZZ])); \
      if (core && r >= j0 && r <= je) {                                             \
        const int q = qb + (r - 1) * sy;                                            \
        st32_nt(nu, q, un); st32_nt(nv, q, vn); st32_nt(nw, q, wn);                 \
      }                                                                             \
    }                                                                               \
    const int s = r - 2;                                                            \
    if (core_s && s >= j0 && s <= je) {                                             \
      const float *Lu = V[0], *Lv = V[1], *Lw = V[2];                               \
      const float u0 = uB, v0 = vC, w0 = wB;                                        \
      const float dux1 = c1 * (Lu[t + 1] - u0), dux2 = c2 * (Lu[t + 2] - Lu[t - 1]); \
      const float dvy1 = c1 * (v0 - vB), dvy2 = c2 * (vD - vA);                      \
      const float dwz1 = c1 * (w0 - Lw[t - PW]), dwz2 = c2 * (Lw[t + PW] - Lw[t - 2 * PW]); \
      const int q = qb + (s - 1) * sy;                                              \
      st32_nt(nxx, q, (xx2 + da * (dux1 + dux2)) + db * (((dvy1 + dvy2) + dwz1) + dwz2)); \
      st32_nt(nyy, q, (yym2 + da * (dvy1 + dvy2)) + db * (((dux1 + dux2) + dwz1) + dwz2)); \
      st32_nt(nzz, q, (zz2 + da * (dwz1 + dwz2)) + db * (((dux1 + dux2) + dvy1) + dvy2)); \
      st32_nt(nxy, q, xym2 + dxy * (((c1 * (uC - u0) + c2 * (un - uA)) + c1 * (v0 - Lv[t - 1])) + \
                                   c2 * (Lv[t + 1] - Lv[t - 2])));                  \
      st32_nt(nxz, q, xz2 + dxz * (((c1 * (Lu[t + PW] - u0) + c2 * (Lu[t + 2 * PW] - Lu[t - PW])) + \
                                   c1 * (w0 - Lw[t - 1])) + c2 * (Lw[t + 1] - Lw[t - 2]))); \
      st32_nt(nyz, q, yzm2 + dxz * (((c1 * (Lv[t + PW] - v0) + c2 * (Lv[t + 2 * PW] - Lv[t - PW])) + \
                                    c1 * (wC - w0)) + c2 * (wn - wA)));             \
    }                                                                               \
    xx2 = xx1; xx1 = xx0; xz2 = xz1; xz1 = xz0; zz2 = zz1; zz1 = zz0;               \
    vA = vB; vB = vC; vC = vD; vD = vn;                                             \
    uA = uB; uB = uC; uC = un;                                                      \
    wA = wB; wB = wC; wC = wn;                                                      \
    xym2 = xym1; xym1 = xy0; xy0 = xyp1;                                            \
    yym2 = yym1; yym1 = yy0; yy0 = yyp1; yyp1 = yyp2;                               \
    yzm2 = yzm1; yzm1 = yz0; yz0 = yzp1;                                            \
  } while (0)

  FUSED_FETCH(A, r0);
  if (r0 + 1 <= r1) FUSED_FETCH(B, r0 + 1);
  for (int r = r0;;) {
    FUSED_ITER(A);
    if (++r > r1) break;
    FUSED_ITER(B);
    if (++r > r1) break;
  }
#undef FUSED_FETCH
#undef FUSED_ITER
#undef FUSED_COMPUTE
#undef FUSED_COL
#undef FUSED_VSRC
#undef FUSED_PLANES
}

}  // namespace fused
}  // namespace fd3d
