// Fused interior sweep: the velocity update (dvel, fd3d_deriv.f90:24-193)
// and the stress update (dstres, :513-622) of the interior cells in one
// pass over HBM.
//
// Separately, the two sweeps move 27 words per interior cell (velocity: 6
// old stresses and 3 old velocities read, 3 new velocities written;
// stress: the 3 new velocities and 6 old stresses read, 6 written).  Fused,
// the old stresses and velocities are read once and the new values written
// once: 18 words.  The two wavefield sets make this race-free: every block
// reads the old set and writes the new one.
//
// Work of a block (64 x 16 lanes).  Lane (tx, ty) owns position
// x = tx - 2, z = ty - 2 of an x-z tile whose core, the cells that get new
// stresses, is x in [0, 59], z in [0, 11]; the block marches along y over
// the rows [j0, je] of its chunk, from j0 - 2 to je + 2.  Iteration r:
//   1. the old-stress planes of row r (loaded into registers during
//      iteration r - 1) go into one of two LDS buffers, and every lane puts
//      its new velocity of row r - 2 (from its register window) into one of
//      two velocity planes; one barrier;
//   2. the loads of row r + 1 are issued; every lane forms the new velocity
//      of its position at row r from the stress planes (x and z neighbours)
//      and its register windows of xy, yy, yz (y neighbours); core lanes
//      form the new stresses of row r - 2 from the velocity plane (x and z
//      neighbours) and their velocity windows (y neighbours).
// Lanes on the 2-cell ring around the core recompute velocities that the
// neighbouring tiles own, so no block waits on another, and every lane
// does the same work.  Positions outside the interior (PML cells, the
// near-fault rows j >= nyt-1, the outer faces) take the values pml_vel and
// vel_fault already wrote into the new set: the kernel runs after them and
// before str_fault and pml_str, which read its new velocities.
//
// Per-point arithmetic is that of tiles.h (the same operands, the same
// association), so the result is bit-identical to the separate sweeps.
// Full-space problems only (with a free surface the ghost velocities of
// fuvw must be formed between the two updates).
#include "stages.h"

namespace fd3d {
namespace {

// global-memory floats: keeps plane loads global_load when their field
// pointer is chosen per wavefront
typedef __attribute__((address_space(1))) const float gfloat;

constexpr int LX = 64, LZ = 16;          // lanes along x and z
constexpr int CX = LX - 4, CZ = LZ - 4;  // core: 60 x 12 cells
constexpr int NT = LX * LZ;

// Old-stress planes of one row, relative to the core origin; every plane
// spans what the velocity stencils of the lane positions x in [-2, 61],
// z in [-2, 13] read.  The five planes lie back to back in one buffer,
// each padded to whole wavefronts (64 cells) so that a wavefront's loads
// of a plane row all come from one field.
template <int X0, int Z0, int W_, int H_, int OFF>
struct SP {
  static constexpr int W = W_, N = W_ * H_, NW = (N + 63) / 64, off = OFF, end = OFF + 64 * NW;
  static constexpr int x0 = X0, z0 = Z0;
  __device__ static int at(int x, int z) { return OFF + (z - Z0) * W_ + (x - X0); }
};
using PXX = SP<-4, -2, 67, 16, 0>;          // xx: x in [-4, 62]
using PXY = SP<-3, -2, 67, 16, PXX::end>;   // xy: x in [-3, 63]
using PXZ = SP<-3, -4, 67, 19, PXY::end>;   // xz: x in [-3, 63], z in [-4, 14]
using PYZ = SP<-2, -4, 64, 19, PXZ::end>;   // yz: z in [-4, 14]
using PZZ = SP<-2, -3, 64, 19, PYZ::end>;   // zz: z in [-3, 15]
constexpr int NSP = PZZ::end;               // 5888 floats
constexpr int NWL = NSP / 64;               // wavefront loads per row (92)
constexpr int NPF = (NWL + LZ - 1) / LZ;    // per wavefront (6)

struct F2Lds {
  float s[2][NSP];        // old stresses of row r (two buffers)
  float v[2][3][NT];      // new u, v, w of row r - 2 at every lane position (two buffers)
};

// Wavefront load m (wave-uniform) and lane -> field (0..4 = xx, xy, xz,
// yz, zz), whether the cell is one of the plane's, and its position.
template <class P>
__device__ inline bool sp_pos(int e, int &x, int &z) {
  z = e / P::W;
  x = e - z * P::W + P::x0;
  z += P::z0;
  return e < P::N;
}
__device__ inline int sp_cell(int m, int lane, bool &ok, int &x, int &z) {
  const int h = m * 64 + lane;
  if (m < PXY::off / 64) { ok = sp_pos<PXX>(h - PXX::off, x, z); return 0; }
  if (m < PXZ::off / 64) { ok = sp_pos<PXY>(h - PXY::off, x, z); return 1; }
  if (m < PYZ::off / 64) { ok = sp_pos<PXZ>(h - PXZ::off, x, z); return 2; }
  if (m < PZZ::off / 64) { ok = sp_pos<PYZ>(h - PYZ::off, x, z); return 3; }
  ok = sp_pos<PZZ>(h - PZZ::off, x, z);
  return 4;
}

__global__ void __launch_bounds__(NT) k_vs_fused(Dims g, Coef c, Fields f0, Fields f1, const Ctl *ctl, int gx,
                                                       int gz, int gy, int NY) {
  __shared__ F2Lds L;
  // the old set (fo) is set it & 1, the new one the other (see field_pair);
  // only the pointers this kernel uses are selected
  const bool odd = ctl->it & 1;
  const int lid = xcd_remap(blockIdx.x, gx * gz * gy);
  const int bx = lid % gx, bz = (lid / gx) % gz, by = lid / (gx * gz);
  const int tx = threadIdx.x, ty = threadIdx.y, t = ty * LX + tx;
  const int wv = __builtin_amdgcn_readfirstlane(ty);   // the wavefront (one lane row)
  const int i0 = g.nabc + 1 + bx * CX, k0 = g.nabc + 1 + bz * CZ;   // core origin
  const int j0 = g.nabc + 1 + by * NY, je = min(j0 + NY - 1, g.nyt - 2);
  const int x = tx - 2, z = ty - 2, i = i0 + x, k = k0 + z;
  const bool in_xz = i >= 1 && i <= g.nxt && k >= 1 && k <= g.nzt;
  const bool int_xz = i >= g.nabc + 1 && i <= g.nxt - g.nabc && k >= g.nabc + 1 && k <= g.nzt - g.nfs;
  const bool core = int_xz && x >= 0 && x < CX && z >= 0 && z < CZ;
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

#define F2_SEL(a) const float *o##a = odd ? f1.a : f0.a; float *n##a = odd ? f0.a : f1.a
  F2_SEL(xx); F2_SEL(yy); F2_SEL(zz); F2_SEL(xy); F2_SEL(yz); F2_SEL(xz); F2_SEL(u); F2_SEL(v); F2_SEL(w);
#undef F2_SEL
  float *nu_ = nu, *nv_ = nv, *nw_ = nw;
#define F2_COL(f, j) ((in_xz && (j) >= 1 && (j) <= g.nyt) ? ld32(f, qb + ((j) - 1) * sy) : 0.0f)
// velocity source of row j: the old set where this kernel updates the
// position, the new set where another kernel already did, 0 outside
#define F2_VSRC(j, a, bb, cc)                                                      \
  do {                                                                              \
    a = bb = cc = 0.0f;                                                             \
    if (in_xz && (j) >= 1 && (j) <= g.nyt) {                                        \
      const int q_ = qb + ((j) - 1) * sy;                                           \
      if (int_xz && (j) >= jlo && (j) <= jhi) { a = ld32(ou, q_); bb = ld32(ov, q_); cc = ld32(ow, q_); } \
      else { a = ld32(nu_, q_); bb = ld32(nv_, q_); cc = ld32(nw_, q_); }            \
    }                                                                               \
  } while (0)
// plane cells of row j into registers: wavefront wv takes wavefront loads
// m = wv, wv + LZ, ...; each is one field (a uniform base pointer).  The
// row-independent part of each cell's offset is formed once (-1: no cell).
  int poff[NPF];
  const gfloat *psrc[NPF];   // wave-uniform field of each load
#pragma unroll
  for (int n = 0; n < NPF; ++n) {
    const int m = wv + n * LZ;
    poff[n] = -1;
    psrc[n] = (const gfloat *)oxx;
    if (m < NWL) {
      int px, pz;
      bool ok;
      const int f = sp_cell(m, tx, ok, px, pz);
      psrc[n] = (const gfloat *)(f == 0 ? oxx : f == 1 ? oxy : f == 2 ? oxz : f == 3 ? oyz : ozz);
      const int gi = i0 + px, gk = k0 + pz;
      if (ok && gi >= 1 && gi <= g.nxt && gk >= 1 && gk <= g.nzt) poff[n] = I3(g, gi, 1, gk);
    }
  }
#define F2_PLANES(pf, j)                                                            \
  do {                                                                              \
    const bool jok_ = (j) >= 1 && (j) <= g.nyt;                                     \
    const int jo_ = ((j) - 1) * sy;                                                 \
    _Pragma("unroll") for (int n = 0; n < NPF; ++n)                                 \
      pf[n] = (jok_ && poff[n] >= 0) ? *(const gfloat *)((const __attribute__((address_space(1))) char *)psrc[n] + ((unsigned)(poff[n] + jo_) << 2)) : 0.0f; \
  } while (0)

  // y windows of the old stresses: xy rows r-2..r, yy rows r-2..r+1, yz
  // rows r-2..r before iteration r (its prefetch set adds xy(r+1),
  // yy(r+2), yz(r+1))
  float xym2 = 0, xym1 = 0, xy0 = 0, yym2 = 0, yym1 = 0, yy0 = 0, yyp1 = 0, yzm2 = 0, yzm1 = 0, yz0 = 0;
  if (int_xz) {
    xym2 = F2_COL(oxy, r0 - 2); xym1 = F2_COL(oxy, r0 - 1); xy0 = F2_COL(oxy, r0);
    yym2 = F2_COL(oyy, r0 - 2); yym1 = F2_COL(oyy, r0 - 1); yy0 = F2_COL(oyy, r0); yyp1 = F2_COL(oyy, r0 + 1);
    yzm2 = F2_COL(oyz, r0 - 2); yzm1 = F2_COL(oyz, r0 - 1); yz0 = F2_COL(oyz, r0);
  }
  // new velocities of the lane's column: v rows r-4..r-1 (vA..vD), u and w
  // rows r-3..r-1 (uA..uC, wA..wC)
  float vA = 0, vB = 0, vC = 0, vD = 0, uA = 0, uB = 0, uC = 0, wA = 0, wB = 0, wC = 0;
  // old xx, xz, zz of the lane's position at rows r-1 and r-2
  float xx1 = 0, xx2 = 0, xz1 = 0, xz2 = 0, zz1 = 0, zz2 = 0;
  // Two prefetch sets, loaded two rows ahead and used alternately: the set
  // of row R holds its plane cells, xy(R+1), yy(R+2), yz(R+1) of the lane's
  // column and the velocity source of row R.
  float pfA[NPF], pfB[NPF];
  float qxyA = 0, qyyA = 0, qyzA = 0, quA, qvA, qwA, qxyB = 0, qyyB = 0, qyzB = 0, quB, qvB, qwB;
#define F2_FETCH(P, R)                                                              \
  do {                                                                              \
    F2_PLANES(pf##P, R);                                                            \
    if (int_xz) { qxy##P = F2_COL(oxy, (R) + 1); qyy##P = F2_COL(oyy, (R) + 2); qyz##P = F2_COL(oyz, (R) + 1); } \
    F2_VSRC(R, qu##P, qv##P, qw##P);                                                \
  } while (0)
// Iteration r with prefetch set P.
#define F2_ITER(P)                                                                  \
  do {                                                                              \
    const int buf = r & 1;                                                          \
    float *S = L.s[buf];                                                            \
    float(*V)[NT] = L.v[buf];                                                       \
    /* 1. planes of row r into LDS; new velocity of row r-2 into the velocity plane */ \
    _Pragma("unroll") for (int n = 0; n < NPF; ++n) {                               \
      const int m = wv + n * LZ;                                                    \
      if (m < NWL) S[m * 64 + tx] = pf##P[n];                                       \
    }                                                                               \
    V[0][t] = uB; V[1][t] = vC; V[2][t] = wB;                                       \
    const float xyp1 = qxy##P, yyp2 = qyy##P, yzp1 = qyz##P;                        \
    const float uo = qu##P, vo = qv##P, wo = qw##P;                                 \
    __syncthreads();                                                                \
    /* 2. loads of row r+2 into this set */                                         \
    if (r + 2 <= r1) F2_FETCH(P, r + 2);                                            \
    F2_COMPUTE();                                                                   \
  } while (0)
// New velocity of the lane's position at row r, new stresses of row r-2,
// then the windows advance to row r+1.
#define F2_COMPUTE()                                                                \
  do {                                                                              \
    float un = uo, vn = vo, wn = wo;                                                \
    const float xx0 = S[PXX::at(x, z)], xz0 = S[PXZ::at(x, z)], zz0 = S[PZZ::at(x, z)]; \
    if (int_xz && r >= jlo && r <= jhi) {                                           \
      const int ax = PXX::at(x, z), ay = PXY::at(x, z), az = PXZ::at(x, z);          \
      const int ayz = PYZ::at(x, z), azz = PZZ::at(x, z);                           \
      constexpr int WXZ = PXZ::W, WYZ = PYZ::W, WZZ = PZZ::W;                       \
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
    if (core && s >= j0 && s <= je) {                                               \
      const float *Lu = V[0], *Lv = V[1], *Lw = V[2];                               \
      const float u0 = uB, v0 = vC, w0 = wB;                                        \
      const float dux1 = c1 * (Lu[t + 1] - u0), dux2 = c2 * (Lu[t + 2] - Lu[t - 1]); \
      const float dvy1 = c1 * (v0 - vB), dvy2 = c2 * (vD - vA);                      \
      const float dwz1 = c1 * (w0 - Lw[t - LX]), dwz2 = c2 * (Lw[t + LX] - Lw[t - 2 * LX]); \
      const int q = qb + (s - 1) * sy;                                              \
      st32(nxx, q, (xx2 + da * (dux1 + dux2)) + db * (((dvy1 + dvy2) + dwz1) + dwz2)); \
      st32(nyy, q, (yym2 + da * (dvy1 + dvy2)) + db * (((dux1 + dux2) + dwz1) + dwz2)); \
      st32(nzz, q, (zz2 + da * (dwz1 + dwz2)) + db * (((dux1 + dux2) + dvy1) + dvy2)); \
      st32(nxy, q, xym2 + dxy * (((c1 * (uC - u0) + c2 * (un - uA)) + c1 * (v0 - Lv[t - 1])) + \
                                   c2 * (Lv[t + 1] - Lv[t - 2])));                  \
      st32(nxz, q, xz2 + dxz * (((c1 * (Lu[t + LX] - u0) + c2 * (Lu[t + 2 * LX] - Lu[t - LX])) + \
                                   c1 * (w0 - Lw[t - 1])) + c2 * (Lw[t + 1] - Lw[t - 2]))); \
      st32(nyz, q, yzm2 + dxz * (((c1 * (Lv[t + LX] - v0) + c2 * (Lv[t + 2 * LX] - Lv[t - LX])) + \
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

  F2_FETCH(A, r0);
  if (r0 + 1 <= r1) F2_FETCH(B, r0 + 1);
  for (int r = r0;;) {
    F2_ITER(A);
    if (++r > r1) break;
    F2_ITER(B);
    if (++r > r1) break;
  }
#undef F2_FETCH
#undef F2_ITER
#undef F2_COMPUTE
#undef F2_COL
#undef F2_VSRC
#undef F2_PLANES
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

hipError_t launch_fused(const Problem &p, hipStream_t s) {
  const Dims &g = p.g;
  if (g.nyt - 2 - g.nabc <= 0) return hipSuccess;
  const int ny = p.tune.ny_vel;
  if (ny < 1) return hipErrorInvalidValue;
  const int gx = cdiv(g.nxt - 2 * g.nabc, CX), gz = cdiv(g.nzt - g.nfs - g.nabc, CZ);
  const int gy = cdiv(g.nyt - 2 - g.nabc, ny);
  hipLaunchKernelGGL(k_vs_fused, dim3(gx * gz * gy), dim3(LX, LZ), 0, s, g, p.c, p.f[0], p.f[1], p.ctl, gx, gz, gy,
                     ny);
  return hipGetLastError();
}

}  // namespace fd3d
