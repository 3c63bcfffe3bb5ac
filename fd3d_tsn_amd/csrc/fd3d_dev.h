// Device-side data layout of one rupture problem and the launch wrappers
// implemented in kernels.hip.
//
// HBM layout.  Every 3D wavefield array (u, v, w, xx, yy, zz, xy, yz, xz,
// two sets) is its own float array in the reference's column-major order
// with a padded x pitch: element (i,j,k), 1-based, lives at
//   off0 + (i-1) + px*((j-1) + nyt*(k-1)),
// where off0 shifts each row so that the first interior point i = nabc+1
// starts a 256-byte boundary and px is a multiple of 64 floats.  A
// wavefront (64 lanes along x) then reads whole 128-byte lines.  The fault
// plane arrays (nxt, nzt) are unpadded.  The medium is depth-layered, as
// the reference builds it (lam1, mu1, d1 constant over each z plane,
// fd3d_init.f90:240-265): lam, mu and rho are tables of one value per
// plane, index k.  PML split fields live in one pool,
// region by region, split by split, each split a dense (ni, nj, nk) box.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fd3d {

struct Dims {
  int nxt, nyt, nzt, nabc, nfs, fspace;
  int px, off0;
  int sy, sz;        // strides of j and k in a 3D array
  long long n3;      // floats per padded 3D array
  int nxT, nzT;      // output slab (nxt-2nabc, nzt-nfs-nabc)
};

__host__ __device__ inline int I3(const Dims &g, int i, int j, int k) {
  return g.off0 + (i - 1) + g.sy * (j - 1) + g.sz * (k - 1);
}
__host__ __device__ inline int I2(const Dims &g, int i, int k) { return (i - 1) + g.nxt * (k - 1); }

// Quotient x / d for 0 <= x < 2^31 from m = ceil(2^(31+l) / d) and
// sh = 31 + l, l = ceil(log2 d): with x * d < 2^(31+l) the rounding of m
// never reaches the next integer, so the result is exact.
__host__ __device__ inline int udiv(int x, unsigned m, int sh) {
  return (int)(((unsigned long long)(unsigned)x * m) >> sh);
}
__host__ inline void udiv_magic(int d, unsigned &m, int &sh) {
  int l = 0;
  while ((1LL << l) < d) ++l;
  sh = 31 + l;
  m = (unsigned)(((1ULL << sh) + (unsigned long long)d - 1) / (unsigned long long)d);
}

// Element access at a 32-bit element index from a wave-uniform base: the
// memory unit adds the byte offset to the scalar base, so no 64-bit vector
// address arithmetic is spent per access.  Every 3D array holds fewer than
// 2^30 floats (checked at create), so byte offsets fit 32 bits.
__device__ __forceinline__ float ld32(const float *base, int idx) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + ((unsigned)idx << 2));
}
__device__ __forceinline__ float ld32_nt(const float *base, int idx) {
  return __builtin_nontemporal_load(reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) +
                                                                     ((unsigned)idx << 2)));
}
__device__ __forceinline__ void st32(float *base, int idx, float v) {
  *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + ((unsigned)idx << 2)) = v;
}
__device__ __forceinline__ void st32_nt(float *base, int idx, float v) {
  __builtin_nontemporal_store(v, reinterpret_cast<float *>(reinterpret_cast<char *>(base) + ((unsigned)idx << 2)));
}

// Float constants formed on the host exactly as the reference forms them.
struct Coef {
  float dh, dt;
  float dth;   // dt/dh
  float dthh;  // (0.5*dt)/dh, the near-fault shear kernels (fd3d_deriv.f90:696,902)
  float dht;   // dh/dt
  float damp_s, c1, c2;
};

struct Fields {
  float *u, *v, *w, *xx, *yy, *zz, *xy, *yz, *xz;
  float *lam, *mu, *d;   // per-plane medium tables, index k = 1..nzt (0 and nzt+1 hold 0)
};

struct Fault {
  // traction_com
  float *tx, *tz, *v1t, *avdx, *avdz, *RFx, *RFz, *au1, *av1, *aw1;
  // friction_com
  float *uZ, *wX, *tabsX, *tabsZ, *T0X, *T0Z, *Dc, *coh;
  float *peakX, *dynX, *DcX, *peakZ, *dynZ, *DcZ;
  double *aX, *bX, *psiX, *vwX, *aZ, *bZ, *psiZ, *vwZ;
  // source_com and the loop-local slip paths
  float *ruptime, *rise, *sliptime, *slipX, *slipZ, *schangeX, *schangeZ, *distX, *distZ;
};

struct FvwScalars {
  double Sn, uini, wini, f0, v0, fw, hx0, hz0, perturb, RR2, TT2, strinixI;
  int newton_max, fvw, nucleation;
};

struct PmlRegion {
  int ib, ie, jb, je_uw, je_v, je_xz, kb, ke;
  int ni, nj, nk, ncell;
  unsigned mdi, mdj;       // x / ni = udiv(x, mdi, sdi) and x / nj likewise, for 0 <= x < 2^31
  int sdi, sdj;
  float *sp;               // the region's 24 splits, split n at sp + n * ncell
  int ztop, nzx, nzy;      // zero xz/yz splits at k2 = nk for j2 <= nzx / nzy
  int blk0;                // first block of this region in the fused PML launch
  long long split_off;     // offset of split 0 in the pool (floats)
  int prof_off;            // offset of the profile block in the profile pool
};

// Logical block id of hardware block pid when blocks are dealt round-robin
// to 8 XCDs: XCD x gets the contiguous logical range [x*N/8, (x+1)*N/8),
// so blocks that work on neighbouring data share that XCD's L2.
__device__ inline int xcd_remap(int pid, int N) {
  const int q = N >> 3, r = N & 7;
  const int xcd = pid & 7, loc = pid >> 3;
  return xcd < r ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
}

// Non-temporal access for values a sweep touches once: the new wavefield
// values it writes and the old centre values it reads.  Their lines then do
// not displace the stencil operands that neighbouring cells re-read from L2
// and the last-level cache, and the next sweep does not start behind the
// write-back of dirty lines.
__device__ __forceinline__ void st_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ float ld_nt(const float *p) { return __builtin_nontemporal_load(p); }

constexpr int kMaxRegions = 5;
constexpr int kSplits = 24;
constexpr int kPmlBlock = 256;
// tiled PML sweeps (pml.hip): 64 x kPmlTZ x-z tiles on the interior tile
// grid, each block marching at most kPmlRows rows along y
constexpr int kPmlTZ = 8;
constexpr int kPmlRows = 32;

struct Pml {
  int nreg, nblk;
  PmlRegion r[kMaxRegions];
  float *splits;
  // per region 12 profile arrays: for p in (ox, oxS, oy, oyS, oz, ozS):
  // (1 - omega) then (1 + omega), 1-based, lengths ni+1 / nj+1 / nk+1
  float *prof;
  // tiled sweeps: one (bx, bz, j0, j1) entry per block, and a device copy
  // of the region table (indexed by a wave-uniform region number)
  int4 *tiles;
  int ntiles;
  const PmlRegion *rdev;
};

struct Ctl {
  int it;              // steps completed since init; the state after step n is in field set n & 1
  int newton_capped;   // Newton solves that hit the cap
  int nonfinite;       // NaN/Inf seen in the slip-rate outputs
  int pad;
};

// Source-time bins for dynamic inversion (SlipRates_com, fd3d_init.f90:
// 116-120; accumulated at fd3d_theo.f90:560-578): mean slip rates over
// NL x NW fault patches (msrx, msrz, patch-major then bin) and the moment
// rate, in bins of dtseis seconds.
struct Rates {
  float *msrx, *msrz, *mr;
  int NL, NW, nbins, on;
  float dL, dW, dtseis;
};

struct Outputs {
  float *ring;         // [ring][FD3D_NOUT][nxT*nzT]
  float *seis;         // [ring][nsta][3]
  const int *sta;      // [nsta][3] 1-based
  int ring_steps, nsta;
};

// Launch shapes of the interior and PML sweeps; chosen per grid size by
// the autotuner.  mode bit 0: tiled interior sweeps (interior.hip) instead
// of one thread per cell; bit 1: tiled PML sweeps (pml.hip) instead of
// one thread per cell (kernels.hip); bit 2: the fused interior sweep
// (fused.h; rows per block from ny_vel; under a free surface followed by
// the top-plane stresses of k_str_top); bits 3 and 4: velocity and stress
// tiles 128 cells wide along x instead of 64.
struct Tuning {
  int mode;
  int tz_vel, ny_vel;   // velocity: x-z tile height (4, 8, 16), rows per block along y
  int tz_str, ny_str;   // stress
};

struct Problem {
  Tuning tune;
  Dims g;
  Coef c;
  Fields f[2];         // the two wavefield sets (u..xz); lam, mu, d shared
  Fault fl;
  FvwScalars fv;
  Pml pml;
  Outputs out;
  Rates rates;
  Ctl *ctl;
};

// In the order the stages run within a step (the batch engine launches its
// stages in this order).  The step counter advances once the wavefield of
// the step is complete (K_ADVANCE); the fault-plane stages after it read
// that step's new set and time, and touch neither the wavefield nor
// anything the next step's velocity sweeps read, so a single problem runs
// them beside the next step's PML and interior velocity sweeps.
enum KernelId {
  K_VEL_INTERIOR = 0, K_VEL_FAULT, K_PML_VEL,
  K_FUSED,             // interior velocities and stresses in one pass
  K_FUVW_UV, K_FUVW_W,
  K_STR_TOP,           // with the fused pass and a free surface: stresses of the top two interior planes
  K_STR_INTERIOR, K_STR_FAULT, K_PML_STR, K_FRES,
  K_ADVANCE,
  K_TRACTION, K_TRACTION_MIRROR, K_TABS, K_FRICTION, K_OUTPUTS, K_RATES, K_NUCLEATION,
  K_COUNT
};
extern const char *kKernelNames[K_COUNT];

// Launch one kernel of the step on a stream; returns hipSuccess or the launch error.
hipError_t launch(KernelId id, const Problem &p, hipStream_t s);
// Tiled interior sweep (velocity or stress), interior.hip.
hipError_t launch_interior(bool stress, const Problem &p, hipStream_t s);
// Fused interior velocity + stress sweep, fused.hip.
hipError_t launch_fused(const Problem &p, hipStream_t s);
// Tiled PML sweep (velocity or stress), pml.hip.
hipError_t launch_pml_tiled(bool stress, const Problem &p, hipStream_t s);
// Fill helpers used by the prologue.
hipError_t fill_f32(float *p, float v, long long n, hipStream_t s);

}  // namespace fd3d
