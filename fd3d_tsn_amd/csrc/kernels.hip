// Kernels of one FD3D_TSN time step for a single problem: one launch per
// stage, each a thin wrapper around the stage bodies of stages.h.
//
// Step structure (kernels in order):
//   velocity : vel_interior | vel_fault | pml_vel | (fuvw_uv, fuvw_w)
//   stress   : str_interior | str_fault | pml_str | (fres)
//   advance  : the step counter (which also selects the wavefield set)
//   fault BC : traction | (traction_mirror) | tabs | friction | outputs | (nucleation)
// The fault-plane kernels after advance read the step just counted (its
// new set: field_pair_done); capi.hip runs them beside the next step's
// velocity sweeps.
// Kernels in parentheses run only with a free surface / rate-and-state
// nucleation.  vel_fault performs the v mirror v(:,nyt,:) = v(:,nyt-1,:)
// for every column before pml_vel updates the PML columns, as bnd2d does
// before pml_uvw in the reference (fd3d_deriv.f90:225-232); str_fault does
// the same for xy and yz (:802-810).
#include "stages.h"

namespace fd3d {

const char *kKernelNames[K_COUNT] = {
    "vel_interior", "vel_fault", "pml_vel", "fused_interior", "fuvw_uv", "fuvw_w", "str_top",
    "str_interior", "str_fault", "pml_str", "fres", "advance",
    "traction", "traction_mirror", "tabs", "friction", "outputs", "rates", "nucleation"};

namespace {

constexpr int BY = 4;  // lane rows of the 64 x 4 blocks of the 2D stages

#define B2 blockIdx.x, blockIdx.y, threadIdx.x, threadIdx.y
#define PAIR                                     \
  const FieldPair FP = field_pair(f0, f1, ctl); \
  const Fields &fo = FP.fo, &fn = FP.fn
// the fault-plane kernels after advance: the step just counted
#define PAIR_DONE                                     \
  const FieldPair FP = field_pair_done(f0, f1, ctl); \
  const Fields &fo = FP.fo, &fn = FP.fn

__global__ void __launch_bounds__(256) k_vel_interior(Dims g, Coef c, Fields f0, Fields f1, const Ctl *ctl) {
  PAIR;
  vel_interior_body(g, c, fo, fn, blockIdx.x, blockIdx.y, blockIdx.z, threadIdx.x, threadIdx.y);
}
__global__ void __launch_bounds__(256) k_vel_fault(Dims g, Coef c, Fields f0, Fields f1, Fault fl, FvwScalars fv,
                                                   Ctl *ctl) {
  PAIR;
  vel_fault_body<BY>(g, c, fo, fn, fl, fv, ctl, B2);
}
__global__ void __launch_bounds__(kPmlBlock) k_pml_vel(Dims g, Coef c, Fields f0, Fields f1, Pml pm, const Ctl *ctl) {
  PAIR;
  pml_vel_body(g, c, fo, fn, pm, blockIdx.x, threadIdx.x);
}
__global__ void __launch_bounds__(256) k_fuvw_uv(Dims g, Fields f0, Fields f1, const Ctl *ctl) {
  PAIR;
  fuvw_uv_body<BY>(g, fo, fn, B2);
}
__global__ void __launch_bounds__(256) k_fuvw_w(Dims g, Fields f0, Fields f1, const Ctl *ctl) {
  PAIR;
  fuvw_w_body<BY>(g, fo, fn, B2);
}
__global__ void __launch_bounds__(256) k_str_interior(Dims g, Coef c, Fields f0, Fields f1, const Ctl *ctl) {
  PAIR;
  str_interior_body(g, c, fo, fn, blockIdx.x, blockIdx.y, blockIdx.z, threadIdx.x, threadIdx.y);
}
// Interior stresses of the planes k0 .. k0 + gridDim.z - 1 (the top two
// interior planes under a free surface, which need the ghost velocities of
// fuvw: the fused sweep leaves them to this kernel).
__global__ void __launch_bounds__(256) k_str_top(Dims g, Coef c, Fields f0, Fields f1, const Ctl *ctl, int gbz0) {
  PAIR;
  str_interior_body(g, c, fo, fn, blockIdx.x, blockIdx.y, gbz0 + (int)blockIdx.z, threadIdx.x, threadIdx.y);
}
__global__ void __launch_bounds__(256) k_str_fault(Dims g, Coef c, Fields f0, Fields f1, Fault fl, const Ctl *ctl) {
  PAIR;
  str_fault_body<BY>(g, c, fo, fn, fl, B2);
}
__global__ void __launch_bounds__(kPmlBlock) k_pml_str(Dims g, Coef c, Fields f0, Fields f1, Pml pm, const Ctl *ctl) {
  PAIR;
  pml_str_body(g, c, fo, fn, pm, blockIdx.x, threadIdx.x);
}
__global__ void __launch_bounds__(256) k_fres(Dims g, Fields f0, Fields f1, const Ctl *ctl) {
  PAIR;
  fres_body<BY>(g, fo, fn, B2);
}
__global__ void __launch_bounds__(256) k_traction(Dims g, Coef c, Fields f0, Fields f1, Fault fl, const Ctl *ctl) {
  PAIR_DONE;
  traction_body<BY>(g, c, fo, fn, fl, B2);
}
__global__ void __launch_bounds__(256) k_traction_mirror(Dims g, Fault fl) {
  traction_mirror_body(g, fl, blockIdx.x * 256 + threadIdx.x);
}
__global__ void __launch_bounds__(256) k_tabs(Dims g, Fault fl) { tabs_body<BY>(g, fl, B2); }
__global__ void __launch_bounds__(256) k_friction(Dims g, Coef c, Fields f0, Fields f1, Fault fl, FvwScalars fv,
                                                  const Ctl *ctl) {
  PAIR_DONE;
  friction_body<BY>(g, c, fo, fn, fl, fv, ctl, B2);
}
// The last block row records the stations.
__global__ void __launch_bounds__(256) k_outputs(Dims g, Fields f0, Fields f1, Fault fl, FvwScalars fv, Outputs o,
                                                 Ctl *ctl) {
  PAIR_DONE;
  if (blockIdx.y == gridDim.y - 1) {
    stations_body(g, fo, fn, o, ctl, blockIdx.x * 256 + threadIdx.y * BX + threadIdx.x);
    return;
  }
  outputs_body<BY>(g, fo, fn, fl, fv, o, ctl, B2);
}
__global__ void __launch_bounds__(256) k_nucleation(Dims g, Coef c, Fault fl, FvwScalars fv, const Ctl *ctl) {
  nucleation_body<BY>(g, c, fl, fv, ctl, B2);
}
__global__ void k_advance(Ctl *ctl) { advance_body(ctl); }
__global__ void __launch_bounds__(1024) k_rates(Dims g, Coef c, Fields f0, Fields f1, Outputs o, Rates R,
                                                const Ctl *ctl) {
  __shared__ double red[1024];
  PAIR_DONE;
  (void)fo;
  rates_body<1024>(g, c, fn, o, R, ctl, blockIdx.x, threadIdx.x, red);
}

#undef PAIR
#undef PAIR_DONE
#undef B2

__global__ void k_fill(float *p, float v, long long n) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; idx < n; idx += stride) p[idx] = v;
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

hipError_t fill_f32(float *p, float v, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  unsigned blocks = cdiv(n, 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch(KernelId id, const Problem &p, hipStream_t s) {
  const Dims &g = p.g;
  const int ni = g.nxt - 2 * g.nabc;              // interior / fault x extent
  const int nyi = g.nyt - 2 - g.nabc;              // interior y extent
  const int nk = g.nzt - g.nfs - g.nabc;           // interior / fault z extent
  const dim3 b2(BX, BY);
  switch (id) {
    case K_VEL_INTERIOR:
      if (nyi <= 0) return hipSuccess;
      if (p.tune.mode & 1) return launch_interior(false, p, s);
      hipLaunchKernelGGL(k_vel_interior, dim3(cdiv(ni, BX), cdiv(nyi, BY), nk), b2, 0, s, g, p.c, p.f[0], p.f[1], p.ctl);
      break;
    case K_VEL_FAULT:
      hipLaunchKernelGGL(k_vel_fault, dim3(cdiv(g.nxt, BX), cdiv(g.nzt, BY)), b2, 0, s, g, p.c, p.f[0], p.f[1],
                         p.fl, p.fv, p.ctl);
      break;
    case K_PML_VEL:
      if (p.pml.nblk == 0) return hipSuccess;
      if (p.tune.mode & 2) return launch_pml_tiled(false, p, s);
      hipLaunchKernelGGL(k_pml_vel, dim3(p.pml.nblk), dim3(kPmlBlock), 0, s, g, p.c, p.f[0], p.f[1], p.pml, p.ctl);
      break;
    case K_FUVW_UV:
      hipLaunchKernelGGL(k_fuvw_uv, dim3(cdiv(g.nxt, BX), cdiv(g.nyt, BY)), b2, 0, s, g, p.f[0], p.f[1], p.ctl);
      break;
    case K_FUVW_W:
      hipLaunchKernelGGL(k_fuvw_w, dim3(cdiv(g.nxt, BX), cdiv(g.nyt, BY)), b2, 0, s, g, p.f[0], p.f[1], p.ctl);
      break;
    case K_STR_INTERIOR:
      if (nyi <= 0) return hipSuccess;
      if (p.tune.mode & 1) return launch_interior(true, p, s);
      hipLaunchKernelGGL(k_str_interior, dim3(cdiv(ni, BX), cdiv(nyi, BY), nk), b2, 0, s, g, p.c, p.f[0], p.f[1], p.ctl);
      break;
    case K_STR_FAULT:
      hipLaunchKernelGGL(k_str_fault, dim3(cdiv(g.nxt, BX), cdiv(g.nzt, BY)), b2, 0, s, g, p.c, p.f[0], p.f[1], p.fl,
                         p.ctl);
      break;
    case K_PML_STR:
      if (p.pml.nblk == 0) return hipSuccess;
      if (p.tune.mode & 2) return launch_pml_tiled(true, p, s);
      hipLaunchKernelGGL(k_pml_str, dim3(p.pml.nblk), dim3(kPmlBlock), 0, s, g, p.c, p.f[0], p.f[1], p.pml, p.ctl);
      break;
    case K_FRES:
      hipLaunchKernelGGL(k_fres, dim3(cdiv(g.nxt, BX), cdiv(g.nyt, BY)), b2, 0, s, g, p.f[0], p.f[1], p.ctl);
      break;
    case K_TRACTION:
      hipLaunchKernelGGL(k_traction, dim3(cdiv(ni, BX), cdiv(nk, BY)), b2, 0, s, g, p.c, p.f[0], p.f[1], p.fl, p.ctl);
      break;
    case K_TRACTION_MIRROR:
      hipLaunchKernelGGL(k_traction_mirror, dim3(cdiv(ni, 256)), dim3(256), 0, s, g, p.fl);
      break;
    case K_TABS:
      hipLaunchKernelGGL(k_tabs, dim3(cdiv(ni, BX), cdiv(nk, BY)), b2, 0, s, g, p.fl);
      break;
    case K_FRICTION:
      hipLaunchKernelGGL(k_friction, dim3(cdiv(ni, BX), cdiv(nk, BY)), b2, 0, s, g, p.c, p.f[0], p.f[1], p.fl,
                         p.fv, p.ctl);
      break;
    case K_OUTPUTS: {
      unsigned gx = cdiv(ni, BX);
      unsigned sta_blocks = cdiv(p.out.nsta, BX * BY);
      if (sta_blocks > gx) gx = sta_blocks;
      hipLaunchKernelGGL(k_outputs, dim3(gx, cdiv(nk, BY) + 1), b2, 0, s, g, p.f[0], p.f[1], p.fl, p.fv, p.out,
                         p.ctl);
      break;
    }
    case K_NUCLEATION: {
      const int nxn = g.nxt - g.nabc - (g.nabc - 1) + 1;
      const int nzn = g.nzt - g.nfs - (g.nabc - 1) + 1;
      hipLaunchKernelGGL(k_nucleation, dim3(cdiv(nxn, BX), cdiv(nzn, BY)), b2, 0, s, g, p.c, p.fl, p.fv, p.ctl);
      break;
    }
    case K_ADVANCE:
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, s, p.ctl);
      break;
    case K_FUSED:
      return launch_fused(p, s);
    case K_STR_TOP: {
      // planes max(nabc+1, ks-1) .. ks, ks = nzt - nfs: the ones the fused
      // sweep does not store under a free surface
      if (nyi <= 0 || nk <= 0) return hipSuccess;
      const int gbz0 = nk >= 2 ? nk - 2 : 0;
      hipLaunchKernelGGL(k_str_top, dim3(cdiv(ni, BX), cdiv(nyi, BY), nk - gbz0), b2, 0, s, g, p.c, p.f[0], p.f[1],
                         p.ctl, gbz0);
      break;
    }
    case K_RATES:
      if (!p.rates.on) return hipSuccess;
      hipLaunchKernelGGL(k_rates, dim3(1 + cdiv(p.rates.NL * p.rates.NW, 1024)), dim3(1024), 0, s, g, p.c, p.f[0],
                         p.f[1], p.out, p.rates, p.ctl);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace fd3d
