// Single-problem launches of the tiled interior sweeps (tiles.h).  Blocks
// are numbered so that each XCD (block id modulo 8 on MI355X) receives a
// contiguous run of the logical tile order, keeping x and z neighbour
// tiles on the same L2.
#include "stages.h"
#include "tiles.h"

namespace fd3d {
namespace {

template <int TX, int TZ>
__global__ void __launch_bounds__(TX * TZ) k_vel_tiled(Dims g, Coef c, Fields f0, Fields f1, const Ctl *ctl, int gx,
                                                       int gz, int gy, int NY) {
  __shared__ float lds[Tile<TX, TZ>::LDS_FLOATS];
  const int lid = xcd_remap(blockIdx.x, gx * gz * gy);
  const FieldPair FP = field_pair(f0, f1, ctl);
  vel_tiled_body<TX, TZ>(g, c, FP.fo, FP.fn, lds, lid % gx, (lid / gx) % gz, lid / (gx * gz), threadIdx.x, threadIdx.y, NY);
}

template <int TX, int TZ>
__global__ void __launch_bounds__(TX * TZ) k_str_tiled(Dims g, Coef c, Fields f0, Fields f1, const Ctl *ctl, int gx,
                                                       int gz, int gy, int NY) {
  __shared__ float lds[2 * 3 * Tile<TX, TZ>::N];
  const int lid = xcd_remap(blockIdx.x, gx * gz * gy);
  const FieldPair FP = field_pair(f0, f1, ctl);
  str_tiled_body<TX, TZ>(g, c, FP.fo, FP.fn, lds, lid % gx, (lid / gx) % gz, lid / (gx * gz), threadIdx.x, threadIdx.y, NY);
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

template <int TX, int TZ>
hipError_t launch_vel(const Problem &p, hipStream_t s, int ny) {
  const Dims &g = p.g;
  const int gx = cdiv(g.nxt - 2 * g.nabc, TX), gz = cdiv(g.nzt - g.nfs - g.nabc, TZ);
  const int gy = cdiv(g.nyt - 2 - g.nabc, ny);
  hipLaunchKernelGGL((k_vel_tiled<TX, TZ>), dim3(gx * gz * gy), dim3(TX, TZ), 0, s, g, p.c, p.f[0], p.f[1], p.ctl, gx, gz,
                     gy, ny);
  return hipGetLastError();
}

template <int TX, int TZ>
hipError_t launch_str(const Problem &p, hipStream_t s, int ny) {
  const Dims &g = p.g;
  const int gx = cdiv(g.nxt - 2 * g.nabc, TX), gz = cdiv(g.nzt - g.nfs - g.nabc, TZ);
  const int gy = cdiv(g.nyt - 2 - g.nabc, ny);
  hipLaunchKernelGGL((k_str_tiled<TX, TZ>), dim3(gx * gz * gy), dim3(TX, TZ), 0, s, g, p.c, p.f[0], p.f[1], p.ctl, gx, gz,
                     gy, ny);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_interior(bool stress, const Problem &p, hipStream_t s) {
  if (p.g.nyt - 2 - p.g.nabc <= 0) return hipSuccess;
  const Tuning &t = p.tune;
  const int tz = stress ? t.tz_str : t.tz_vel;
  const int ny = stress ? t.ny_str : t.ny_vel;
  if (ny < 1) return hipErrorInvalidValue;
  const bool wide = (t.mode & (stress ? 16 : 8)) != 0;
  switch (wide ? -tz : tz) {
    case 4: return stress ? launch_str<64, 4>(p, s, ny) : launch_vel<64, 4>(p, s, ny);
    case 8: return stress ? launch_str<64, 8>(p, s, ny) : launch_vel<64, 8>(p, s, ny);
    case 16: return stress ? launch_str<64, 16>(p, s, ny) : launch_vel<64, 16>(p, s, ny);
    case -4: return stress ? launch_str<128, 4>(p, s, ny) : launch_vel<128, 4>(p, s, ny);
    case -8: return stress ? launch_str<128, 8>(p, s, ny) : launch_vel<128, 8>(p, s, ny);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fd3d
