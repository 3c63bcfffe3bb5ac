// Fused interior velocity + stress sweep of one problem (body in fused.h).
#include "fused.h"

namespace fd3d {
namespace {

using namespace fused;

// Tile order within a y chunk: strips kStrip tiles wide in x, each walked
// in z, so the 32 blocks an XCD runs at once (xcd_remap gives each XCD a
// contiguous range) cover 8 z-tiles of 4 x-tiles instead of 2 z-rows of
// 16 x-tiles.  Measured on the 1024x512x512 sweep (5 processes each): 4.60
// ms against 4.79 ms for the x-fastest order, though the L2 then fetches
// slightly more (1.42 against 1.36 times the algorithmic reads); equal on
// 512x256x256.  Strips 2 tiles wide: 4.53 ms there but 8% slower on
// 512x256x256; z-fastest order (1 tile wide): 5.93 ms.
constexpr int kStrip = 4;

__global__ void __launch_bounds__(NT) k_vs_fused(Dims g, Coef c, Fields f0, Fields f1, const Ctl *ctl, int gx,
                                                       int gz, int gy, int NY) {
  __shared__ FusedLds L;
  const int lid = xcd_remap(blockIdx.x, gx * gz * gy);
  const int kmax_s = g.nzt - g.nfs - (g.fspace ? 0 : 2);
  const int per = gx * gz, by = lid / per, r = lid - by * per;
  const int st = r / (kStrip * gz), w = min(kStrip, gx - st * kStrip), q = r - st * kStrip * gz;
  fused_body(g, c, f0, f1, ctl, st * kStrip + q % w, q / w, by, NY, kmax_s, L);
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

hipError_t launch_fused(const Problem &p, hipStream_t s) {
  const Dims &g = p.g;
  if (g.nyt - 2 - g.nabc <= 0) return hipSuccess;
  const int ny = p.tune.ny_vel;
  if (ny < 1) return hipErrorInvalidValue;
  const int gx = cdiv(g.nxt - 2 * g.nabc, LX), gz = cdiv(g.nzt - g.nfs - g.nabc, ZC);
  const int gy = cdiv(g.nyt - 2 - g.nabc, ny);
  hipLaunchKernelGGL(k_vs_fused, dim3(gx * gz * gy), dim3(LX, NWV), 0, s, g, p.c, p.f[0], p.f[1], p.ctl, gx, gz, gy,
                     ny);
  return hipGetLastError();
}

}  // namespace fd3d
