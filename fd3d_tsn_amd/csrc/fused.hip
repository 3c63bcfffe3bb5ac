// Fused interior velocity + stress sweep of one problem (body in fused.h).
#include "fused.h"

namespace fd3d {
namespace {

using namespace fused;

__global__ void __launch_bounds__(NT) k_vs_fused(Dims g, Coef c, Fields f0, Fields f1, const Ctl *ctl, int gx,
                                                       int gz, int gy, int NY, int strip) {
  __shared__ FusedLds L;
  const int lid = xcd_remap(blockIdx.x, gx * gz * gy);
  const int kmax_s = g.nzt - g.nfs - (g.fspace ? 0 : 2);
  const int per = gx * gz, by = lid / per;
  int bx, bz;
  strip_tile(lid - by * per, gx, gz, strip, bx, bz);
  fused_body(g, c, f0, f1, ctl, bx, bz, by, NY, kmax_s, L);
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
                     ny, strip_of_mode(p.tune.mode));
  return hipGetLastError();
}

}  // namespace fd3d
