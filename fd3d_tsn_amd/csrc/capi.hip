// Host side of the C ABI declared in include/fd3d_tsn.h.
//
// One handle = one rupture problem resident in HBM + one HIP stream.
// Responsibilities: device allocation (fd3d_create), uploads with the
// padded layout, the fd3d() prologue (zeroing, PML profiles, fault
// parameter staggering), the per-step kernel sequence, the device output
// ring and its transfer to the caller, per-kernel event timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "handle.h"

using namespace fd3d;

namespace {

int fail(fd3d_handle *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((h), e_ == hipErrorOutOfMemory ? FD3D_E_NOMEM : FD3D_E_HIP,        \
                  std::string(#call) + ": " + hipGetErrorString(e_));                \
  } while (0)

Dims make_dims(const fd3d_config &c) {
  Dims g{};
  g.nxt = c.nxt; g.nyt = c.nyt; g.nzt = c.nzt; g.nabc = c.nabc;
  g.fspace = c.fspace ? 1 : 0;
  g.nfs = g.fspace ? c.nabc : 2;  // fd3d_init.f90:161-165
  // row offset so that i = nabc+1 starts a 64-float (256 B) boundary
  g.off0 = (64 - (c.nabc % 64)) % 64;
  g.px = ((g.off0 + c.nxt + 63) / 64) * 64;
  g.sy = g.px;
  g.sz = g.px * c.nyt;
  g.n3 = (long long)g.sz * c.nzt;
  g.nxT = c.nxt - 2 * c.nabc;
  g.nzT = c.nzt - g.nfs - c.nabc;
  return g;
}

// Default launch shape: tall x-z tiles where the grid allows, and enough
// y chunks for about 2048 blocks (8 per CU) on large grids.  Grids with at
// least kFusedTileRows tile-rows of the fused sweep (its x-z
// tiles of 64 x 11 cells times the interior rows) run the fused interior
// sweep.  Its y chunks minimise waves x (rows + 4), waves = ceil(blocks /
// CUs): each workgroup holds a CU for its whole chunk and re-reads a 4-row
// ring (fused_rows; the autotuner uses the same rule).  Measured
// (interior sweeps, fused against separate): 1024x512x512 4.85 against
// 5.85 ms, 256x256x256 0.36 against 0.45, 128x256x256 0.21 against 0.25;
// even at 256x128x128 (6048 tile-rows); slower below, 192x128x128 0.107
// against 0.086, 128x64x64 0.067 against 0.032.
constexpr long long kFusedTileRows = 8192;

int fused_rows(long long tiles, int nyi, int ncu) {
  long long best = -1;
  int brows = nyi;
  for (int fgy = 1; fgy <= 32; ++fgy) {
    const int rows = (nyi + fgy - 1) / fgy;
    const long long chunks = (nyi + rows - 1) / rows;
    const long long cost = ((tiles * chunks + ncu - 1) / ncu) * (rows + 4);
    if (best < 0 || cost < best) { best = cost; brows = rows; }
  }
  return brows;
}

// Strip width of the fused sweep's tile order (mode bits 5-6): the 8 XCDs
// take contiguous ranges of the tile numbers, and the measured optima are
// the widths at which each XCD's range is exactly one strip (1024x512x512:
// 16 x 47 tiles, 94 per XCD, strips of 2, 4.61 against 4.70 ms in strips of
// 4; 512x256x256: 8 x 24 tiles in 4 y chunks, 96 per XCD, strips of 4).
// Other grids keep strips of 4 (the autotuner measures the widths).
int fused_strip_bits(int fgx, int fgz, int fgy) {
  const long long n = (long long)fgx * fgz * fgy;
  if (n % 8 != 0 || (n / 8) % fgz != 0) return 0;
  return n / 8 / fgz == 2 ? 32 : 0;
}

Tuning default_tuning(const Dims &g, int ncu) {
  Tuning t{};
  t.mode = 1;
  const int nx = g.nxt - 2 * g.nabc, nz = g.nzt - g.nfs - g.nabc, ny = g.nyt - 2 - g.nabc;
  const int tz = nz >= 64 ? 16 : (nz >= 16 ? 8 : 4);
  const long long tiles = (long long)((nx + 63) / 64) * ((nz + tz - 1) / tz);
  long long gy = (2048 + tiles - 1) / tiles;
  if (gy < 1) gy = 1;
  int rows = ny > 0 ? (int)((ny + gy - 1) / gy) : 1;
  if (rows < 8) rows = 8;
  t.tz_vel = t.tz_str = tz;
  t.ny_vel = t.ny_str = rows;
  const long long ftiles = (long long)((nx + 63) / 64) * ((nz + 10) / 11);
  if (ny > 0 && ftiles * ny >= kFusedTileRows) {
    t.ny_vel = fused_rows(ftiles, ny, ncu > 0 ? ncu : 256);
    t.mode = 1 | 4 | fused_strip_bits((nx + 63) / 64, (nz + 10) / 11, (ny + t.ny_vel - 1) / t.ny_vel);
  }
  return t;
}

Coef make_coef(const fd3d_config &c) {
  Coef k{};
  k.dh = c.dh; k.dt = c.dt;
  k.dth = c.dt / c.dh;
  k.dthh = (0.5f * c.dt) / c.dh;
  k.dht = c.dh / c.dt;
  k.damp_s = c.damp_s;
  k.c1 = 9.0f / 8.0f;
  k.c2 = -1.0f / 24.0f;
  return k;
}

// PML region table (allocation extents fd3d_deriv.f90:1324-1437).
int build_regions(const fd3d_config &c, const Dims &g, Pml &pm, long long &nsplit, long long &nprof) {
  const int nabc = c.nabc, nxt = c.nxt, nyt = c.nyt, nzt = c.nzt;
  const int kb = nabc + 1, ke = nzt - g.nfs;
  struct Spec { int ib, ie, jb, je_uw, je_v, je_xz, kb, ke, ztop, nzx, nzy; };
  std::vector<Spec> sp = {
      {2, nabc, nabc + 1, nyt, nyt - 1, nyt, kb, ke, 1, nyt - nabc, nyt - nabc - 1},
      {nxt - nabc + 1, nxt - 1, nabc + 1, nyt, nyt - 1, nyt, kb, ke, g.fspace ? 0 : 1, nyt - nabc, nyt - nabc - 1},
      {2, nxt - 1, 2, nabc, nabc, nabc, kb, ke, g.fspace ? 0 : 1, nabc - 1, nabc - 1},
      {2, nxt - 1, 2, nyt, nyt - 1, nyt, 2, nabc, 0, 0, 0},
  };
  if (g.fspace) sp.push_back({2, nxt - 1, 2, nyt, nyt - 1, nyt, nzt - nabc + 1, nzt - 1, 0, 0, 0});
  pm.nreg = (int)sp.size();
  nsplit = 0;
  nprof = 0;
  int blk = 0;
  for (int r = 0; r < pm.nreg; ++r) {
    PmlRegion &R = pm.r[r];
    const Spec &s = sp[r];
    R.ib = s.ib; R.ie = s.ie; R.jb = s.jb; R.je_uw = s.je_uw; R.je_v = s.je_v; R.je_xz = s.je_xz;
    R.kb = s.kb; R.ke = s.ke;
    R.ni = s.ie - s.ib + 1; R.nj = s.je_uw - s.jb + 1; R.nk = s.ke - s.kb + 1;
    if (R.ni <= 0 || R.nj <= 0 || R.nk <= 0) return -1;
    R.ncell = R.ni * R.nj * R.nk;
    if ((long long)kSplits * R.ncell >= (1LL << 30)) return -1;   // 32-bit byte offsets into the split block
    udiv_magic(R.ni, R.mdi, R.sdi);
    udiv_magic(R.nj, R.mdj, R.sdj);
    R.ztop = s.ztop; R.nzx = s.nzx; R.nzy = s.nzy;
    R.blk0 = blk;
    blk += (R.ncell + kPmlBlock - 1) / kPmlBlock;
    R.split_off = nsplit;
    nsplit += (long long)kSplits * R.ncell;
    R.prof_off = (int)nprof;
    nprof += 4LL * (R.ni + 1) + 4LL * (R.nj + 1) + 4LL * (R.nk + 1);
  }
  pm.nblk = blk;
  return 0;
}

// Blocks of the tiled PML sweeps (pml.hip): 64 x kPmlTZ x-z tiles on the
// interior tile grid (origins i = nabc+1 + 64 bx, k = nabc+1 + kPmlTZ bz)
// that hold PML cells, each with the rows of the regions it meets, cut
// into runs of at most kPmlRows rows: (bx, bz, j0, j1).
std::vector<int4> build_pml_tiles(const Dims &g, const Pml &pm) {
  std::vector<int4> t;
  const int o = g.nabc + 1;
  auto fdiv = [](int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); };
  int ilo = g.nxt, ihi = 0, klo = g.nzt, khi = 0;
  for (int r = 0; r < pm.nreg; ++r) {
    ilo = std::min(ilo, pm.r[r].ib); ihi = std::max(ihi, pm.r[r].ie);
    klo = std::min(klo, pm.r[r].kb); khi = std::max(khi, pm.r[r].ke);
  }
  for (int bz = fdiv(klo - o, kPmlTZ); bz <= fdiv(khi - o, kPmlTZ); ++bz) {
    const int k0 = o + kPmlTZ * bz, k1 = k0 + kPmlTZ - 1;
    for (int bx = fdiv(ilo - o, 64); bx <= fdiv(ihi - o, 64); ++bx) {
      const int i0 = o + 64 * bx, i1 = i0 + 63;
      int jlo = g.nyt + 1, jhi = 0;
      for (int r = 0; r < pm.nreg; ++r) {
        const PmlRegion &R = pm.r[r];
        if (std::max(i0, R.ib) <= std::min(i1, R.ie) && std::max(k0, R.kb) <= std::min(k1, R.ke)) {
          jlo = std::min(jlo, R.jb);
          jhi = std::max(jhi, R.je_uw);
        }
      }
      for (int j = jlo; j <= jhi; j += kPmlRows) t.push_back(make_int4(bx, bz, j, std::min(jhi, j + kPmlRows - 1)));
    }
  }
  return t;
}

// init_pml damping profiles (fd3d_deriv.f90:1439-1491), stored as (1-w) and
// (1+w) per profile, region by region, in the order the kernels read them.
std::vector<float> build_profiles(const fd3d_config &c, const Dims &g, const Pml &pm) {
  const int nabc = c.nabc, nxt = c.nxt;
  const float omegaM = (c.pml_fact * c.pml_vp) / ((2.0f * c.dh) * (float)(nabc - 1));
  std::vector<float> om(nabc + 1, 0), omR(nabc + 1, 0), omM(nabc + 1, 0), omRM(nabc + 1, 0);
  const float scale = (0.5f * c.dt) * omegaM;
  for (int i = 1; i <= nabc - 1; ++i) {
    float q = (float)i / (float)(nabc - 1);
    om[i] = scale * (((q * q) * q) * q);
    omR[nabc - i] = om[i];
    float qm = ((float)i - 0.5f) / (float)(nabc - 1);
    omM[i] = scale * (((qm * qm) * qm) * qm);
    omRM[nabc - i] = omM[i];
  }
  long long total = 0;
  for (int r = 0; r < pm.nreg; ++r)
    total += 4LL * (pm.r[r].ni + 1) + 4LL * (pm.r[r].nj + 1) + 4LL * (pm.r[r].nk + 1);
  std::vector<float> out((size_t)total, 0.0f);
  for (int r = 0; r < pm.nreg; ++r) {
    const PmlRegion &R = pm.r[r];
    std::vector<float> ox(R.ni + 1, 0), oxS(R.ni + 1, 0), oy(R.nj + 1, 0), oyS(R.nj + 1, 0),
        oz(R.nk + 1, 0), ozS(R.nk + 1, 0);
    for (int i = 1; i <= nabc - 1; ++i) {
      switch (r) {
        case 0: ox[i] = omR[i]; oxS[i] = omRM[i]; break;
        case 1: ox[i] = omM[i]; oxS[i] = om[i]; break;
        case 2:
          oy[i] = omR[i]; ox[i] = omR[i]; ox[nxt - i - 1] = omM[i];
          oyS[i] = omRM[i]; oxS[i] = omRM[i]; oxS[nxt - i - 1] = om[i];
          break;
        case 3:
          oz[i] = omR[i]; ox[i] = omR[i]; ox[nxt - i - 1] = omM[i]; oy[i] = omR[i];
          ozS[i] = omRM[i]; oxS[i] = omRM[i]; oxS[nxt - i - 1] = om[i]; oyS[i] = omRM[i];
          break;
        default:
          oz[i] = omM[i]; ox[i] = omR[i]; ox[nxt - i - 1] = omM[i]; oy[i] = omR[i];
          ozS[i] = om[i]; oxS[i] = omRM[i]; oxS[nxt - i - 1] = om[i]; oyS[i] = omRM[i];
          break;
      }
    }
    // regions 1-2 read omegay/omegaz where later regions have yS/zS profiles,
    // region 3 reads omegaz (pml_xyz, fd3d_deriv.f90:2398-2455, 2785-2815)
    if (r <= 1) { oyS = oy; ozS = oz; }
    if (r == 2) ozS = oz;
    float *dst = out.data() + R.prof_off;
    auto put = [&](const std::vector<float> &o) {
      for (size_t n = 0; n < o.size(); ++n) *dst++ = 1.0f - o[n];
      for (size_t n = 0; n < o.size(); ++n) *dst++ = 1.0f + o[n];
    };
    put(ox); put(oxS); put(oy); put(oyS); put(oz); put(ozS);
  }
  return out;
}

// Device allocation of n elements.  skew shifts the array's start by that
// many bytes (a multiple of 256) inside a slightly larger allocation: the
// 3D fields are read and written in lock step at equal offsets, and
// staggering their starts spreads those streams over the HBM channels.
template <typename T>
int dalloc(fd3d_handle *h, T **ptr, size_t n, size_t skew = 0) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T) + skew);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, e == hipErrorOutOfMemory ? FD3D_E_NOMEM : FD3D_E_HIP,
                std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  h->allocs.push_back(p);
  h->bytes += n * sizeof(T) + skew;
  *ptr = reinterpret_cast<T *>(static_cast<char *>(p) + skew);
  return 0;
}

// Start offset of the q-th 3D field: odd multiples of 256 bytes modulo 64 KiB.
// The unit ($FD3D_FIELD_SKEW, bytes) is rounded down to whole 256-byte
// units: the sweeps' row segments and 16-byte loads rely on that alignment.
size_t field_skew(int q) {
  static const char *env = std::getenv("FD3D_FIELD_SKEW");
  const size_t unit = env ? ((size_t)std::max(0L, std::atol(env)) / 256) * 256 : 2304;
  return (unit * (size_t)q) % 65536;
}

int validate(const fd3d_config *c) {
  if (!c) return FD3D_E_INVALID;
  if (c->nabc < 3) return FD3D_E_INVALID;
  const int nfs = c->fspace ? c->nabc : 2;
  if (c->nxt < 2 * c->nabc + 4 || c->nyt < c->nabc + 4 || c->nzt < c->nabc + nfs + 4) return FD3D_E_INVALID;
  if (!(c->dh > 0) || !(c->dt > 0)) return FD3D_E_INVALID;
  if (c->friction != FD3D_SW && c->friction != FD3D_FVW) return FD3D_E_INVALID;
  if (c->nstations < 0 || (c->nstations > 0 && !c->stations)) return FD3D_E_INVALID;
  for (int s = 0; s < c->nstations; ++s) {
    const int *t = c->stations + 3 * s;
    if (t[0] < 1 || t[0] > c->nxt || t[1] < 1 || t[1] > c->nyt || t[2] < 1 || t[2] > c->nzt)
      return FD3D_E_INVALID;
  }
  Dims g = make_dims(*c);
  if (g.n3 >= (1LL << 30)) return FD3D_E_INVALID;  // 32-bit byte offsets into each 3D array
  return 0;
}

size_t footprint(const fd3d_config &c) {
  Dims g = make_dims(c);
  Pml pm{};
  long long nsplit = 0, nprof = 0;
  if (build_regions(c, g, pm, nsplit, nprof) != 0) return 0;
  const size_t n2 = (size_t)c.nxt * c.nzt;
  const int ring = c.ring_steps > 0 ? c.ring_steps : 64;
  size_t b = 18 * (size_t)g.n3 * 4;                 // 3D fields: two wavefield sets
  b += 3 * (size_t)(c.nzt + 2) * 4;                 // medium: one value per layer (z plane)
  b += (size_t)nsplit * 4 + (size_t)nprof * 4;      // PML
  b += n2 * 4 * 33 + n2 * 8 * 8;                    // fault plane
  b += (size_t)ring * FD3D_NOUT * g.nxT * g.nzT * 4 + (size_t)ring * c.nstations * 3 * 4;
  return b;
}

void record(fd3d_handle *h, KernelId id, bool start, hipEvent_t &ev) {
  (void)id;
  (void)start;
  if (h->pool.empty()) {
    (void)hipEventCreate(&ev);
  } else {
    ev = h->pool.back();
    h->pool.pop_back();
  }
  (void)hipEventRecord(ev, h->stream);
}

int launch_one(fd3d_handle *h, KernelId id) {
  hipEvent_t a = nullptr, b = nullptr;
  if (h->prof) record(h, id, true, a);
  hipError_t e = launch(id, h->p, h->stream);
  if (e != hipSuccess) return fail(h, FD3D_E_HIP, std::string("launch ") + kKernelNames[id] + ": " + hipGetErrorString(e));
  if (h->prof) {
    record(h, id, false, b);
    h->evs.push_back({(int)id, a, b});
  }
  return 0;
}

void harvest(fd3d_handle *h) {
  for (auto &e : h->evs) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      h->kms[e.id] += ms;
      h->kcount[e.id] += 1;
    }
    h->pool.push_back(e.a);
    h->pool.push_back(e.b);
  }
  h->evs.clear();
}

// Wait for the work queued on h: its own stream, and the stream of the
// batch that advances it.
hipError_t owner_sync(fd3d_handle *h) {
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && h->batch) e = hipStreamSynchronize(fd3d_batch_stream_of(h->batch));
  return e;
}

int check_ctl(fd3d_handle *h) { return fd3d_handle_check_ctl(h, h->stream); }

// Kernel sequence of one step (fd3d_theo.f90:208-542).  The velocity
// kernels read only the old wavefield set and the stress kernels only the
// new velocities and old stresses, so vel_fault and pml_vel may run before
// the fused interior sweep, which needs their new velocities.
void drop_graph(fd3d_handle *h) {
  for (int n = 0; n < 3; ++n) {
    if (h->gexec[n]) (void)hipGraphExecDestroy(h->gexec[n]);
    if (h->graph[n]) (void)hipGraphDestroy(h->graph[n]);
    h->gexec[n] = nullptr;
    h->graph[n] = nullptr;
  }
}

void build_seq(fd3d_handle *h) {
  drop_graph(h);
  const Problem &p = h->p;
  const bool fused = p.tune.mode & 4;
  h->seq.clear();
  if (fused) {
    // under a free surface the ghost velocities come after the fused sweep
    // and before the stresses of the top two planes (str_top)
    h->seq = {K_VEL_FAULT, K_PML_VEL, K_FUSED};
    if (!p.g.fspace) { h->seq.push_back(K_FUVW_UV); h->seq.push_back(K_FUVW_W); h->seq.push_back(K_STR_TOP); }
  } else {
    h->seq = {K_VEL_INTERIOR, K_VEL_FAULT, K_PML_VEL};
    if (!p.g.fspace) { h->seq.push_back(K_FUVW_UV); h->seq.push_back(K_FUVW_W); }
    h->seq.push_back(K_STR_INTERIOR);
  }
  h->seq.push_back(K_STR_FAULT); h->seq.push_back(K_PML_STR);
  if (!p.g.fspace) h->seq.push_back(K_FRES);
  h->seq.push_back(K_ADVANCE);
  h->seq.push_back(K_TRACTION);
  if (!p.g.fspace) h->seq.push_back(K_TRACTION_MIRROR);
  h->seq.push_back(K_TABS); h->seq.push_back(K_FRICTION); h->seq.push_back(K_OUTPUTS);
  if (p.rates.on) h->seq.push_back(K_RATES);
  if (p.fv.nucleation) h->seq.push_back(K_NUCLEATION);
}

}  // namespace

int fd3d_handle_check_ctl(fd3d_handle *h, hipStream_t s) {
  Ctl c{};
  HIPCHK(h, hipMemcpyAsync(&c, h->p.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  if (c.newton_capped > 0)
    return fail(h, FD3D_E_NEWTON, "rate-and-state Newton solve hit the iteration cap at " +
                                      std::to_string(c.newton_capped) + " node-steps");
  if (c.nonfinite) return fail(h, FD3D_E_NONFINITE, "non-finite slip rate in the fault outputs");
  return 0;
}

extern "C" {

size_t fd3d_footprint(const fd3d_config *cfg) {
  if (validate(cfg) != 0) return 0;
  return footprint(*cfg);
}

int fd3d_device_memory(size_t *free_bytes, size_t *total_bytes) {
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return FD3D_E_HIP;
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return 0;
}

int fd3d_create(const fd3d_config *cfg, fd3d_handle **out) {
  if (!out) return FD3D_E_INVALID;
  *out = nullptr;
  int v = validate(cfg);
  if (v != 0) return v;
  fd3d_handle *h = new fd3d_handle();
  h->cfg = *cfg;
  if (cfg->nstations > 0) h->sta.assign(cfg->stations, cfg->stations + 3 * cfg->nstations);
  h->cfg.stations = h->sta.data();
  if (h->cfg.ring_steps <= 0) h->cfg.ring_steps = 64;
  if (h->cfg.newton_max <= 0) h->cfg.newton_max = 200;
  Problem &p = h->p;
  p.g = make_dims(*cfg);
  p.c = make_coef(*cfg);
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      ncu = 0;
    (void)hipGetLastError();
    p.tune = default_tuning(p.g, ncu);
  }
  long long nsplit = 0, nprof = 0;
  if (build_regions(*cfg, p.g, p.pml, nsplit, nprof) != 0) {
    delete h;
    return FD3D_E_INVALID;
  }
  int rc = 0;
  auto bail = [&](int code) {
    fd3d_destroy(h);
    return code;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return bail(FD3D_E_HIP);
  if (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->join, hipEventDisableTiming) != hipSuccess)
    return bail(FD3D_E_HIP);
  const size_t n3 = (size_t)p.g.n3, n2 = (size_t)cfg->nxt * cfg->nzt;
  // two wavefield sets (the step reads one and writes the other)
  Fields &A = p.f[0], &B = p.f[1];
  float **f3[] = {&A.u, &A.v, &A.w, &A.xx, &A.yy, &A.zz, &A.xy, &A.yz, &A.xz,
                  &B.u, &B.v, &B.w, &B.xx, &B.yy, &B.zz, &B.xy, &B.yz, &B.xz};
  for (int q = 0; q < (int)(sizeof(f3) / sizeof(f3[0])); ++q)
    if ((rc = dalloc(h, f3[q], n3, field_skew(q)))) return bail(rc);
  // the depth-layered medium: lam, mu, rho of plane k at index k (1..nzt;
  // entries 0 and nzt+1 stay zero)
  const size_t nlay = (size_t)cfg->nzt + 2;
  for (float **pp : {&A.lam, &A.mu, &A.d}) {
    if ((rc = dalloc(h, pp, nlay))) return bail(rc);
    if (hipMemsetAsync(*pp, 0, nlay * sizeof(float), h->stream) != hipSuccess) return bail(FD3D_E_HIP);
  }
  B.lam = A.lam; B.mu = A.mu; B.d = A.d;
  Fault &fl = p.fl;
  float **f2[] = {&fl.tx, &fl.tz, &fl.v1t, &fl.avdx, &fl.avdz, &fl.RFx, &fl.RFz, &fl.au1, &fl.av1,
                  &fl.aw1, &fl.uZ, &fl.wX, &fl.tabsX, &fl.tabsZ, &fl.T0X, &fl.T0Z, &fl.Dc, &fl.coh,
                  &fl.peakX, &fl.dynX, &fl.DcX, &fl.peakZ, &fl.dynZ, &fl.DcZ, &fl.ruptime, &fl.rise,
                  &fl.sliptime, &fl.slipX, &fl.slipZ, &fl.schangeX, &fl.schangeZ, &fl.distX, &fl.distZ};
  for (auto pp : f2)
    if ((rc = dalloc(h, pp, n2))) return bail(rc);
  double **d2[] = {&fl.aX, &fl.bX, &fl.psiX, &fl.vwX, &fl.aZ, &fl.bZ, &fl.psiZ, &fl.vwZ};
  for (auto pp : d2)
    if ((rc = dalloc(h, pp, n2))) return bail(rc);
  if ((rc = dalloc(h, &p.pml.splits, (size_t)nsplit))) return bail(rc);
  for (int r = 0; r < p.pml.nreg; ++r) p.pml.r[r].sp = p.pml.splits + p.pml.r[r].split_off;
  {
    const std::vector<int4> tiles = build_pml_tiles(p.g, p.pml);
    p.pml.ntiles = (int)tiles.size();
    if ((rc = dalloc(h, &p.pml.tiles, std::max<size_t>(tiles.size(), 1)))) return bail(rc);
    if (!tiles.empty() && hipMemcpy(p.pml.tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice) !=
                              hipSuccess)
      return bail(FD3D_E_HIP);
    PmlRegion *rd = nullptr;
    if ((rc = dalloc(h, &rd, (size_t)kMaxRegions))) return bail(rc);
    if (hipMemcpy(rd, p.pml.r, sizeof(p.pml.r), hipMemcpyHostToDevice) != hipSuccess) return bail(FD3D_E_HIP);
    p.pml.rdev = rd;
  }
  if ((rc = dalloc(h, &p.pml.prof, (size_t)nprof))) return bail(rc);
  const size_t slab = (size_t)p.g.nxT * p.g.nzT;
  if ((rc = dalloc(h, &p.out.ring, (size_t)h->cfg.ring_steps * FD3D_NOUT * slab))) return bail(rc);
  if ((rc = dalloc(h, &p.out.seis, (size_t)h->cfg.ring_steps * std::max(1, cfg->nstations) * 3))) return bail(rc);
  if ((rc = dalloc(h, &h->d_sta, (size_t)std::max(1, cfg->nstations) * 3))) return bail(rc);
  if ((rc = dalloc(h, &p.ctl, 1))) return bail(rc);
  if (cfg->nstations > 0 &&
      hipMemcpy(h->d_sta, h->sta.data(), h->sta.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return bail(FD3D_E_HIP);
  p.out.sta = h->d_sta;
  p.out.nsta = cfg->nstations;
  p.out.ring_steps = h->cfg.ring_steps;
  // zero everything once (pads included); the prologue zeroes state again
  for (auto pp : f3)
    if (hipMemsetAsync(*pp, 0, n3 * sizeof(float), h->stream) != hipSuccess) return bail(FD3D_E_HIP);
  std::vector<float> prof = build_profiles(*cfg, p.g, p.pml);
  if (hipMemcpy(p.pml.prof, prof.data(), prof.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return bail(FD3D_E_HIP);
  FvwScalars &fv = p.fv;
  fv.Sn = cfg->Sn; fv.uini = cfg->uini; fv.wini = cfg->wini; fv.f0 = cfg->f0; fv.v0 = cfg->v0;
  fv.fw = cfg->fw; fv.hx0 = cfg->hx0; fv.hz0 = cfg->hz0; fv.perturb = cfg->perturb; fv.RR2 = cfg->RR2;
  fv.TT2 = cfg->TT2; fv.strinixI = cfg->strinixI; fv.newton_max = h->cfg.newton_max;
  fv.fvw = cfg->friction == FD3D_FVW; fv.nucleation = fv.fvw && cfg->nucleation;
  build_seq(h);
  if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(FD3D_E_HIP);
  *out = h;
  return 0;
}

// 3D fields cross the boundary in the reference's dense layout; on the
// device every x row is padded to px.  The host packs / unpacks the rows
// and the transfer is one contiguous copy (a strided 2D copy from pageable
// memory moves one row at a time).
static int upload3(fd3d_handle *h, float *dst, const float *src) {
  const Dims &g = h->p.g;
  std::vector<float> &buf = h->staging;
  buf.assign((size_t)g.n3, 0.0f);
  const size_t rows = (size_t)g.nyt * g.nzt;
  for (size_t r = 0; r < rows; ++r)
    std::memcpy(buf.data() + g.off0 + r * g.px, src + r * g.nxt, (size_t)g.nxt * sizeof(float));
  HIPCHK(h, hipMemcpy(dst, buf.data(), (size_t)g.n3 * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

static int download3(fd3d_handle *h, float *dst, const float *src) {
  const Dims &g = h->p.g;
  HIPCHK(h, owner_sync(h));
  std::vector<float> &buf = h->staging;
  buf.resize((size_t)g.n3);
  HIPCHK(h, hipMemcpy(buf.data(), src, (size_t)g.n3 * sizeof(float), hipMemcpyDeviceToHost));
  const size_t rows = (size_t)g.nyt * g.nzt;
  for (size_t r = 0; r < rows; ++r)
    std::memcpy(dst + r * g.nxt, buf.data() + g.off0 + r * g.px, (size_t)g.nxt * sizeof(float));
  return 0;
}

// The reference's medium is depth-layered: readcrustalmodel fills whole
// x-y planes, lam1(1:nxt,1:nyt,k) = ... (fd3d_init.f90:240-265).  The
// device keeps one value per plane, so the sweeps read no medium arrays.
int fd3d_upload_medium(fd3d_handle *h, const float *lam1, const float *mu1, const float *d1) {
  if (!h || !lam1 || !mu1 || !d1) return fail(h, FD3D_E_INVALID, "null argument");
  if (h->batch) return fail(h, FD3D_E_STATE, "remove the problem from its batch first");
  HIPCHK(h, owner_sync(h));
  const int nzt = h->cfg.nzt;
  const size_t plane = (size_t)h->cfg.nxt * h->cfg.nyt;
  std::vector<float> lay(3 * ((size_t)nzt + 2), 0.0f);
  float *L = lay.data(), *M = L + nzt + 2, *D = M + nzt + 2;
  // CFL check on the largest P velocity (fd3d_init.f90:274-278)
  float vmax = 0.0f;
  for (int k = 1; k <= nzt; ++k) {
    const size_t q0 = (size_t)(k - 1) * plane;
    const float l = lam1[q0], m = mu1[q0], d = d1[q0];
    for (size_t q = q0; q < q0 + plane; ++q)
      if (lam1[q] != l || mu1[q] != m || d1[q] != d)
        return fail(h, FD3D_E_INVALID, "the medium must be constant over each z plane (depth-layered, as "
                                        "fd3d_init builds it from crustal.dat); plane k = " + std::to_string(k) +
                                        " varies");
    if (!(d > 0.0f)) return fail(h, FD3D_E_INVALID, "density must be positive");
    const float vp = std::sqrt((l + 2.0f * m) / d);
    if (vp > vmax) vmax = vp;
    L[k] = l; M[k] = m; D[k] = d;
  }
  const float cfl = vmax * h->cfg.dt / h->cfg.dh;
  if (cfl > 0.25f) return fail(h, FD3D_E_CFL, "numerically unstable: CFL = " + std::to_string(cfl));
  const Fields &f = h->p.f[0];
  HIPCHK(h, hipMemcpy(f.lam, L, ((size_t)nzt + 2) * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(f.mu, M, ((size_t)nzt + 2) * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(f.d, D, ((size_t)nzt + 2) * sizeof(float), hipMemcpyHostToDevice));
  h->medium_ok = true;
  return 0;
}

int fd3d_upload_fault_sw(fd3d_handle *h, const float *striniX, const float *striniZ, const float *peak_xz,
                         const float *dyn_xz, const float *Dc, const float *coh) {
  if (!h || !striniX || !striniZ || !peak_xz || !dyn_xz || !Dc || !coh) return fail(h, FD3D_E_INVALID, "null argument");
  if (h->cfg.friction != FD3D_SW) return fail(h, FD3D_E_INVALID, "handle is rate-and-state");
  const size_t n2 = (size_t)h->cfg.nxt * h->cfg.nzt;
  h->h_striniX.assign(striniX, striniX + n2);
  h->h_striniZ.assign(striniZ, striniZ + n2);
  h->h_peak.assign(peak_xz, peak_xz + n2);
  h->h_dyn.assign(dyn_xz, dyn_xz + n2);
  h->h_Dc.assign(Dc, Dc + n2);
  h->h_coh.assign(coh, coh + n2);
  h->fault_ok = true;
  return 0;
}

int fd3d_upload_fault_fvw(fd3d_handle *h, const double *Dc, const double *striniX, const double *striniZ,
                          const double *a, const double *b, const double *psi, const double *vw,
                          const double *aX, const double *bX, const double *psiX, const double *vwX) {
  if (!h || !Dc || !striniX || !striniZ || !a || !b || !psi || !vw || !aX || !bX || !psiX || !vwX)
    return fail(h, FD3D_E_INVALID, "null argument");
  if (h->cfg.friction != FD3D_FVW) return fail(h, FD3D_E_INVALID, "handle is slip weakening");
  const size_t n2 = (size_t)h->cfg.nxt * h->cfg.nzt;
  auto tof = [&](const double *s) { std::vector<float> o(n2); for (size_t q = 0; q < n2; ++q) o[q] = (float)s[q]; return o; };
  h->h_Dc = tof(Dc);
  h->h_striniX = tof(striniX);
  h->h_striniZ = tof(striniZ);
  h->h_coh.assign(n2, 0.0f);
  h->h_a.assign(a, a + n2); h->h_b.assign(b, b + n2); h->h_psi.assign(psi, psi + n2); h->h_vw.assign(vw, vw + n2);
  h->h_aX.assign(aX, aX + n2); h->h_bX.assign(bX, bX + n2); h->h_psiX.assign(psiX, psiX + n2);
  h->h_vwX.assign(vwX, vwX + n2);
  h->fault_ok = true;
  return 0;
}

int fd3d_init_state(fd3d_handle *h) {
  if (!h) return FD3D_E_INVALID;
  if (!h->medium_ok || !h->fault_ok) return fail(h, FD3D_E_STATE, "upload the medium and the fault first");
  if (h->batch) return fail(h, FD3D_E_STATE, "remove the problem from its batch first");
  if (h->begun != h->waited) return fail(h, FD3D_E_STATE, "output chunks in flight: fd3d_step_wait first");
  Problem &p = h->p;
  const Dims &g = p.g;
  const size_t n3 = (size_t)g.n3;
  const int nxt = h->cfg.nxt, nzt = h->cfg.nzt;
  const size_t n2 = (size_t)nxt * nzt;
  hipStream_t s = h->stream;
  for (const Fields &F : p.f) {
    float *f3[] = {F.u, F.v, F.w, F.xx, F.yy, F.zz, F.xy, F.yz, F.xz};
    for (auto a : f3) HIPCHK(h, hipMemsetAsync(a, 0, n3 * sizeof(float), s));
  }
  Fault &fl = p.fl;
  float *z2[] = {fl.tx, fl.tz, fl.v1t, fl.avdx, fl.avdz, fl.RFx, fl.RFz, fl.au1, fl.av1, fl.aw1, fl.uZ,
                 fl.wX, fl.rise, fl.slipX, fl.slipZ, fl.schangeX, fl.schangeZ, fl.distX, fl.distZ};
  for (auto a : z2) HIPCHK(h, hipMemsetAsync(a, 0, n2 * sizeof(float), s));
  HIPCHK(h, fill_f32(fl.ruptime, 1.e4f, (long long)n2, s));
  HIPCHK(h, fill_f32(fl.sliptime, 1.e4f, (long long)n2, s));
  const float tabs0 = p.fv.fvw ? 1.0f : 0.0f;  // fd3d_theo.f90:72,79-80
  HIPCHK(h, fill_f32(fl.tabsX, tabs0, (long long)n2, s));
  HIPCHK(h, fill_f32(fl.tabsZ, tabs0, (long long)n2, s));
  long long nsplit = 0;
  for (int r = 0; r < p.pml.nreg; ++r) nsplit += (long long)kSplits * p.pml.r[r].ncell;
  HIPCHK(h, hipMemsetAsync(p.pml.splits, 0, (size_t)nsplit * sizeof(float), s));
  HIPCHK(h, hipMemsetAsync(p.ctl, 0, sizeof(Ctl), s));
  if (p.rates.on) {   // MSRX = MSRZ = MomentRate = 0 (fd3d_theo.f90:67)
    const size_t nb = (size_t)p.rates.NL * p.rates.NW * p.rates.nbins;
    if (nb) {
      HIPCHK(h, hipMemsetAsync(p.rates.msrx, 0, nb * sizeof(float), s));
      HIPCHK(h, hipMemsetAsync(p.rates.msrz, 0, nb * sizeof(float), s));
    }
    HIPCHK(h, hipMemsetAsync(p.rates.mr, 0, (size_t)p.rates.nbins * sizeof(float), s));
  }
  // interp_fric (fd3d_deriv.f90:3150-3190) on the host, float arithmetic
  std::vector<float> peakX(n2, 0), dynX(n2, 0), DcX(n2, 0), T0X(n2, 0);
  auto I = [&](int i, int k) { return (size_t)(i - 1) + (size_t)nxt * (size_t)(k - 1); };
  const bool fvw = p.fv.fvw;
  for (int k = 2; k <= nzt - 1; ++k)
    for (int i = 2; i <= nxt - 1; ++i) {
      size_t c = I(i, k), im = I(i - 1, k), km = I(i, k - 1), ikm = I(i - 1, k - 1);
      if (!fvw) {
        peakX[c] = (((h->h_peak[c] + h->h_peak[im]) + h->h_peak[km]) + h->h_peak[ikm]) / 4.0f;
        dynX[c] = (((h->h_dyn[c] + h->h_dyn[im]) + h->h_dyn[km]) + h->h_dyn[ikm]) / 4.0f;
        DcX[c] = (((h->h_Dc[c] + h->h_Dc[im]) + h->h_Dc[km]) + h->h_Dc[ikm]) / 4.0f;
      }
      T0X[c] = (((h->h_striniX[c] + h->h_striniX[im]) + h->h_striniX[km]) + h->h_striniX[ikm]) / 4.0f;
    }
  auto up = [&](void *dst, const void *src, size_t bytes) -> int {
    HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return 0;
  };
  int rc;
  const size_t fb = n2 * sizeof(float), db = n2 * sizeof(double);
  std::vector<float> zeros(n2, 0.0f);
  if ((rc = up(fl.T0X, T0X.data(), fb))) return rc;
  if ((rc = up(fl.T0Z, h->h_striniZ.data(), fb))) return rc;
  if ((rc = up(fl.Dc, h->h_Dc.data(), fb))) return rc;
  if ((rc = up(fl.coh, h->h_coh.data(), fb))) return rc;
  if (!fvw) {
    if ((rc = up(fl.peakX, peakX.data(), fb))) return rc;
    if ((rc = up(fl.dynX, dynX.data(), fb))) return rc;
    if ((rc = up(fl.DcX, DcX.data(), fb))) return rc;
    if ((rc = up(fl.peakZ, h->h_peak.data(), fb))) return rc;
    if ((rc = up(fl.dynZ, h->h_dyn.data(), fb))) return rc;
    if ((rc = up(fl.DcZ, h->h_Dc.data(), fb))) return rc;
  } else {
    if ((rc = up(fl.aZ, h->h_a.data(), db))) return rc;
    if ((rc = up(fl.bZ, h->h_b.data(), db))) return rc;
    if ((rc = up(fl.psiZ, h->h_psi.data(), db))) return rc;
    if ((rc = up(fl.vwZ, h->h_vw.data(), db))) return rc;
    if ((rc = up(fl.aX, h->h_aX.data(), db))) return rc;
    if ((rc = up(fl.bX, h->h_bX.data(), db))) return rc;
    if ((rc = up(fl.psiX, h->h_psiX.data(), db))) return rc;
    if ((rc = up(fl.vwX, h->h_vwX.data(), db))) return rc;
  }
  HIPCHK(h, hipStreamSynchronize(s));
  h->steps = 0;
  h->inited = true;
  return 0;
}

// Kernels of one half-step that read only the old set (velocities) or
// only the new velocities and old stresses (stresses) and write disjoint
// cells: interior, fault-plane and PML updates.
static int half_step_group(KernelId id, bool small) {
  switch (id) {
    case K_VEL_INTERIOR: return small ? 1 : 0;
    case K_VEL_FAULT: case K_PML_VEL: return 1;
    case K_STR_INTERIOR: return small ? 2 : 0;
    case K_STR_FAULT: case K_PML_STR: return 2;
    default: return 0;
  }
}

// Problems below this many cells capture overlapping half-steps, interior
// sweeps included: their launches are too small to fill the GPU one after
// another.  Above it the interior sweeps run alone (overlapping them only
// makes the sweeps compete for HBM: 1024x512x512 step 8.75 -> 9.15 ms),
// and only the small fault-plane kernels run beside the PML sweeps they
// are independent of (vel_fault with pml_vel, str_fault with pml_str:
// disjoint writes, neither reads what the other writes).
constexpr long long kOverlapCells = 4LL << 20;

// The step captured into three graphs on h->stream (kernel arguments are
// fixed for the life of the handle: the wavefield set and the time come from
// the device step counter, so the graphs serve every step):
//   head    the step up to and including advance;
//   rotated the previous step's fault-plane kernels (traction .. nucleation)
//           and then vel_fault on the side stream, beside this step's other
//           velocity sweeps (pml_vel, and vel_interior when the interior
//           runs as two sweeps) on the main stream; then the rest of the
//           step up to advance;
//   tail    the fault-plane kernels of the last step.
// The fault-plane kernels read the finished step's new set and the fault
// arrays; the velocity sweeps read the same set and write the other one
// and the PML splits, so the two groups share no written data.  vel_fault
// reads what friction wrote and follows it on the side stream.  In a run
// of kernels of the same half-step group (half_step_group) the first goes
// on the main stream and the rest on the side stream (fork and join by
// events), so they run side by side.
static int capture_graphs(fd3d_handle *h) {
  const bool small = (long long)h->cfg.nxt * h->cfg.nyt * h->cfg.nzt < kOverlapCells;
  const std::vector<KernelId> &seq = h->seq;
  const size_t nadv = (size_t)(std::find(seq.begin(), seq.end(), K_ADVANCE) - seq.begin());
  if (nadv >= seq.size()) return fail(h, FD3D_E_STATE, "step sequence without advance");
  const std::vector<KernelId> pre(seq.begin(), seq.begin() + nadv + 1), chain(seq.begin() + nadv + 1, seq.end());
  size_t nlead = 0;   // leading velocity sweeps of the step
  while (nlead < pre.size() && (pre[nlead] == K_VEL_INTERIOR || pre[nlead] == K_VEL_FAULT || pre[nlead] == K_PML_VEL))
    ++nlead;
  for (int kind = 0; kind < 3; ++kind) {
    const hipError_t eb = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (eb != hipSuccess) {
      drop_graph(h);   // all three graphs or none
      return fail(h, FD3D_E_HIP, std::string("begin capture: ") + hipGetErrorString(eb));
    }
    hipError_t e = hipSuccess;
    std::string what;
    auto on = [&](KernelId id, hipStream_t st) {
      if (e == hipSuccess && (e = launch(id, h->p, st)) != hipSuccess) what = kKernelNames[id];
    };
    auto ev = [&](hipEvent_t evt, hipStream_t rec, hipStream_t wait) {
      if (e == hipSuccess && (e = hipEventRecord(evt, rec)) == hipSuccess) e = hipStreamWaitEvent(wait, evt, 0);
      if (e != hipSuccess && what.empty()) what = "fork/join";
    };
    // kernels q[from..] with the half-step groups side by side
    auto grouped = [&](const std::vector<KernelId> &q, size_t from) {
      for (size_t n = from; n < q.size() && e == hipSuccess;) {
        size_t m = n + 1;
        const int grp = half_step_group(q[n], small);
        while (grp && m < q.size() && half_step_group(q[m], small) == grp) ++m;
        if (m - n > 1) ev(h->fork, h->stream, h->side);
        on(q[n], h->stream);
        for (size_t r = n + 1; r < m; ++r) on(q[r], h->side);
        if (m - n > 1) ev(h->join, h->side, h->stream);
        n = m;
      }
    };
    if (kind == 0) {
      grouped(pre, 0);
    } else if (kind == 1) {
      ev(h->fork, h->stream, h->side);
      for (KernelId id : chain) on(id, h->side);
      on(K_VEL_FAULT, h->side);
      for (size_t n = 0; n < nlead; ++n)
        if (pre[n] != K_VEL_FAULT) on(pre[n], h->stream);
      ev(h->join, h->side, h->stream);
      grouped(pre, nlead);
    } else {
      for (KernelId id : chain) on(id, h->stream);
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess || ee != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      drop_graph(h);
      return fail(h, FD3D_E_HIP, "capture " + what + ": " + hipGetErrorString(e != hipSuccess ? e : ee));
    }
    h->graph[kind] = g;
    const hipError_t ei = hipGraphInstantiate(&h->gexec[kind], g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      drop_graph(h);
      return fail(h, FD3D_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    }
  }
  return 0;
}

// Queue n steps on h->stream: kernel by kernel with events when profiling,
// else head, n-1 rotated and the tail graph.
static int run_steps(fd3d_handle *h, int n) {
  if (n <= 0) return 0;
  if (h->prof) {
    for (int s = 0; s < n; ++s)
      for (KernelId id : h->seq) {
        int rc = launch_one(h, id);
        if (rc) return rc;
      }
    return 0;
  }
  if (!h->gexec[0] || !h->gexec[1] || !h->gexec[2]) {
    drop_graph(h);
    int rc = capture_graphs(h);
    if (rc) return rc;
  }
  HIPCHK(h, hipGraphLaunch(h->gexec[0], h->stream));
  for (int s = 1; s < n; ++s) HIPCHK(h, hipGraphLaunch(h->gexec[1], h->stream));
  HIPCHK(h, hipGraphLaunch(h->gexec[2], h->stream));
  return 0;
}

int fd3d_step(fd3d_handle *h, int nsteps, float *out, float *seis) {
  if (!h) return FD3D_E_INVALID;
  if (!h->inited) return fail(h, FD3D_E_STATE, "fd3d_init_state first");
  if (h->batch) return fail(h, FD3D_E_STATE, "the problem is advanced by a batch (fd3d_batch_step)");
  if (nsteps < 0) return fail(h, FD3D_E_INVALID, "negative step count");
  if (h->begun != h->waited) return fail(h, FD3D_E_STATE, "output chunks in flight: fd3d_step_wait first");
  const int ring = h->cfg.ring_steps;
  const size_t slab = (size_t)h->p.g.nxT * h->p.g.nzT;
  const int nsta = h->cfg.nstations;
  int done = 0;
  while (done < nsteps) {
    const int pos = h->steps % ring;
    const int chunk = std::min(nsteps - done, ring - pos);
    int rc = run_steps(h, chunk);
    if (rc) return rc;
    if (out) {
      HIPCHK(h, hipMemcpyAsync(out + (size_t)done * FD3D_NOUT * slab, h->p.out.ring + (size_t)pos * FD3D_NOUT * slab,
                               (size_t)chunk * FD3D_NOUT * slab * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    }
    if (seis && nsta > 0) {
      HIPCHK(h, hipMemcpyAsync(seis + (size_t)done * nsta * 3, h->p.out.seis + (size_t)pos * nsta * 3,
                               (size_t)chunk * nsta * 3 * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    }
    if (out || seis) HIPCHK(h, hipStreamSynchronize(h->stream));
    h->steps += chunk;
    done += chunk;
  }
  if (out || seis) return check_ctl(h);
  return 0;
}

int fd3d_sync(fd3d_handle *h) {
  if (!h) return FD3D_E_INVALID;
  HIPCHK(h, owner_sync(h));
  return check_ctl(h);
}

// Pinned buffers, events and copy stream of the output stream, made at the
// first fd3d_step_begin.
static int stream_setup(fd3d_handle *h) {
  if (h->copy) return 0;
  const size_t slab = (size_t)h->p.g.nxT * h->p.g.nzT;
  const size_t half = (size_t)h->cfg.ring_steps / 2;
  const size_t floats = half * FD3D_NOUT * slab + half * (size_t)std::max(1, h->cfg.nstations) * 3;
  for (auto &c : h->chunk) {   // (a call after a failed one completes what is missing)
    if (!c.host) {
      void *p = nullptr;
      HIPCHK(h, hipHostMalloc(&p, floats * sizeof(float) + sizeof(Ctl), hipHostMallocDefault));
      c.host = static_cast<float *>(p);
      c.ctl = reinterpret_cast<Ctl *>(c.host + floats);
    }
    if (!c.computed) HIPCHK(h, hipEventCreateWithFlags(&c.computed, hipEventDisableTiming));
    if (!c.copied) HIPCHK(h, hipEventCreateWithFlags(&c.copied, hipEventDisableTiming));
  }
  HIPCHK(h, hipStreamCreateWithFlags(&h->copy, hipStreamNonBlocking));
  return 0;
}

// Advance nsteps and queue the copy of their outputs (see the header).
int fd3d_step_begin(fd3d_handle *h, int nsteps) {
  if (!h) return FD3D_E_INVALID;
  if (!h->inited) return fail(h, FD3D_E_STATE, "fd3d_init_state first");
  if (h->batch) return fail(h, FD3D_E_STATE, "the problem is advanced by a batch (fd3d_batch_step)");
  const int half = h->cfg.ring_steps / 2;
  if (nsteps < 1 || nsteps > half)
    return fail(h, FD3D_E_INVALID, "fd3d_step_begin: 1 <= nsteps <= ring_steps/2 = " + std::to_string(half));
  if (h->begun - h->waited >= 2)
    return fail(h, FD3D_E_STATE, "two chunks are in flight: fd3d_step_wait first");
  int rc = stream_setup(h);
  if (rc) return rc;
  fd3d_handle::Chunk &c = h->chunk[h->begun & 1];
  // the chunk that used this host buffer (two chunks ago) has been handed
  // out, and its copy has finished; the device ring slots of this chunk
  // were last copied by that chunk or an earlier one (two chunks span at
  // most ring_steps slots), so the steps may overwrite them
  if (h->begun >= 2) HIPCHK(h, hipStreamWaitEvent(h->stream, c.copied, 0));
  const int ring = h->cfg.ring_steps;
  c.first = h->steps + 1;
  c.n = nsteps;
  if ((rc = run_steps(h, nsteps))) return rc;
  HIPCHK(h, hipEventRecord(c.computed, h->stream));
  HIPCHK(h, hipStreamWaitEvent(h->copy, c.computed, 0));
  const size_t slab = (size_t)h->p.g.nxT * h->p.g.nzT;
  const int nsta = h->cfg.nstations;
  float *seis = c.host + (size_t)half * FD3D_NOUT * slab;
  // step s sits in ring slot (s-1) % ring: at most two contiguous runs.
  // (All five slabs of a step, the zero psi slab of slip weakening too: a
  // strided copy that skips it measured far slower than the whole run.)
  int done = 0;
  while (done < nsteps) {
    const int slot = (c.first - 1 + done) % ring;
    const int n = std::min(nsteps - done, ring - slot);
    HIPCHK(h, hipMemcpyAsync(c.host + (size_t)done * FD3D_NOUT * slab, h->p.out.ring + (size_t)slot * FD3D_NOUT * slab,
                             (size_t)n * FD3D_NOUT * slab * sizeof(float), hipMemcpyDeviceToHost, h->copy));
    if (nsta > 0)
      HIPCHK(h, hipMemcpyAsync(seis + (size_t)done * nsta * 3, h->p.out.seis + (size_t)slot * nsta * 3,
                               (size_t)n * nsta * 3 * sizeof(float), hipMemcpyDeviceToHost, h->copy));
    done += n;
  }
  // the flags as of this chunk (or a later step: the next chunk may already
  // run; they only grow)
  HIPCHK(h, hipMemcpyAsync(c.ctl, h->p.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->copy));
  HIPCHK(h, hipEventRecord(c.copied, h->copy));
  h->steps += nsteps;
  h->begun += 1;
  return 0;
}

// Wait for the oldest chunk's copy and hand out its host buffer.
int fd3d_step_wait(fd3d_handle *h, int *nsteps, const float **out, const float **seis) {
  if (!h) return FD3D_E_INVALID;
  if (h->waited >= h->begun) return fail(h, FD3D_E_STATE, "no chunk in flight: fd3d_step_begin first");
  fd3d_handle::Chunk &c = h->chunk[h->waited & 1];
  HIPCHK(h, hipEventSynchronize(c.copied));
  h->waited += 1;
  const size_t slab = (size_t)h->p.g.nxT * h->p.g.nzT;
  if (nsteps) *nsteps = c.n;
  if (out) *out = c.host;
  if (seis) *seis = c.host + (size_t)(h->cfg.ring_steps / 2) * FD3D_NOUT * slab;
  if (c.ctl->newton_capped > 0)
    return fail(h, FD3D_E_NEWTON, "rate-and-state Newton solve hit the iteration cap at " +
                                      std::to_string(c.ctl->newton_capped) + " node-steps");
  if (c.ctl->nonfinite) return fail(h, FD3D_E_NONFINITE, "non-finite slip rate in the fault outputs");
  return 0;
}

int fd3d_steps_done(const fd3d_handle *h) { return h ? h->steps : 0; }

int fd3d_read_outputs(fd3d_handle *h, int nsteps, float *out, float *seis) {
  if (!h) return FD3D_E_INVALID;
  const int ring = h->cfg.ring_steps;
  if (nsteps < 0 || nsteps > ring || nsteps > h->steps)
    return fail(h, FD3D_E_INVALID, "nsteps must be within the steps taken and the output ring");
  HIPCHK(h, owner_sync(h));
  const size_t slab = (size_t)h->p.g.nxT * h->p.g.nzT;
  const int nsta = h->cfg.nstations;
  // step s (1-based) sits in slot (s-1) % ring
  int done = 0;
  while (done < nsteps) {
    const int first = h->steps - nsteps + done + 1;
    const int slot = (first - 1) % ring;
    const int n = std::min(nsteps - done, ring - slot);
    if (out)
      HIPCHK(h, hipMemcpy(out + (size_t)done * FD3D_NOUT * slab, h->p.out.ring + (size_t)slot * FD3D_NOUT * slab,
                          (size_t)n * FD3D_NOUT * slab * sizeof(float), hipMemcpyDeviceToHost));
    if (seis && nsta > 0)
      HIPCHK(h, hipMemcpy(seis + (size_t)done * nsta * 3, h->p.out.seis + (size_t)slot * nsta * 3,
                          (size_t)n * nsta * 3 * sizeof(float), hipMemcpyDeviceToHost));
    done += n;
  }
  return check_ctl(h);
}

int fd3d_download_final(fd3d_handle *h, float *out7) {
  if (!h || !out7) return fail(h, FD3D_E_INVALID, "null argument");
  HIPCHK(h, owner_sync(h));
  const size_t n2 = (size_t)h->cfg.nxt * h->cfg.nzt;
  std::vector<float> t0x(n2), t0z(n2), rise(n2), schx(n2), schz(n2);
  Fault &fl = h->p.fl;
  HIPCHK(h, hipMemcpy(out7, fl.ruptime, n2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(rise.data(), fl.rise, n2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(out7 + 2 * n2, fl.sliptime, n2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(out7 + 3 * n2, fl.slipX, n2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(out7 + 4 * n2, fl.slipZ, n2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(schx.data(), fl.schangeX, n2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(schz.data(), fl.schangeZ, n2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(t0x.data(), fl.T0X, n2 * 4, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(t0z.data(), fl.T0Z, n2 * 4, hipMemcpyDeviceToHost));
  for (size_t q = 0; q < n2; ++q) {
    out7[n2 + q] = rise[q] - out7[q];          // fd3d_theo.f90:658
    out7[5 * n2 + q] = schx[q] - t0x[q];       // :607
    out7[6 * n2 + q] = schz[q] - t0z[q];       // :606
  }
  return 0;
}

// Field `which` of wavefield set `set` (the state after the last step is
// in set steps & 1).
static float *field_ptr(fd3d_handle *h, int which, int set) {
  Fields &f = h->p.f[set];
  switch (which) {
    case FD3D_U: return f.u; case FD3D_V: return f.v; case FD3D_W: return f.w;
    case FD3D_XX: return f.xx; case FD3D_YY: return f.yy; case FD3D_ZZ: return f.zz;
    case FD3D_XY: return f.xy; case FD3D_YZ: return f.yz; case FD3D_XZ: return f.xz;
    default: return nullptr;
  }
}

int fd3d_download_field(fd3d_handle *h, int which, float *out) {
  if (!h || !out) return fail(h, FD3D_E_INVALID, "null argument");
  float *src = field_ptr(h, which, h->steps & 1);
  if (!src) return fail(h, FD3D_E_INVALID, "unknown field");
  return download3(h, out, src);
}

int fd3d_upload_field(fd3d_handle *h, int which, const float *in) {
  if (!h || !in) return fail(h, FD3D_E_INVALID, "null argument");
  if (!field_ptr(h, which, 0)) return fail(h, FD3D_E_INVALID, "unknown field");
  HIPCHK(h, owner_sync(h));
  // both sets: cells no stage ever writes (the outer faces) must read the
  // same in either set
  int rc = upload3(h, field_ptr(h, which, 0), in);
  if (rc == 0) rc = upload3(h, field_ptr(h, which, 1), in);
  return rc;
}

int fd3d_set_source_bins(fd3d_handle *h, int NL, int NW, float dL, float dW, float dtseis, int nbins) {
  if (!h) return FD3D_E_INVALID;
  if (NL < 0 || NW < 0 || nbins < 1 || !(dtseis > 0) || ((NL > 0 || NW > 0) && !(dL > 0 && dW > 0)))
    return fail(h, FD3D_E_INVALID, "source bins: NL, NW >= 0, dL, dW, dtseis > 0, nbins >= 1");
  if (h->batch) return fail(h, FD3D_E_STATE, "remove the problem from its batch first");
  if (h->p.rates.on) return fail(h, FD3D_E_STATE, "source bins are already set");
  Rates &R = h->p.rates;
  const size_t nb = (size_t)NL * NW * nbins;
  int rc;
  if ((rc = dalloc(h, &R.msrx, std::max<size_t>(nb, 1)))) return rc;
  if ((rc = dalloc(h, &R.msrz, std::max<size_t>(nb, 1)))) return rc;
  if ((rc = dalloc(h, &R.mr, (size_t)nbins))) return rc;
  HIPCHK(h, hipMemset(R.msrx, 0, std::max<size_t>(nb, 1) * sizeof(float)));
  HIPCHK(h, hipMemset(R.msrz, 0, std::max<size_t>(nb, 1) * sizeof(float)));
  HIPCHK(h, hipMemset(R.mr, 0, (size_t)nbins * sizeof(float)));
  R.NL = NL; R.NW = NW; R.nbins = nbins; R.dL = dL; R.dW = dW; R.dtseis = dtseis; R.on = 1;
  build_seq(h);
  return 0;
}

int fd3d_download_source_bins(fd3d_handle *h, float *msrx, float *msrz, float *moment_rate) {
  if (!h) return FD3D_E_INVALID;
  const Rates &R = h->p.rates;
  if (!R.on) return fail(h, FD3D_E_STATE, "fd3d_set_source_bins first");
  HIPCHK(h, owner_sync(h));
  const size_t nb = (size_t)R.NL * R.NW * R.nbins;
  if (msrx && nb) HIPCHK(h, hipMemcpy(msrx, R.msrx, nb * sizeof(float), hipMemcpyDeviceToHost));
  if (msrz && nb) HIPCHK(h, hipMemcpy(msrz, R.msrz, nb * sizeof(float), hipMemcpyDeviceToHost));
  if (moment_rate) HIPCHK(h, hipMemcpy(moment_rate, R.mr, (size_t)R.nbins * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int fd3d_set_tuning(fd3d_handle *h, int mode, int tz_vel, int ny_vel, int tz_str, int ny_str) {
  if (!h) return FD3D_E_INVALID;
  auto ok_tz = [](int tz) { return tz == 4 || tz == 8 || tz == 16; };
  if (mode < 0 || mode > 127) return fail(h, FD3D_E_INVALID, "tuning mode must be 0..127");
  if ((mode & 96) && !(mode & 4)) return fail(h, FD3D_E_INVALID, "the strip width (bits 5, 6) needs bit 2");
  if ((mode & 24) && !(mode & 1)) return fail(h, FD3D_E_INVALID, "tile widths (bits 3, 4) need bit 0");
  if (((mode & 8) && tz_vel > 8) || ((mode & 16) && tz_str > 8))
    return fail(h, FD3D_E_INVALID, "128-wide tiles take tz 4 or 8");
  if ((mode & 4) && ny_vel < 1) return fail(h, FD3D_E_INVALID, "rows per block must be >= 1");
  if ((mode & 1) && (!ok_tz(tz_vel) || !ok_tz(tz_str) || ny_vel < 1 || ny_str < 1))
    return fail(h, FD3D_E_INVALID, "tile height must be 4, 8 or 16 and rows >= 1");
  if (h->batch) return fail(h, FD3D_E_STATE, "remove the problem from its batch first");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->p.tune.mode = mode;
  if (mode & 5) {
    h->p.tune.tz_vel = tz_vel; h->p.tune.ny_vel = ny_vel;
    h->p.tune.tz_str = tz_str; h->p.tune.ny_str = ny_str;
  }
  build_seq(h);
  return 0;
}

int fd3d_get_tuning(const fd3d_handle *h, int *out5) {
  if (!h || !out5) return FD3D_E_INVALID;
  const Tuning &t = h->p.tune;
  out5[0] = t.mode; out5[1] = t.tz_vel; out5[2] = t.ny_vel; out5[3] = t.tz_str; out5[4] = t.ny_str;
  return 0;
}

int fd3d_set_profiling(fd3d_handle *h, int on) {
  if (!h) return FD3D_E_INVALID;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  harvest(h);
  for (int k = 0; k < K_COUNT; ++k) { h->kms[k] = 0; h->kcount[k] = 0; }
  h->prof = on != 0;
  return 0;
}

int fd3d_kernel_times(fd3d_handle *h, int n, const char **names, double *ms, long *launches) {
  if (!h) return FD3D_E_INVALID;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  harvest(h);
  int c = 0;
  for (int k = 0; k < K_COUNT && c < n; ++k) {
    if (h->kcount[k] == 0) continue;
    if (names) names[c] = kKernelNames[k];
    if (ms) ms[c] = h->kms[k];
    if (launches) launches[c] = h->kcount[k];
    ++c;
  }
  return c;
}

double fd3d_step_bytes(const fd3d_handle *h, long *bulk_cells, long *pml_cells) {
  if (!h) return 0.0;
  const Dims &g = h->p.g;
  long pml = 0;
  for (int r = 0; r < h->p.pml.nreg; ++r) pml += h->p.pml.r[r].ncell;
  // bulk: the 4th-order interior and the near-fault rows j = nyt-1, nyt
  const long bulk = (long)(g.nxt - 2 * g.nabc) * (long)(g.nyt - g.nabc) * (long)(g.nzt - g.nfs - g.nabc);
  if (bulk_cells) *bulk_cells = bulk;
  if (pml_cells) *pml_cells = pml;
  return 120.0 * (double)bulk + 276.0 * (double)pml;
}

const char *fd3d_last_error(const fd3d_handle *h) { return h ? h->err.c_str() : "null handle"; }

void fd3d_destroy(fd3d_handle *h) {
  if (!h) return;
  if (h->batch) (void)fd3d_batch_remove(h->batch, h);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto &e : h->evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
  for (auto e : h->pool) (void)hipEventDestroy(e);
  drop_graph(h);
  if (h->copy) (void)hipStreamSynchronize(h->copy);
  for (auto &c : h->chunk) {
    if (c.host) (void)hipHostFree(c.host);
    if (c.computed) (void)hipEventDestroy(c.computed);
    if (c.copied) (void)hipEventDestroy(c.copied);
  }
  if (h->copy) (void)hipStreamDestroy(h->copy);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->fork) (void)hipEventDestroy(h->fork);
  if (h->join) (void)hipEventDestroy(h->join);
  if (h->side) (void)hipStreamDestroy(h->side);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"
