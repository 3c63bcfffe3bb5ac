// Multi-domain batch engine: many resident problems of different sizes
// advanced together, their steps replayed from captured HIP graphs.
//
// Work decomposition.  Every stage of one problem's step (the kernel
// sequence of capi.hip) is cut into work items of one 64 x TZ workgroup:
// interior tiles (tiles.h), 2D blocks of the fault and free-surface planes
// and 256-cell chunks of the PML (stages.h).  Items are numbered
// stage-major: all problems' velocity tiles first, then all problems'
// near-fault velocity blocks, and so on.  A (problem, stage) pair is a
// segment of consecutive item numbers.
//
// Persistent mode (FD3D_BATCH_PERSISTENT): one launch per step with as
// many workgroups as fit on the device at once.  A workgroup takes the next
// item from a device queue, waits until the stage its problem runs before
// this one has finished all its items, runs the item and counts it done.
// Items are handed out in increasing order, so every item an item waits
// for was handed out earlier to a running workgroup, whose own waits are on
// still earlier items: the earliest unfinished item can always proceed and
// the launch drains.  Problems overlap freely: while one problem's fault
// stages run, other problems' sweeps fill the machine, with no launch
// boundary between stages.  Cross-workgroup visibility uses device-scope
// fences around the counters (writers release, waiters acquire), which
// also covers the per-XCD L2 caches.  A wait is bounded (about a second);
// a wait that runs out sets a sticky flag that fd3d_batch_sync reports.
//
// Staged mode (FD3D_BATCH_STAGED): one launch per stage covering every
// problem; kernel boundaries order the stages.
//
// Either way the launches are captured into HIP graphs, re-captured when
// the membership changes: the persistent form's step is one graph; the
// staged form's run of n steps is a head, n-1 rotated and a tail graph, so
// that each step's fault-plane stages run beside the next step's velocity
// stages (capture_staged_t).  The graphs hold kernel nodes only: the
// per-step reset of the queue and counters is a kernel, not a memset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "handle.h"
#include "fused.h"
#include "stages.h"
#include "tiles.h"

namespace fd3d {

struct BatchDomain {
  Problem p;
  int n[K_COUNT];     // work items of each stage (0: the problem has no such stage)
  int dep[K_COUNT];   // stage of the same problem that must finish first (-1: none)
  int gx, gz;         // interior tiles along x and z
  int nyv, nys;       // interior rows per item, velocity and stress
  int gyv;            // interior chunks along y, velocity (stress: n / (gx*gz))
  int fpx;            // 64-wide blocks along x over nxt (near-fault planes, free surface)
  int fix;            // 64-wide blocks along x over the fault interior
  int nux;            // 64-wide blocks along x over the nucleation patch
  int nout;           // fault-node items of the output stage; station items follow
  int fgx, fgz, nyf;  // fused sweep: tiles along x and z, rows per item
  int sty, stz0;      // top-plane stresses (free surface with the fused sweep):
                      // row blocks of tz rows, first plane's index past nabc
};

struct BatchSeg {
  int start, dom, stage, pad;
};

struct BatchArgs {
  const BatchSeg *seg;
  const BatchDomain *dom;
  unsigned *cnt;      // [0] queue head, [1] failed waits (sticky), [2 + d*K_COUNT + s] items done
  int nseg, total;
  long long wait_ticks;   // bound of one dependency wait in wall-clock ticks
};

constexpr int kCntBase = 2;
constexpr double kWaitSeconds = 0.25;   // a dependency wait longer than this is a failure

namespace {

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// One work item.  FVW = false leaves out the rate-and-state code (batches
// whose members all use slip weakening), which lowers register use.
template <int TZ, bool FVW>
__device__ __forceinline__ void run_item(const BatchDomain &D, int stage, int b, float *lds, int lx, int ly) {
  constexpr int BLK = 64 * TZ;
  const Problem &P = D.p;
  const int t = ly * 64 + lx;
  // the fault-plane stages after advance read the step it has just counted
  const FieldPair FP = stage > K_ADVANCE ? field_pair_done(P.f[0], P.f[1], P.ctl) : field_pair(P.f[0], P.f[1], P.ctl);
  const Fields &fo = FP.fo, &fn = FP.fn;
  switch (stage) {
    case K_VEL_INTERIOR:
      vel_tiled_body<64, TZ>(P.g, P.c, fo, fn, lds, b % D.gx, (b / D.gx) % D.gz, b / (D.gx * D.gz), lx, ly, D.nyv);
      break;
    case K_VEL_FAULT:
      vel_fault_body<TZ, FVW>(P.g, P.c, fo, fn, P.fl, P.fv, P.ctl, b % D.fpx, b / D.fpx, lx, ly);
      break;
    case K_PML_VEL: {
      const int gi = b * BLK + t;
      pml_vel_body(P.g, P.c, fo, fn, P.pml, gi / kPmlBlock, gi % kPmlBlock);
      break;
    }
    case K_FUVW_UV:
      fuvw_uv_body<TZ>(P.g, fo, fn, b % D.fpx, b / D.fpx, lx, ly);
      break;
    case K_FUVW_W:
      fuvw_w_body<TZ>(P.g, fo, fn, b % D.fpx, b / D.fpx, lx, ly);
      break;
    case K_STR_TOP: {
      // one 64 x TZ block of interior (i, j) cells of one of the two top
      // planes: the one-thread-per-cell body with its rows of BYF lanes
      const int bx = b % D.fix, r = b / D.fix, jb = r % D.sty, z = r / D.sty;
      str_interior_body(P.g, P.c, fo, fn, bx, jb * (TZ / BYF), D.stz0 + z, lx, ly);
      break;
    }
    case K_STR_INTERIOR:
      str_tiled_body<64, TZ>(P.g, P.c, fo, fn, lds, b % D.gx, (b / D.gx) % D.gz, b / (D.gx * D.gz), lx, ly, D.nys);
      break;
    case K_STR_FAULT:
      str_fault_body<TZ>(P.g, P.c, fo, fn, P.fl, b % D.fpx, b / D.fpx, lx, ly);
      break;
    case K_PML_STR: {
      const int gi = b * BLK + t;
      pml_str_body(P.g, P.c, fo, fn, P.pml, gi / kPmlBlock, gi % kPmlBlock);
      break;
    }
    case K_FRES:
      fres_body<TZ>(P.g, fo, fn, b % D.fpx, b / D.fpx, lx, ly);
      break;
    case K_TRACTION:
      traction_body<TZ>(P.g, P.c, fo, fn, P.fl, b % D.fix, b / D.fix, lx, ly);
      break;
    case K_TRACTION_MIRROR:
      traction_mirror_body(P.g, P.fl, b * BLK + t);
      break;
    case K_TABS:
      tabs_body<TZ>(P.g, P.fl, b % D.fix, b / D.fix, lx, ly);
      break;
    case K_FRICTION:
      friction_body<TZ, FVW>(P.g, P.c, fo, fn, P.fl, P.fv, P.ctl, b % D.fix, b / D.fix, lx, ly);
      break;
    case K_OUTPUTS:
      if (b < D.nout)
        outputs_body<TZ>(P.g, fo, fn, P.fl, P.fv, P.out, P.ctl, b % D.fix, b / D.fix, lx, ly);
      else
        stations_body(P.g, fo, fn, P.out, P.ctl, (b - D.nout) * BLK + t);
      break;
    case K_ADVANCE:
      if (t == 0) advance_body(P.ctl);
      break;
    case K_RATES:
      rates_body<BLK>(P.g, P.c, fn, P.out, P.rates, P.ctl, b, t, reinterpret_cast<double *>(lds));
      break;
    case K_NUCLEATION:
      nucleation_body<TZ>(P.g, P.c, P.fl, P.fv, P.ctl, b % D.nux, b / D.nux, lx, ly);
      break;
    default:
      break;
  }
}

// Segment holding item `item` among segments [lo, hi] (last start <= item).
__device__ inline int find_seg(const BatchSeg *seg, int lo, int hi, int item) {
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg[mid].start <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int TZ, bool FVW>
__global__ void __launch_bounds__(64 * TZ) k_batch_persistent(BatchArgs A) {
  __shared__ float lds[Tile<64, TZ>::LDS_FLOATS];
  __shared__ int s_item, s_seg;
  const int t = threadIdx.y * 64 + threadIdx.x;
  for (;;) {
    if (t == 0) {
      const int item = (int)atomicAdd(&A.cnt[0], 1u);
      int sg = 0;
      if (item < A.total) {
        sg = find_seg(A.seg, 0, A.nseg - 1, item);
        const BatchSeg S = A.seg[sg];
        const BatchDomain &D = A.dom[S.dom];
        const int dep = D.dep[S.stage];
        if (dep >= 0) {
          // Spin on a relaxed device-scope load (no cache maintenance per
          // poll); a single acquire fence follows.  The wait is bounded in
          // wall-clock time, and once one wait has failed every other wait
          // gives up at once, so the launch always drains quickly.
          const unsigned need = (unsigned)D.n[dep];
          const unsigned *c = &A.cnt[kCntBase + S.dom * K_COUNT + dep];
          const long long t0 = wall_clock64();
          while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            __builtin_amdgcn_s_sleep(1);
            if (__hip_atomic_load(&A.cnt[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
            if (wall_clock64() - t0 > A.wait_ticks) {
              atomicOr(&A.cnt[1], 1u);
              break;
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
      }
      s_item = item;
      s_seg = sg;
    }
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= A.total) break;
    const int sg = __builtin_amdgcn_readfirstlane(s_seg);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // each wave: no stale L1 lines of the producers' data
    const int dom = A.seg[sg].dom, stage = A.seg[sg].stage;
    run_item<TZ, FVW>(A.dom[dom], stage, item - A.seg[sg].start, lds, threadIdx.x, threadIdx.y);
    // each wave waits for its own stores to complete (no cache maintenance),
    // then lane 0 releases them all with one device-scope release
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0); expcnt, lgkmcnt not waited
    __syncthreads();  // also frees LDS and s_item for the next item
    if (t == 0) {
      // one release per item: the workgroup's stores reach memory before the count
      __hip_atomic_fetch_add(&A.cnt[kCntBase + dom * K_COUNT + stage], 1u, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// The fused interior sweep of every member that runs it: one 64 x 16
// workgroup per fused tile and y chunk (fused.h), tiles in strip order.
__global__ void __launch_bounds__(fused::NT) k_batch_fused(BatchArgs A, int seg_lo, int seg_hi, int item0, int n) {
  __shared__ fused::FusedLds L;
  const int item = item0 + xcd_remap((int)blockIdx.x, n);
  const int sg = find_seg(A.seg, seg_lo, seg_hi, item);
  const BatchDomain &D = A.dom[A.seg[sg].dom];
  const int b = item - A.seg[sg].start;
  // under a free surface the top two planes' stresses are the str_top stage
  const int kmax_s = D.p.g.nzt - D.p.g.nfs - (D.p.g.fspace ? 0 : 2);
  const int per = D.fgx * D.fgz, by = b / per;
  int bx, bz;
  fused::strip_tile(b - by * per, D.fgx, D.fgz, fused::kStrip, bx, bz);
  fused::fused_body(D.p.g, D.p.c, D.p.f[0], D.p.f[1], D.p.ctl, bx, bz, by, D.nyf, kmax_s, L);
}

// One stage of every member; the stage is a template argument, so each
// instantiation holds (and allocates registers for) one stage body only.
template <int TZ, bool FVW, int STAGE>
__global__ void __launch_bounds__(64 * TZ) k_batch_stage(BatchArgs A, int seg_lo, int seg_hi, int item0) {
  __shared__ float lds[STAGE == K_VEL_INTERIOR || STAGE == K_STR_INTERIOR ? Tile<64, TZ>::LDS_FLOATS
                       : STAGE == K_RATES                                  ? 2 * 64 * TZ
                                                                           : 1];
  const int item = item0 + (int)blockIdx.x;
  const int sg = find_seg(A.seg, seg_lo, seg_hi, item);
  run_item<TZ, FVW>(A.dom[A.seg[sg].dom], STAGE, item - A.seg[sg].start, lds, threadIdx.x, threadIdx.y);
}

template <int TZ, bool FVW, int S>
void launch_stage(int n, hipStream_t st, const BatchArgs &A, int lo, int hi, int i0) {
  hipLaunchKernelGGL((k_batch_stage<TZ, FVW, S>), dim3(n), dim3(64, TZ), 0, st, A, lo, hi, i0);
}

// Stage launch by run-time stage id: a table of the per-stage
// instantiations over every KernelId (0 .. K_COUNT-1), so it follows the
// enum as stages are added.
template <int TZ, bool FVW, int... S>
void launch_stage_tab(int s, int n, hipStream_t st, const BatchArgs &A, int lo, int hi, int i0,
                      std::integer_sequence<int, S...>) {
  using Fn = void (*)(int, hipStream_t, const BatchArgs &, int, int, int);
  static constexpr Fn tab[] = {&launch_stage<TZ, FVW, S>...};
  static_assert(sizeof(tab) / sizeof(tab[0]) == K_COUNT, "one launcher per stage");
  if (s >= 0 && s < K_COUNT) tab[s](n, st, A, lo, hi, i0);
}

template <int TZ, bool FVW>
void launch_stage_rt(int s, int n, hipStream_t st, const BatchArgs &A, int lo, int hi, int i0) {
  launch_stage_tab<TZ, FVW>(s, n, st, A, lo, hi, i0, std::make_integer_sequence<int, K_COUNT>{});
}

// Per-step reset of the queue head and the done counters (a kernel, so
// that the captured graph holds no memset node).
__global__ void k_batch_reset(unsigned *cnt, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) cnt[0] = 0u;
  if (i < n) cnt[kCntBase + i] = 0u;
}

__global__ void k_zero_u32(unsigned *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// The batch's report after a chunk, gathered so that the host reads it in
// one copy: rep[0] the sticky failed-wait count, then per member q its
// control block's Newton-cap count and non-finite flag.
__global__ void k_batch_report(const BatchDomain *dom, int n, const unsigned *cnt, unsigned *rep) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q == 0) rep[0] = cnt[1];
  if (q < n) {
    const Ctl *c = dom[q].p.ctl;
    rep[1 + 2 * q] = (unsigned)c->newton_capped;
    rep[2 + 2 * q] = (unsigned)c->nonfinite;
  }
}

}  // namespace

// Fixed device tables of the batch that allocates nothing
// (fd3d_batch_create_static): the member descriptors, the segment table and
// the counters of up to kStaticMaxProblems problems, in one array per GPU
// that lives until the process ends.  The array's name is a host-side
// stand-in; host code uses only the device address hipGetSymbolAddress
// returns for the current device.
constexpr int kStaticMaxProblems = 64;
constexpr size_t align256(size_t n) { return (n + 255) / 256 * 256; }
constexpr size_t kStaticDomBytes = align256(sizeof(BatchDomain) * kStaticMaxProblems);
constexpr size_t kStaticSegBytes = align256(sizeof(BatchSeg) * K_COUNT * kStaticMaxProblems);
constexpr size_t kStaticCntBytes =
    align256(sizeof(unsigned) * (kCntBase + K_COUNT * kStaticMaxProblems + 1 + 2 * kStaticMaxProblems));
constexpr size_t kScratchBytes = kStaticDomBytes + kStaticSegBytes + kStaticCntBytes;
static_assert(kScratchBytes <= (1u << 20), "the fixed batch tables stay under 1 MiB per GPU");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

}  // namespace fd3d

struct fd3d_batch {
  int maxdom = 0, tz = 16, mode = 0;
  std::vector<fd3d_handle *> mem;
  hipStream_t stream = nullptr;
  BatchDomain *d_dom = nullptr;
  BatchSeg *d_seg = nullptr;
  unsigned *d_cnt = nullptr;
  unsigned *d_rep = nullptr;    // report after a chunk (k_batch_report), behind the counters in d_cnt
  std::vector<unsigned> rep;    // its host copy: 1 + 2 * maxdom words
  int seg_cap = 0;
  std::vector<BatchSeg> segs;
  int stage_seg0[K_COUNT + 1] = {0};
  int stage_item0[K_COUNT + 1] = {0};
  int total = 0, workers = 0, launches = 0;
  bool dirty = true;
  bool fvw = false;   // some member uses rate-and-state friction
  int fused_items = 8;   // least fused work items per CU for the fused stage
  int fixed_dev = -1;    // fd3d_batch_create_static: tables in the device's fixed scratch, else -1
  long long wait_ticks = 0;
  // staged mode: head, rotated and tail graphs of a step (see capture_staged_t);
  // persistent mode: graph 0 only
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
  hipStream_t side = nullptr;               // the captured fork's second stream
  hipEvent_t fork = nullptr, join = nullptr;
  struct Ev { hipEvent_t a, b; int n; };
  std::vector<Ev> evs;
  std::vector<hipEvent_t> pool;
  std::string err;
};

namespace {

int bfail(fd3d_batch *b, int code, const std::string &m) {
  if (b) b->err = m;
  return code;
}

#define BCHK(b, call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return bfail((b), e_ == hipErrorOutOfMemory ? FD3D_E_NOMEM : FD3D_E_HIP,               \
                   std::string(#call) + ": " + hipGetErrorString(e_));                       \
  } while (0)

// Work items of each stage of h at tile height tz.  With fused allowed
// (staged launches, enough fused work in the batch) a member whose tuning
// asks for the fused interior sweep (mode bit 2) runs it as a stage, and
// under a free surface the top-plane stresses as the str_top stage after
// the ghost velocities; otherwise the member runs the separate sweeps
// (work items of one sweep), whatever form it uses alone.
BatchDomain describe(const fd3d_handle *h, int tz, bool allow_fused) {
  BatchDomain D{};
  D.p = h->p;
  const Dims &g = h->p.g;
  const int blk = 64 * tz;
  const int ni = g.nxt - 2 * g.nabc, nyi = g.nyt - 2 - g.nabc, nk = g.nzt - g.nfs - g.nabc;
  D.gx = cdiv(ni, 64);
  D.gz = cdiv(nk, tz);
  const bool fz = allow_fused && (h->p.tune.mode & 4) && nyi > 0;
  D.nyv = std::max(1, h->p.tune.ny_vel);
  D.nys = std::max(1, h->p.tune.ny_str);
  // a member that asks for the fused sweep holds its fused rows in ny_vel;
  // when it runs the separate sweeps here instead, both use ny_str
  if ((h->p.tune.mode & 4) && !fz) D.nyv = D.nys;
  D.gyv = nyi > 0 ? cdiv(nyi, D.nyv) : 0;
  const int gys = nyi > 0 ? cdiv(nyi, D.nys) : 0;
  D.fpx = cdiv(g.nxt, 64);
  D.fix = cdiv(ni, 64);
  const int nxn = g.nxt - g.nabc - (g.nabc - 1) + 1, nzn = g.nzt - g.nfs - (g.nabc - 1) + 1;
  D.nux = cdiv(nxn, 64);
  D.nout = D.fix * cdiv(nk, tz);
  D.fgx = cdiv(ni, fused::LX);
  D.fgz = cdiv(nk, fused::ZC);
  D.nyf = std::max(1, h->p.tune.ny_vel);
  int items[K_COUNT] = {0};
  items[K_VEL_INTERIOR] = D.gx * D.gz * D.gyv;
  items[K_VEL_FAULT] = D.fpx * cdiv(g.nzt, tz);
  items[K_PML_VEL] = cdiv(h->p.pml.nblk * kPmlBlock, blk);
  items[K_FUVW_UV] = items[K_FUVW_W] = items[K_FRES] = D.fpx * cdiv(g.nyt, tz);
  items[K_STR_INTERIOR] = D.gx * D.gz * gys;
  items[K_STR_FAULT] = items[K_VEL_FAULT];
  items[K_PML_STR] = items[K_PML_VEL];
  items[K_TRACTION] = items[K_TABS] = items[K_FRICTION] = D.nout;
  items[K_TRACTION_MIRROR] = cdiv(ni, blk);
  items[K_OUTPUTS] = D.nout + cdiv(h->p.out.nsta, blk);
  items[K_NUCLEATION] = D.nux * cdiv(nzn, tz);
  items[K_ADVANCE] = 1;
  items[K_RATES] = h->p.rates.on ? 1 + cdiv(h->p.rates.NL * h->p.rates.NW, blk) : 0;
  items[K_FUSED] = fz ? D.fgx * D.fgz * cdiv(nyi, D.nyf) : 0;
  D.sty = cdiv(std::max(nyi, 1), tz);
  D.stz0 = nk >= 2 ? nk - 2 : 0;
  items[K_STR_TOP] = (fz && !g.fspace && nk > 0) ? D.fix * D.sty * (nk - D.stz0) : 0;
  for (int s = 0; s < K_COUNT; ++s) { D.n[s] = 0; D.dep[s] = -1; }
  std::vector<KernelId> seq;
  for (KernelId id : h->seq) {
    if (id == K_FUSED && !fz) { seq.push_back(K_VEL_INTERIOR); continue; }
    if (id == K_STR_TOP && !fz) continue;
    seq.push_back(id);
  }
  if (!fz) {
    auto it = std::find(seq.begin(), seq.end(), K_STR_FAULT);
    if (std::find(seq.begin(), seq.end(), K_STR_INTERIOR) == seq.end() && it != seq.end())
      seq.insert(it, K_STR_INTERIOR);
  }
  std::sort(seq.begin(), seq.end());
  int prev = -1;
  for (KernelId id : seq) {
    if (items[id] <= 0) continue;
    D.n[id] = items[id];
    D.dep[id] = prev;
    prev = id;
  }
  return D;
}

// Workgroups of the persistent launch: as many as are resident at once.
template <int TZ, bool FVW>
int size_workers(fd3d_batch *b) {
  int nb = 0, ncu = 0, dev = 0;
  BCHK(b, hipGetDevice(&dev));
  BCHK(b, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  BCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_persistent<TZ, FVW>, 64 * TZ, 0));
  b->workers = std::max(1, std::min(std::max(nb, 1) * ncu, b->total));
  int khz = 0;
  BCHK(b, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  b->wait_ticks = (long long)(kWaitSeconds * 1e3 * (double)std::max(khz, 1));
  return 0;
}

template <int TZ, bool FVW>
int capture_t(fd3d_batch *b) {
  BatchArgs A{b->d_seg, b->d_dom, b->d_cnt, (int)b->segs.size(), b->total, b->wait_ticks};
  const dim3 blk(64, TZ);
  const int nc = (int)b->mem.size() * K_COUNT;
  hipLaunchKernelGGL(k_batch_reset, dim3(cdiv(std::max(nc, 1), 256)), dim3(256), 0, b->stream, b->d_cnt, nc);
  hipLaunchKernelGGL((k_batch_persistent<TZ, FVW>), dim3(b->workers), blk, 0, b->stream, A);
  b->launches = 2;
  return 0;
}

// Staged mode: one launch per stage covering every member, captured as
//   head    a step up to and including advance;
//   rotated the previous step's fault-plane stages (traction .. nucleation)
//           and then vel_fault on the side stream, beside vel_interior and
//           pml_vel; then the rest of the step up to advance;
//   tail    the fault-plane stages of the last step.
// Within a step vel_fault runs beside vel_interior and pml_vel, and
// str_fault beside str_top, str_interior and pml_str: each group reads the
// same old data and writes disjoint cells.  The fault-plane stages read the
// finished step's new set and the fault arrays, which the velocity sweeps
// of the next step neither write nor read (the kernels after advance take
// the step it has counted).  A run of n steps replays head, n-1 rotated,
// tail.
template <int TZ, bool FVW>
int capture_staged_t(fd3d_batch *b, int kind) {
  BatchArgs A{b->d_seg, b->d_dom, b->d_cnt, (int)b->segs.size(), b->total, b->wait_ticks};
  int launches = 0;
  auto stage = [&](int s, hipStream_t st) {
    const int n = b->stage_item0[s + 1] - b->stage_item0[s];
    if (n <= 0) return;
    if (s == K_FUSED)
      hipLaunchKernelGGL(k_batch_fused, dim3(n), dim3(fused::LX, fused::NWV), 0, st, A, b->stage_seg0[s],
                         b->stage_seg0[s + 1] - 1, b->stage_item0[s], n);
    else
      launch_stage_rt<TZ, FVW>(s, n, st, A, b->stage_seg0[s], b->stage_seg0[s + 1] - 1, b->stage_item0[s]);
    ++launches;
  };
  hipError_t e = hipSuccess;
  auto fork = [&]() {
    if (e == hipSuccess) e = hipEventRecord(b->fork, b->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->side, b->fork, 0);
  };
  auto join = [&]() {
    if (e == hipSuccess) e = hipEventRecord(b->join, b->side);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join, 0);
  };
  static const int chain[] = {K_TRACTION, K_TRACTION_MIRROR, K_TABS, K_FRICTION, K_OUTPUTS, K_RATES, K_NUCLEATION};
  if (kind == 2) {
    for (int s : chain) stage(s, b->stream);
  } else {
    fork();
    if (kind == 1)
      for (int s : chain) stage(s, b->side);
    stage(K_VEL_FAULT, b->side);
    stage(K_VEL_INTERIOR, b->stream);
    stage(K_PML_VEL, b->stream);
    join();
    stage(K_FUSED, b->stream);
    stage(K_FUVW_UV, b->stream);
    stage(K_FUVW_W, b->stream);
    fork();
    stage(K_STR_FAULT, b->side);
    stage(K_STR_TOP, b->stream);
    stage(K_STR_INTERIOR, b->stream);
    stage(K_PML_STR, b->stream);
    join();
    stage(K_FRES, b->stream);
    stage(K_ADVANCE, b->stream);
  }
  if (e != hipSuccess) return bfail(b, FD3D_E_HIP, std::string("batch graph fork/join: ") + hipGetErrorString(e));
  if (kind == 1) b->launches = launches;   // a step in a run: every stage once
  return 0;
}
static_assert(K_ADVANCE + 8 == K_COUNT && K_NUCLEATION == K_COUNT - 1,
              "capture_staged_t names every stage: advance, then the seven fault-plane stages");

void drop_graphs(fd3d_batch *b) {
  for (int n = 0; n < 3; ++n) {
    if (b->exec[n]) { (void)hipGraphExecDestroy(b->exec[n]); b->exec[n] = nullptr; }
    if (b->graph[n]) { (void)hipGraphDestroy(b->graph[n]); b->graph[n] = nullptr; }
  }
}

int rebuild(fd3d_batch *b) {
  drop_graphs(b);
  std::vector<BatchDomain> doms;
  b->fvw = false;
  bool fused = b->mode == FD3D_BATCH_STAGED;
  if (fused) {
    // members that ask for the fused sweep get it when the batch's fused
    // items fill the GPU, otherwise every member runs the separate sweeps
    long long items = 0;
    for (fd3d_handle *h : b->mem) items += describe(h, b->tz, true).n[K_FUSED];
    int dev = 0, ncu = 0;
    BCHK(b, hipGetDevice(&dev));
    BCHK(b, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    fused = items >= (long long)b->fused_items * std::max(ncu, 1);
  }
  for (fd3d_handle *h : b->mem) {
    doms.push_back(describe(h, b->tz, fused));
    if (h->p.fv.fvw) b->fvw = true;
  }
  b->segs.clear();
  int total = 0;
  for (int s = 0; s < K_COUNT; ++s) {
    b->stage_seg0[s] = (int)b->segs.size();
    b->stage_item0[s] = total;
    for (int d = 0; d < (int)doms.size(); ++d) {
      if (doms[d].n[s] <= 0) continue;
      b->segs.push_back({total, d, s, 0});
      total += doms[d].n[s];
    }
  }
  b->stage_seg0[K_COUNT] = (int)b->segs.size();
  b->stage_item0[K_COUNT] = total;
  b->total = total;
  if (b->mem.empty()) {
    b->dirty = false;
    return 0;
  }
  if ((int)b->segs.size() > b->seg_cap) {
    if (b->fixed_dev >= 0)   // cannot happen: one segment per (member, stage) fits the fixed table
      return bfail(b, FD3D_E_INVALID, "batch: segment table of the fixed scratch exceeded");
    if (b->d_seg) (void)hipFree(b->d_seg);
    b->d_seg = nullptr;
    b->seg_cap = (int)b->segs.size() * 2;
    BCHK(b, hipMalloc(&b->d_seg, (size_t)b->seg_cap * sizeof(BatchSeg)));
  }
  BCHK(b, hipMemcpyAsync(b->d_dom, doms.data(), doms.size() * sizeof(BatchDomain), hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipMemcpyAsync(b->d_seg, b->segs.data(), b->segs.size() * sizeof(BatchSeg), hipMemcpyHostToDevice,
                         b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  b->dirty = false;   // the device descriptors are the members' (fd3d_batch_sync relies on it)
  int rc = 0;
  b->workers = 0;
  if (b->mode == FD3D_BATCH_PERSISTENT) {
    switch (b->tz * 2 + (b->fvw ? 1 : 0)) {
      case 8: rc = size_workers<4, false>(b); break;
      case 9: rc = size_workers<4, true>(b); break;
      case 16: rc = size_workers<8, false>(b); break;
      case 17: rc = size_workers<8, true>(b); break;
      case 32: rc = size_workers<16, false>(b); break;
      default: rc = size_workers<16, true>(b); break;
    }
    if (rc) { b->dirty = true; return rc; }
  }
  const int kinds = b->mode == FD3D_BATCH_PERSISTENT ? 1 : 3;
  for (int kind = 0; kind < kinds; ++kind) {
    const hipError_t eb = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (eb != hipSuccess) {
      drop_graphs(b);   // all graphs or none
      b->dirty = true;
      return bfail(b, FD3D_E_HIP, std::string("begin capture: ") + hipGetErrorString(eb));
    }
    const int key = b->tz * 2 + (b->fvw ? 1 : 0);
    if (b->mode == FD3D_BATCH_PERSISTENT) {
      switch (key) {
        case 8: rc = capture_t<4, false>(b); break;
        case 9: rc = capture_t<4, true>(b); break;
        case 16: rc = capture_t<8, false>(b); break;
        case 17: rc = capture_t<8, true>(b); break;
        case 32: rc = capture_t<16, false>(b); break;
        default: rc = capture_t<16, true>(b); break;
      }
    } else {
      switch (key) {
        case 8: rc = capture_staged_t<4, false>(b, kind); break;
        case 9: rc = capture_staged_t<4, true>(b, kind); break;
        case 16: rc = capture_staged_t<8, false>(b, kind); break;
        case 17: rc = capture_staged_t<8, true>(b, kind); break;
        case 32: rc = capture_staged_t<16, false>(b, kind); break;
        default: rc = capture_staged_t<16, true>(b, kind); break;
      }
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      drop_graphs(b);
      b->dirty = true;
      return rc ? rc : bfail(b, FD3D_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
    b->graph[kind] = g;
    e = hipGraphInstantiate(&b->exec[kind], g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graphs(b);
      b->dirty = true;
      return bfail(b, FD3D_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
  }
  return 0;
}

hipEvent_t take_event(fd3d_batch *b) {
  hipEvent_t e = nullptr;
  if (!b->pool.empty()) { e = b->pool.back(); b->pool.pop_back(); }
  else (void)hipEventCreate(&e);
  return e;
}

}  // namespace

hipStream_t fd3d_batch_stream_of(const fd3d_batch *b) { return b ? b->stream : nullptr; }

namespace {

// Device address of the fixed scratch, per device, looked up once with
// that device current; and whether a live batch holds it.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];
bool g_scratch_busy[kMaxDevices];

// Claims the current device's scratch; returns its device address or null.
void *claim_scratch(int &dev, int &code) {
  code = FD3D_E_HIP;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  if (g_scratch_busy[dev]) {
    code = FD3D_E_STATE;
    return nullptr;
  }
  g_scratch_busy[dev] = true;
  code = 0;
  return g_scratch_ptr[dev];
}

// The batch's stream, the side stream of its captured forks and the fork
// and join events; nonzero on failure.
int make_streams(fd3d_batch *b) {
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
    return 1;
  return 0;
}

void release_scratch(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev >= 0 && dev < kMaxDevices) g_scratch_busy[dev] = false;
}

}  // namespace

extern "C" {

int fd3d_batch_create(int max_problems, int tz, int mode, fd3d_batch **out) {
  if (!out) return FD3D_E_INVALID;
  *out = nullptr;
  if (max_problems < 1 || (tz != 4 && tz != 8 && tz != 16) ||
      (mode != FD3D_BATCH_PERSISTENT && mode != FD3D_BATCH_STAGED))
    return FD3D_E_INVALID;
  // The persistent form stalled on MI355X (its first steps did not finish
  // within two minutes on a five-problem batch, twice); until the cause of
  // the stall is found it is not offered.
  if (mode == FD3D_BATCH_PERSISTENT) return FD3D_E_STATE;
  fd3d_batch *b = new fd3d_batch();
  b->maxdom = max_problems;
  b->tz = tz;
  b->mode = mode;
  auto bail = [&](int code) { fd3d_batch_destroy(b); return code; };
  if (make_streams(b)) return bail(FD3D_E_HIP);
  if (hipMalloc(&b->d_dom, (size_t)max_problems * sizeof(BatchDomain)) != hipSuccess) return bail(FD3D_E_NOMEM);
  const size_t ncnt = kCntBase + (size_t)max_problems * K_COUNT;
  if (hipMalloc(&b->d_cnt, (ncnt + 1 + 2 * (size_t)max_problems) * sizeof(unsigned)) != hipSuccess)
    return bail(FD3D_E_NOMEM);
  b->d_rep = b->d_cnt + ncnt;
  b->rep.assign(1 + 2 * (size_t)max_problems, 0u);
  if (hipMemsetAsync(b->d_cnt, 0, ncnt * sizeof(unsigned), b->stream) != hipSuccess) return bail(FD3D_E_HIP);
  if (hipStreamSynchronize(b->stream) != hipSuccess) return bail(FD3D_E_HIP);
  *out = b;
  return 0;
}

int fd3d_batch_create_static(int max_problems, int tz, fd3d_batch **out) {
  if (!out) return FD3D_E_INVALID;
  *out = nullptr;
  if (max_problems < 1 || max_problems > kStaticMaxProblems || (tz != 4 && tz != 8 && tz != 16))
    return FD3D_E_INVALID;
  int dev = -1, code = 0;
  void *base = claim_scratch(dev, code);
  if (!base) return code;
  fd3d_batch *b = new fd3d_batch();
  b->maxdom = max_problems;
  b->tz = tz;
  b->mode = FD3D_BATCH_STAGED;
  b->fixed_dev = dev;
  char *c = static_cast<char *>(base);
  b->d_dom = reinterpret_cast<BatchDomain *>(c);
  b->d_seg = reinterpret_cast<BatchSeg *>(c + kStaticDomBytes);
  b->d_cnt = reinterpret_cast<unsigned *>(c + kStaticDomBytes + kStaticSegBytes);
  b->d_rep = b->d_cnt + kCntBase + K_COUNT * kStaticMaxProblems;
  b->rep.assign(1 + 2 * (size_t)max_problems, 0u);
  b->seg_cap = K_COUNT * kStaticMaxProblems;
  auto bail = [&](int rc) { fd3d_batch_destroy(b); return rc; };
  if (make_streams(b)) return bail(FD3D_E_HIP);
  // the last batch's counters are still in the array: clear them
  const int ncnt = kCntBase + K_COUNT * kStaticMaxProblems;
  hipLaunchKernelGGL(k_zero_u32, dim3(cdiv(ncnt, 256)), dim3(256), 0, b->stream, b->d_cnt, ncnt);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) return bail(FD3D_E_HIP);
  *out = b;
  return 0;
}

int fd3d_batch_add(fd3d_batch *b, fd3d_handle *h) {
  if (!b || !h) return FD3D_E_INVALID;
  if (!h->inited) return bfail(b, FD3D_E_STATE, "fd3d_init_state the problem first");
  if (h->batch) return bfail(b, FD3D_E_STATE, "the problem already belongs to a batch");
  if (h->begun != h->waited) return bfail(b, FD3D_E_STATE, "the problem has output chunks in flight");
  if ((int)b->mem.size() >= b->maxdom) return bfail(b, FD3D_E_INVALID, "batch is full");
  BCHK(b, hipStreamSynchronize(h->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  b->mem.push_back(h);
  h->batch = b;
  b->dirty = true;
  return 0;
}

int fd3d_batch_remove(fd3d_batch *b, fd3d_handle *h) {
  if (!b || !h) return FD3D_E_INVALID;
  auto it = std::find(b->mem.begin(), b->mem.end(), h);
  if (it == b->mem.end()) return bfail(b, FD3D_E_INVALID, "not a member of this batch");
  BCHK(b, hipStreamSynchronize(b->stream));
  b->mem.erase(it);
  h->batch = nullptr;
  b->dirty = true;
  return 0;
}

int fd3d_batch_size(const fd3d_batch *b) { return b ? (int)b->mem.size() : 0; }

int fd3d_batch_step(fd3d_batch *b, int nsteps) {
  if (!b || nsteps < 0) return FD3D_E_INVALID;
  if (b->mem.empty() || nsteps == 0) return 0;
  if (b->dirty) {
    int rc = rebuild(b);
    if (rc) return rc;
  }
  hipEvent_t e0 = take_event(b), e1 = take_event(b);
  BCHK(b, hipEventRecord(e0, b->stream));
  if (b->mode == FD3D_BATCH_PERSISTENT) {
    for (int s = 0; s < nsteps; ++s) BCHK(b, hipGraphLaunch(b->exec[0], b->stream));
  } else {
    BCHK(b, hipGraphLaunch(b->exec[0], b->stream));
    for (int s = 1; s < nsteps; ++s) BCHK(b, hipGraphLaunch(b->exec[1], b->stream));
    BCHK(b, hipGraphLaunch(b->exec[2], b->stream));
  }
  BCHK(b, hipEventRecord(e1, b->stream));
  b->evs.push_back({e0, e1, nsteps});
  for (fd3d_handle *h : b->mem) h->steps += nsteps;
  return 0;
}

int fd3d_batch_sync(fd3d_batch *b) {
  if (!b) return FD3D_E_INVALID;
  const int n = (int)b->mem.size();
  if (b->dirty) {
    // membership changed since the last step: the device descriptors are
    // not the members', so read each member's control block
    BCHK(b, hipStreamSynchronize(b->stream));
    unsigned flag = 0;
    BCHK(b, hipMemcpyAsync(&flag, b->d_cnt + 1, sizeof(unsigned), hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    if (flag) return bfail(b, FD3D_E_HIP, "batch: a stage dependency wait ran out of time");
    for (int q = 0; q < n; ++q) {
      fd3d_handle *h = b->mem[q];
      int rc = fd3d_handle_check_ctl(h, b->stream);
      if (rc) return bfail(b, rc, "member " + std::to_string(q) + ": " + h->err);
    }
    return 0;
  }
  // one gather kernel and one copy after the queued steps, instead of a
  // round trip per member
  hipLaunchKernelGGL(k_batch_report, dim3(cdiv(std::max(n, 1), 64)), dim3(64), 0, b->stream, b->d_dom, n, b->d_cnt,
                     b->d_rep);
  BCHK(b, hipGetLastError());
  BCHK(b, hipMemcpyAsync(b->rep.data(), b->d_rep, (1 + 2 * (size_t)n) * sizeof(unsigned), hipMemcpyDeviceToHost,
                         b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  if (b->rep[0]) return bfail(b, FD3D_E_HIP, "batch: a stage dependency wait ran out of time");
  for (int q = 0; q < n; ++q) {
    if (b->rep[1 + 2 * q] == 0u && b->rep[2 + 2 * q] == 0u) continue;
    fd3d_handle *h = b->mem[q];
    int rc = fd3d_handle_check_ctl(h, b->stream);   // the member's error code and message
    if (rc) return bfail(b, rc, "member " + std::to_string(q) + ": " + h->err);
  }
  return 0;
}

// The fused stage runs one 1024-lane workgroup per CU at a time; with few
// items per CU its last wave leaves CUs idle for longer than the saved
// bytes gain (measured: a 64-domain batch of sides 96-320, about 30 items
// per CU, steps 7% faster fused; the first 48 served requests, 2-4 items
// per CU, 5% slower).  The default threshold is 8 items per CU.
int fd3d_batch_set_fused_items(fd3d_batch *b, int min_items_per_cu) {
  if (!b || min_items_per_cu < 0) return FD3D_E_INVALID;
  BCHK(b, hipStreamSynchronize(b->stream));
  b->fused_items = min_items_per_cu;
  b->dirty = true;
  return 0;
}

int fd3d_batch_info(const fd3d_batch *b, int *out4) {
  if (!b || !out4) return FD3D_E_INVALID;
  out4[0] = b->workers;
  out4[1] = b->total;
  out4[2] = (int)b->segs.size();
  out4[3] = b->launches;
  return 0;
}

double fd3d_batch_elapsed(fd3d_batch *b, int *nsteps) {
  if (nsteps) *nsteps = 0;
  if (!b) return 0.0;
  if (hipStreamSynchronize(b->stream) != hipSuccess) return 0.0;
  double ms = 0.0;
  int n = 0;
  for (auto &e : b->evs) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, e.a, e.b) == hipSuccess) { ms += t; n += e.n; }
    b->pool.push_back(e.a);
    b->pool.push_back(e.b);
  }
  b->evs.clear();
  if (nsteps) *nsteps = n;
  return ms;
}

const char *fd3d_batch_last_error(const fd3d_batch *b) { return b ? b->err.c_str() : "null batch"; }

void fd3d_batch_destroy(fd3d_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (fd3d_handle *h : b->mem) h->batch = nullptr;
  drop_graphs(b);
  for (auto &e : b->evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
  for (auto e : b->pool) (void)hipEventDestroy(e);
  if (b->fixed_dev >= 0) {
    release_scratch(b->fixed_dev);   // the tables live in the fixed scratch
  } else {
    if (b->d_dom) (void)hipFree(b->d_dom);
    if (b->d_seg) (void)hipFree(b->d_seg);
    if (b->d_cnt) (void)hipFree(b->d_cnt);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"
