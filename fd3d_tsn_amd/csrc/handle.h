// Host-side state of one problem (fd3d_handle), shared by the single-problem
// C ABI (capi.hip) and the multi-domain batch engine (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fd3d_tsn.h"
#include "fd3d_dev.h"

using namespace fd3d;

struct fd3d_handle {
  fd3d_config cfg{};
  std::vector<int> sta;
  Problem p{};
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  size_t bytes = 0;
  bool medium_ok = false, fault_ok = false, inited = false;
  int steps = 0;
  std::string err;
  // host copies of the fault inputs (staggering is done on the host)
  std::vector<float> h_striniX, h_striniZ, h_peak, h_dyn, h_Dc, h_coh;
  std::vector<double> h_a, h_b, h_psi, h_vw, h_aX, h_bX, h_psiX, h_vwX;
  int *d_sta = nullptr;
  std::vector<float> staging;   // padded host copy of one 3D field during transfers
  // profiling
  bool prof = false;
  struct Ev { int id; hipEvent_t a, b; };
  std::vector<Ev> evs;
  std::vector<hipEvent_t> pool;
  double kms[K_COUNT] = {0};
  long kcount[K_COUNT] = {0};
  // step sequence, and three graphs of it (replayed by fd3d_step when
  // profiling is off; rebuilt when the sequence changes): a run of n steps
  // is the head graph (a step up to advance), n-1 rotated graphs (the
  // previous step's fault-plane kernels beside this step's velocity
  // sweeps, then this step up to advance) and the tail graph (the last
  // step's fault-plane kernels)
  std::vector<KernelId> seq;
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  // second stream and fork/join events of the captured step: the kernels of
  // a half-step that write disjoint cells run side by side
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // the batch this problem is advanced in (fd3d_batch_add), or null
  fd3d_batch *batch = nullptr;
  // overlapped output stream (fd3d_step_begin / fd3d_step_wait): two pinned
  // host buffers, each holding one chunk's slabs, station samples and the
  // control block after the chunk; the copies run on their own stream
  struct Chunk {
    float *host = nullptr;       // [ring/2][FD3D_NOUT][slab] slabs, then [ring/2][nsta][3] samples
    Ctl *ctl = nullptr;          // control block after the chunk
    hipEvent_t computed = nullptr, copied = nullptr;
    int first = 0, n = 0;        // first step (1-based) and step count of the chunk
  };
  Chunk chunk[2];
  hipStream_t copy = nullptr;
  long long begun = 0, waited = 0;   // chunks begun and handed out, in order
};


// The stream a batch runs its steps on (batch.hip).
hipStream_t fd3d_batch_stream_of(const fd3d_batch *b);

// Reads the device control block of h on stream s and turns its flags into
// FD3D_E_NEWTON / FD3D_E_NONFINITE (capi.hip).
int fd3d_handle_check_ctl(fd3d_handle *h, hipStream_t s);
