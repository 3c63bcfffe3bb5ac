/*
 * fd3d_tsn.h -- C ABI of the MI355X time-stepping core for FD3D_TSN.
 *
 * Replaces the reference's time loop SUBROUTINE fd3d() (src/fd3d_theo.f90:22,
 * called from src/dynamicsolver.f90:66) together with the kernels it calls
 * in src/fd3d_deriv.f90.  The reference passes all state through Fortran
 * MODULE arrays (src/fd3d_init.f90:15-124); here the caller hands the same
 * arrays over once as plain pointers, in the same column-major layout
 * (x fastest, then y, then z; element (i,j,k), 1-based, at
 * (i-1) + nxt*((j-1) + nyt*(k-1))), and the library keeps the whole
 * wavefield resident in HBM for the lifetime of the handle.
 *
 * Every function returns 0 on success or a negative FD3D_E* code; the
 * message of the last failure on a handle is fd3d_last_error(h).
 * A handle is used from one host thread at a time.  Calls on different
 * handles are independent (each handle owns its HIP stream).
 *
 * Fortran binding: every entry point takes plain pointers and integers by
 * value, so an interface block with bind(C) and type(c_ptr) / value
 * arguments maps 1:1 (see fd3d_tsn_amd/driver/fd3d_tsn_mod.f90).
 */
#ifndef FD3D_TSN_H
#define FD3D_TSN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FD3D_OK 0
#define FD3D_E_INVALID (-1)      /* bad configuration or argument          */
#define FD3D_E_CFL (-2)          /* Vp*dt/dh > 0.25 (fd3d_init.f90:274-278) */
#define FD3D_E_NOMEM (-3)        /* device memory exhausted                */
#define FD3D_E_HIP (-4)          /* HIP runtime error                      */
#define FD3D_E_STATE (-5)        /* call out of order (e.g. step before init) */
#define FD3D_E_NEWTON (-6)       /* rate-and-state Newton hit its iteration cap */
#define FD3D_E_NONFINITE (-7)    /* NaN/Inf detected in the fault outputs  */

enum fd3d_friction { FD3D_SW = 0, FD3D_FVW = 1 };

/* Per-step fault output slabs, in the order of the reference's .res streams
 * (fd3d_theo.f90:548-553).  Each slab is nxtT*nztT floats, x fastest,
 * covering i in [nabc+1, nxt-nabc], k in [nabc+1, nzt-nfs]. */
enum fd3d_output {
  FD3D_OUT_SLIPRATE_X = 0,   /* result/sliprateX.res    */
  FD3D_OUT_SLIPRATE_Z = 1,   /* result/sliprateZ.res    */
  FD3D_OUT_SHEAR_X = 2,      /* result/shearstressX.res */
  FD3D_OUT_SHEAR_Z = 3,      /* result/shearstressZ.res */
  FD3D_OUT_PSI = 4,          /* result/psi.res (rate-and-state only) */
  FD3D_NOUT = 5
};

/* 3D wavefield ids for fd3d_download_field / fd3d_upload_field */
enum fd3d_field {
  FD3D_U = 0, FD3D_V, FD3D_W, FD3D_XX, FD3D_YY, FD3D_ZZ, FD3D_XY, FD3D_YZ, FD3D_XZ
};

/* Final 2D arrays (nxt*nzt each), fd3d_theo.f90:604-658 */
enum fd3d_final {
  FD3D_RUPTIME = 0, FD3D_RISE, FD3D_SLIPTIME, FD3D_SLIP_X, FD3D_SLIP_Z,
  FD3D_STRESSDROP_X, FD3D_STRESSDROP_Z, FD3D_NFINAL
};

/* Problem definition: the values fd3d_init() reads from inputfd3d.dat
 * (fd3d_init.f90:145-171) plus the macro choices of the reference build
 * (-DFSPACE, -DFVW, -DTPV104). */
typedef struct fd3d_config {
  int nxt, nyt, nzt;      /* full grid, PML and ghost layers included          */
  int nabc;               /* PML thickness (>= 3)                              */
  int fspace;             /* 1: full space (nfs = nabc, z-high PML); 0: free surface (nfs = 2) */
  int friction;           /* FD3D_SW or FD3D_FVW                               */
  int nucleation;         /* FVW only: TPV104 smooth nucleation in time (fd3d_theo.f90:515-542) */
  int nstations;          /* stations recorded every step (u, v, w)           */
  const int *stations;    /* 3*nstations 1-based grid indices (staX, staY, staZ) */
  float dh, dt;           /* grid step [m], time step [s]                      */
  float damp_s;           /* Dalguer-Day damping near the fault                */
  float pml_vp, pml_fact; /* PML: omegaM = pml_fact*pml_vp/(2*dh*(nabc-1))      */
  /* rate-and-state scalars (friction_com, fd3d_init.f90:61), double precision */
  double Sn, uini, wini, f0, v0, fw;
  double hx0, hz0, perturb, RR2, TT2, strinixI;
  int newton_max;         /* Newton iteration cap (0 -> 200); the reference has none */
  int ring_steps;         /* device output ring length in steps (0 -> 64)      */
} fd3d_config;

typedef struct fd3d_handle fd3d_handle;

/* Replaces the allocation half of fd3d() (fd3d_theo.f90:53-57) and
 * init_pml's allocations (fd3d_deriv.f90:1319-1437).  Allocates every
 * device array; fails with FD3D_E_NOMEM when HBM does not hold the
 * problem.  Only checks the configuration; the CFL check needs the medium
 * (fd3d_upload_medium). */
int fd3d_create(const fd3d_config *cfg, fd3d_handle **out);

/* Device bytes fd3d_create would allocate for cfg (admission control). */
size_t fd3d_footprint(const fd3d_config *cfg);

/* Free and total device memory of the current GPU in bytes (admission of
 * new problems against free HBM); the reference has no counterpart. */
int fd3d_device_memory(size_t *free_bytes, size_t *total_bytes);

/* medium_com arrays lam1, mu1, d1 (fd3d_init.f90:262-264), nxt*nyt*nzt each.
 * The reference fills them plane by plane from the layered crustal model
 * (fd3d_init.f90:240-265); the device keeps one value per z plane, so a
 * medium that varies within a plane is refused with FD3D_E_INVALID.
 * Returns FD3D_E_CFL when max(Vp)*dt/dh > 0.25 (fd3d_init.f90:274-278);
 * FD3D_E_STATE while the problem is a batch member. */
int fd3d_upload_medium(fd3d_handle *h, const float *lam1, const float *mu1, const float *d1);

/* Slip-weakening fault inputs at Z-nodes, nxt*nzt each: striniX, striniZ,
 * peak_xz, dyn_xz, Dc, coh (friction_com, fd3d_init.f90:191). */
int fd3d_upload_fault_sw(fd3d_handle *h, const float *striniX, const float *striniZ,
                         const float *peak_xz, const float *dyn_xz, const float *Dc,
                         const float *coh);

/* Rate-and-state inputs, nxt*nzt each (fd3d_init.f90:182-186): Dc, striniX,
 * striniZ, then a, b, psi, vw at Z-nodes and aX, bX, psiX, vwX at X-nodes.
 * Double precision: the Newton solve and state update run in FP64. */
int fd3d_upload_fault_fvw(fd3d_handle *h, const double *Dc, const double *striniX,
                          const double *striniZ, const double *a, const double *b,
                          const double *psi, const double *vw, const double *aX,
                          const double *bX, const double *psiX, const double *vwX);

/* The prologue of fd3d() (fd3d_theo.f90:59-88): zero the wavefield and the
 * fault state, build the PML profiles (init_pml) and the staggered fault
 * parameters (interp_fric).  May be called again to restart. */
int fd3d_init_state(fd3d_handle *h);

/* Advance nsteps time steps (the body of fd3d_theo.f90:187-602).
 * out:  NULL, or room for nsteps*FD3D_NOUT*nxtT*nztT floats, filled step by
 *       step with the output slabs in fd3d_output order.
 * seis: NULL, or room for nsteps*nstations*3 floats (u, v, w per station).
 * Kernels run asynchronously on the handle's stream; the call returns after
 * the last step's outputs have reached the host (or immediately when both
 * pointers are NULL: use fd3d_sync). */
int fd3d_step(fd3d_handle *h, int nsteps, float *out, float *seis);

/* Wait for queued steps; reports FD3D_E_NEWTON / FD3D_E_NONFINITE. */
int fd3d_sync(fd3d_handle *h);

/* Overlapped output stream.  The reference copies the fault slabs to the
 * host and writes the result/ .res files every step (fd3d_theo.f90:199-204,545-558),
 * with the GPU idle meanwhile.  Here the caller writes one chunk of steps
 * while the GPU computes the next:
 *   fd3d_step_begin(h, n)   advance n steps (1 <= n <= ring_steps/2) and
 *                           queue the copy of their output slabs, station
 *                           samples and error flags into one of two pinned
 *                           host buffers owned by the handle, on a copy
 *                           stream of its own; returns without waiting.
 *                           At most two chunks are in flight (FD3D_E_STATE).
 *   fd3d_step_wait(h, &n, &out, &seis)
 *                           wait for the oldest chunk's copy; *out points to
 *                           n*FD3D_NOUT*nxtT*nztT floats in the layout of
 *                           fd3d_step's out, *seis to n*nstations*3 floats.
 *                           Valid until the fd3d_step_begin after the next
 *                           (the buffer is then reused).  Returns
 *                           FD3D_E_NEWTON / FD3D_E_NONFINITE from the flags
 *                           after that chunk.
 * Usage: begin(c0); begin(c1); loop { wait -> write; begin(next) }; wait the
 * rest.  fd3d_step, fd3d_init_state and fd3d_batch_add refuse a handle with
 * chunks in flight. */
int fd3d_step_begin(fd3d_handle *h, int nsteps);
int fd3d_step_wait(fd3d_handle *h, int *nsteps, const float **out, const float **seis);

/* Steps taken since fd3d_init_state. */
int fd3d_steps_done(const fd3d_handle *h);

/* Copy the output slabs (out, FD3D_NOUT*nxtT*nztT floats per step) and
 * station samples (seis, nstations*3 floats per step) of the last nsteps
 * steps from the device ring, oldest first; nsteps <= ring_steps.  Either
 * pointer may be NULL.  The same data fd3d_step hands back per call, for
 * problems advanced by fd3d_batch_step (the reference writes them every
 * step, fd3d_theo.f90:546-557). */
int fd3d_read_outputs(fd3d_handle *h, int nsteps, float *out, float *seis);

/* The arrays fd3d() leaves in source_com after the loop (fd3d_theo.f90:604-658),
 * in fd3d_final order, each nxt*nzt: ruptime, rise (minus ruptime), sliptime,
 * slipX, slipZ, schangeX-T0X, schangeZ-T0Z.  out7 holds 7*nxt*nzt floats. */
int fd3d_download_final(fd3d_handle *h, float *out7);

/* Copy one 3D field to / from the host (nxt*nyt*nzt floats, reference layout). */
int fd3d_download_field(fd3d_handle *h, int which, float *out);
int fd3d_upload_field(fd3d_handle *h, int which, const float *in);

/* Source-time bins for dynamic inversion, accumulated on the device every
 * step as fd3d_theo.f90:560-578 does into SlipRates_com (MSRX, MSRZ,
 * MomentRate; allocated in inversion_init, fd3d_init.f90:786-800):
 *   bin k = int((it-1)*dt/dtseis)+1 of nbins (the reference: dtseis = 1 s,
 *   nbins = nSR = ntfd);
 *   msrx/msrz[(patch)*nbins + k-1] += mean slip rate over patch (i, j) of an
 *   NL x NW grid of dL x dW patches, times dt/dtseis;
 *   moment_rate[k-1] += sum over the fault of |slip rate| * mu * dh^2 * dt/dtseis.
 * Call before fd3d_init_state (which zeroes them); NL = NW = 0 keeps only
 * the moment rate.  The patch sums follow the reference's float order; the
 * moment-rate sum over the fault is a fixed-order double-precision tree
 * (the reference adds node by node in float). */
int fd3d_set_source_bins(fd3d_handle *h, int NL, int NW, float dL, float dW, float dtseis, int nbins);
/* msrx, msrz: NL*NW*nbins floats each (or NULL); moment_rate: nbins floats (or NULL). */
int fd3d_download_source_bins(fd3d_handle *h, float *msrx, float *msrz, float *moment_rate);

/* Launch shapes of the 3D sweeps.  mode bit 0: interior sweeps where a
 * block owns a 64 x tz tile of the x-z plane and marches ny rows along y
 * (tz in {4, 8, 16}, ny >= 1), else one thread per cell; mode bit 1: PML
 * sweeps as 64 x 8 x-z tiles marching y, else one thread per cell; mode
 * bit 2: one fused velocity + stress pass over the interior (ny_vel rows
 * per block; under a free surface the top two interior planes get their
 * stresses after the ghost velocities, in a small pass of their own); mode
 * bits 3 and 4 (with bit 0): velocity and stress tiles 128 wide instead of
 * 64 (tz in {4, 8}); mode bits 5-6 (with bit 2): the fused sweep walks its
 * tiles in z within strips 4 (0), 2 (1), 8 (2) or 3 (3) x-tiles wide.
 * fd3d_create picks a default from the grid size (the
 * fused sweep for grids whose 64 x 11 x-z tiles times interior rows reach
 * 8192, such as 256x256x256 and up, in strips of 2 tiles when that makes
 * each of the 8 XCDs' tile ranges exactly one strip, else of 4);
 * fd3d_tsn_amd/autotune.py measures and caches the best.
 * Results do not depend on the shape (bit-identical). */
int fd3d_set_tuning(fd3d_handle *h, int mode, int tz_vel, int ny_vel, int tz_str, int ny_str);
int fd3d_get_tuning(const fd3d_handle *h, int *out5);

/* Per-kernel timing with HIP events on the handle's stream (0 = off).
 * fd3d_kernel_times fills up to n entries of names/ms/launches for the
 * kernels launched since profiling was switched on; returns the count. */
int fd3d_set_profiling(fd3d_handle *h, int on);
int fd3d_kernel_times(fd3d_handle *h, int n, const char **names, double *ms, long *launches);

/* Algorithmic HBM bytes of one step of this problem: 120 B per bulk cell
 * and 276 B per PML cell (fault plane excluded), and the cell counts. */
double fd3d_step_bytes(const fd3d_handle *h, long *bulk_cells, long *pml_cells);

const char *fd3d_last_error(const fd3d_handle *h);

/* Release the handle and its device memory (the deallocations at
 * fd3d_theo.f90:609-625); removes it from its batch first. */
void fd3d_destroy(fd3d_handle *h);

/* ------------------------------------------------------------------------
 * Batched problems.  The reference runs one rupture per process
 * (dynamicsolver.f90:66 calls fd3d() once); a dynamic inversion runs many
 * such forward models back to back.  A batch advances many resident
 * problems of different sizes, friction laws and free-surface settings
 * together (FD3D_BATCH_STAGED): each stage of a step is one launch covering
 * every problem, whose workgroups take (problem, stage, block) items from a
 * table; the launches of a step are captured in HIP graphs and replayed,
 * with each step's fault-plane stages beside the next step's velocity
 * stages.  Results are bit-identical to fd3d_step on each problem alone.
 * ------------------------------------------------------------------------ */
typedef struct fd3d_batch fd3d_batch;

enum fd3d_batch_mode {
  FD3D_BATCH_PERSISTENT = 0,  /* one persistent launch per step, dependency counters per problem;
                                 currently refused with FD3D_E_STATE (stalled on MI355X) */
  FD3D_BATCH_STAGED = 1       /* one launch per stage covering every problem */
};

/* tz: x-z tile height of the interior sweeps of every member (4, 8 or 16);
 * each member keeps its own rows-per-tile (ny_vel, ny_str of fd3d_set_tuning).
 * mode FD3D_BATCH_PERSISTENT is refused with FD3D_E_STATE for now. */
int fd3d_batch_create(int max_problems, int tz, int mode, fd3d_batch **out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (a serving loop that must not call hipMalloc between requests,
 * code under HIP graph capture, harnesses that forbid allocation): the same
 * batch as fd3d_batch_create(max_problems, tz, FD3D_BATCH_STAGED, out),
 * whose member descriptors, segment table and counters live in a fixed
 * device array of the library (one per GPU, 125.5 KiB, held until the
 * process ends) instead of device allocations.  Creating, adding, removing,
 * stepping and destroying such a batch call no hipMalloc or hipFree; the
 * members' own memory comes from fd3d_create as before.  At most 64
 * problems (max_problems > 64: FD3D_E_INVALID).  One such batch per GPU at
 * a time: the array is shared by every caller on that GPU, so a second one
 * while the first lives gets FD3D_E_STATE (overlapping use would race).
 * Results are bit-identical to the allocating batch. */
int fd3d_batch_create_static(int max_problems, int tz, fd3d_batch **out);

/* Add an initialised problem (fd3d_init_state done).  While it is a member,
 * fd3d_step, fd3d_set_tuning and fd3d_init_state on it fail with
 * FD3D_E_STATE; downloads and fd3d_read_outputs wait for the batch. */
int fd3d_batch_add(fd3d_batch *b, fd3d_handle *h);
int fd3d_batch_remove(fd3d_batch *b, fd3d_handle *h);
int fd3d_batch_size(const fd3d_batch *b);

/* Advance every member nsteps time steps (asynchronous; fd3d_batch_sync). */
int fd3d_batch_step(fd3d_batch *b, int nsteps);

/* Wait for queued steps; reports a failed dependency wait (FD3D_E_HIP) and
 * the members' FD3D_E_NEWTON / FD3D_E_NONFINITE flags, the message naming
 * the first failing member (its position in the order of fd3d_batch_add).
 * fd3d_sync on a member reports that member's own flags, so a caller can
 * remove the failed members and go on stepping the others. */
int fd3d_batch_sync(fd3d_batch *b);

/* Members whose tuning asks for the fused interior sweep (fd3d_set_tuning
 * mode bit 2; ny_vel rows per work item) run it as a stage of the batch
 * (under a free surface with the top-plane stresses as a stage of their
 * own) when the batch's fused work items reach
 * min_items_per_cu per compute unit (default 8; 0: whenever a member asks);
 * otherwise every member runs the separate sweeps, such members with ny_str
 * rows for both.  Results are bit-identical either way.  Marks the step
 * graph for re-capture. */
int fd3d_batch_set_fused_items(fd3d_batch *b, int min_items_per_cu);

/* out4: resident workgroups of the persistent launch, work items per step,
 * (problem, stage) segments per step, kernel launches per step. */
int fd3d_batch_info(const fd3d_batch *b, int *out4);

/* Milliseconds of GPU time of the steps since the last call (HIP events
 * around each replayed step graph); *nsteps receives the step count. */
double fd3d_batch_elapsed(fd3d_batch *b, int *nsteps);

const char *fd3d_batch_last_error(const fd3d_batch *b);
void fd3d_batch_destroy(fd3d_batch *b);

#ifdef __cplusplus
}
#endif
#endif
