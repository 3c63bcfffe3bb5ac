/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * A CPU restatement of the FD3D_TSN time loop (src/fd3d_theo.f90:187-602
 * and the kernels of src/fd3d_deriv.f90 it calls).  It is the checker for
 * the HIP product path: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.  Nothing in fd3d_tsn_amd/ links it.
 *
 * Precision modes (prec argument of orc_create):
 *   ORC_F32   all arithmetic in float  (the reference's default-REAL build)
 *   ORC_F64   all arithmetic in double (the reference built -fdefault-real-8,
 *             the analogue of -r8 in examples/TPV104/compile.sh:13)
 *   ORC_MIXED wavefield and slip-weakening arithmetic in float, the
 *             rate-and-state Newton solve, state-variable update and
 *             friction value in double (what the HIP path computes)
 *
 * Every array crossing this API is double (lossless for float values) and
 * laid out like the reference's Fortran arrays: x fastest, then y, then z,
 * with the 1-based Fortran index (i,j,k) at offset
 * (i-1) + nxt*((j-1) + nyt*(k-1)).  2D fault arrays are (nxt,nzt).
 *
 * Parity status of each routine (what pins it) is stated at the routine in
 * fd3d_oracle.cpp.
 */
#ifndef FD3D_ORACLE_H
#define FD3D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_F32 = 0, ORC_F64 = 1, ORC_MIXED = 2 };

/* integer parameters, index into ip[] */
enum {
  ORC_NXT = 0, ORC_NYT, ORC_NZT, ORC_NABC, ORC_FSPACE, ORC_FVW, ORC_NUCL,
  ORC_NSTA, ORC_NEWTON_MAX, ORC_NIP
};
/* real parameters, index into rp[] */
enum {
  ORC_DH = 0, ORC_DT, ORC_DAMP_S, ORC_PML_VP, ORC_PML_FACT,
  ORC_SN, ORC_UINI, ORC_WINI, ORC_F0, ORC_V0, ORC_FW,
  ORC_HX0, ORC_HZ0, ORC_PERTURB, ORC_RR2, ORC_TT2, ORC_STRINIXI, ORC_NRP
};

/* output slab ids, in the order of the per-step .res streams */
enum { ORC_OUT_SRX = 0, ORC_OUT_SRZ, ORC_OUT_SCX, ORC_OUT_SCZ, ORC_OUT_PSI, ORC_NOUT };

/* 3D field ids for orc_get_field */
enum { ORC_U = 0, ORC_V, ORC_W, ORC_XX, ORC_YY, ORC_ZZ, ORC_XY, ORC_YZ, ORC_XZ };
/* 2D fault-plane ids for orc_get_plane */
enum {
  ORC_P_TX = 0, ORC_P_TZ, ORC_P_RFX, ORC_P_RFZ, ORC_P_AVDX, ORC_P_AVDZ,
  ORC_P_TABSX, ORC_P_TABSZ, ORC_P_UZ, ORC_P_WX, ORC_P_DISTX, ORC_P_DISTZ,
  ORC_P_PSIX, ORC_P_PSIZ, ORC_P_T0X, ORC_P_T0Z, ORC_P_V1T,
  ORC_P_RUPTIME, ORC_P_RISE, ORC_P_SLIPTIME, ORC_P_SLIPX, ORC_P_SLIPZ,
  ORC_P_SCHANGEX, ORC_P_SCHANGEZ, ORC_NPLANE
};

void *orc_create(int prec, const int *ip, const double *rp, const int *sta /* 3*nsta, 1-based */);
void orc_destroy(void *h);
/* lam1, mu1, d1: nxt*nyt*nzt */
int orc_set_medium(void *h, const double *lam, const double *mu, const double *d);
/* slip weakening inputs at Z-nodes (fd3d_init.f90:191 module arrays) */
int orc_set_fault_sw(void *h, const double *striniX, const double *striniZ,
                     const double *peak_xz, const double *dyn_xz,
                     const double *Dc, const double *coh);
/* rate-and-state inputs (fd3d_init.f90:182-186 module arrays) */
int orc_set_fault_fvw(void *h, const double *Dc, const double *striniX,
                      const double *striniZ, const double *a, const double *b,
                      const double *psi, const double *vw, const double *aX,
                      const double *bX, const double *psiX, const double *vwX);
/* fd3d() prologue: zero state, PML profiles, fault-parameter staggering */
int orc_init(void *h);
/* advance nsteps; out: nsteps*ORC_NOUT*nxtT*nztT (may be NULL),
   seis: nsteps*nsta*3 (may be NULL).  Returns the number of Newton solves
   that hit the iteration cap (0 when converged), or <0 on error. */
long orc_step(void *h, int nsteps, double *out, double *seis);
/* epilogue arrays after the loop: ruptime, rise(-ruptime), sliptime, slipX,
   slipZ, schangeX-T0X, schangeZ-T0Z; 7*nxt*nzt (fd3d_theo.f90:606-607,658) */
int orc_final(void *h, double *out7);
int orc_get_field(void *h, int which, double *out);
int orc_get_plane(void *h, int which, double *out);
/* overwrite a 3D field (tests that start from random wavefields) */
int orc_set_field(void *h, int which, const double *in);

#ifdef __cplusplus
}
#endif
#endif
