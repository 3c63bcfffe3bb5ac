! Test infrastructure only: a driver PROGRAM that links the unmodified
! reference sources (fd3d_init.f90, fd3d_deriv.f90, fd3d_theo.f90 from
! the FD3D_TSN tree) in place of src/dynamicsolver.f90, so that the
! reference time loop fd3d() can run on seeded per-node fault fields.
!
! The reference's own setup routines cannot express seeded random stress
! or friction fields (forwardspecialTPV5 only knows rectangular patches,
! readinversionresult forces dyn_xz = coh = 0, fd3d_init.f90:830-835), so
! this program reads the fault fields from 'fault_in.bin' after
! fd3d_init()/inversion_init() (dynamicsolver.f90:35-36) and then calls
! fd3d() exactly as dynamicsolver.f90:62-66 does.  After the loop it
! writes the 2D result arrays as one raw stream, 'result/final.bin'.
!
! fault_in.bin layout (Fortran order, default REAL kind of the build):
!   slip weakening : striniX striniZ peak_xz dyn_xz Dc coh  (nxt x nzt each)
!   FVW            : Dc striniX striniZ a b psi vw aX bX psiX vwX
!                    then scalars Sn uini wini f0 v0 fw hx0 hz0 perturb
!                    RR2 TT2 strinixI
      PROGRAM fd3d_ref_harness
      USE medium_com
      USE fd3dparam_com
      USE friction_com
      USE source_com
      USE pml_com
      USE SlipRates_com
      USE inversion_com
      IMPLICIT NONE
      integer :: u

      call fd3d_init()
      call inversion_init()

      open(newunit=u, file='fault_in.bin', form='unformatted', access='stream', status='old')
#if defined FVW
      read(u) Dc, striniX, striniZ, a, b, psi, vw, aX, bX, psiX, vwX
      read(u) Sn, uini, wini, f0, v0, fw, hx0, hz0, perturb, RR2, TT2, strinixI
#else
      read(u) striniX, striniZ, peak_xz, dyn_xz, Dc, coh
#endif
      close(u)

      ioutput = 1
      call fd3d()

      open(newunit=u, file='result/final.bin', form='unformatted', access='stream', status='replace')
      write(u) ruptime, rise, sliptime, slipX, slipZ, schangeX, schangeZ
      close(u)
      END PROGRAM
