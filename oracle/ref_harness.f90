! Test infrastructure only: a driver PROGRAM that links the unmodified
! reference sources (fd3d_init.f90, fd3d_deriv.f90, fd3d_theo.f90 from
! the FD3D_TSN tree) in place of src/dynamicsolver.f90, so that the
! reference time loop fd3d() can run on seeded per-node fault fields.
!
! The reference's own setup routines cannot express seeded random stress
! or friction fields (forwardspecialTPV5 only knows rectangular patches,
! readinversionresult forces dyn_xz = coh = 0, fd3d_init.f90:830-835), so
! by default this program reads the fault fields from 'fault_in.bin' after
! fd3d_init()/inversion_init() (dynamicsolver.f90:35-36) and then calls
! fd3d() exactly as dynamicsolver.f90:62-66 does.  After the loop it
! writes the 2D result arrays as one raw stream, 'result/final.bin', and
! the inversion accumulators MSRX, MSRZ, MomentRate, 'result/rates.bin'.
! Optional file 'rates_in.txt' (NL NW dL dW dtseis) sets the patch grid
! and bin width of those accumulators before inversion_init allocates them.
!
! Optional file 'harness_setup.txt' (first word):
!   file       read fault_in.bin (default)
!   tpv5       call the reference's forwardspecialTPV5   (slip weakening builds)
!   tpv104     call the reference's forwardspecialTPV104 (rate-and-state builds)
!   inversion  call the reference's readinversionresult (forwardmodel.dat)
! followed optionally by the word 'dump': write the fault arrays the setup
! produced to 'setup_dump.bin' and stop before the time loop.
!
! fault_in.bin / setup_dump.bin layout (Fortran order, default REAL kind):
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
      integer :: u, ios
      character(len=32) :: setup, action
      logical :: have, rates
      real :: dts

      setup = 'file'
      action = ''
      inquire(file='harness_setup.txt', exist=have)
      if (have) then
        open(newunit=u, file='harness_setup.txt', status='old')
        read(u, *, iostat=ios) setup, action
        close(u)
      endif

      call fd3d_init()
      ! optional source-time bins of the inversion (SlipRates_com): the
      ! reference leaves NL = NW = 0 and dtseis = 1 (inversion_init)
      inquire(file='rates_in.txt', exist=rates)
      if (rates) then
        open(newunit=u, file='rates_in.txt', status='old')
        read(u, *) NL, NW, dL, dW, dts
        close(u)
      endif
      call inversion_init()
      if (rates) dtseis = dts

      select case (trim(setup))
      case ('file')
        open(newunit=u, file='fault_in.bin', form='unformatted', access='stream', status='old')
#if defined FVW
        read(u) Dc, striniX, striniZ, a, b, psi, vw, aX, bX, psiX, vwX
        read(u) Sn, uini, wini, f0, v0, fw, hx0, hz0, perturb, RR2, TT2, strinixI
#else
        read(u) striniX, striniZ, peak_xz, dyn_xz, Dc, coh
#endif
        close(u)
#if defined FVW
      case ('tpv104')
        call forwardspecialTPV104()
#else
      case ('tpv5')
        call forwardspecialTPV5()
#endif
      case ('inversion')
        call readinversionresult()
      case default
        print *, 'unknown setup ', trim(setup)
        stop 1
      end select

      if (trim(action) == 'dump') then
        open(newunit=u, file='setup_dump.bin', form='unformatted', access='stream', status='replace')
#if defined FVW
        write(u) Dc, striniX, striniZ, a, b, psi, vw, aX, bX, psiX, vwX
        write(u) Sn, uini, wini, f0, v0, fw, hx0, hz0, perturb, RR2, TT2, strinixI
#else
        write(u) striniX, striniZ, peak_xz, dyn_xz, Dc, coh
#endif
        write(u) lam1, mu1, d1
        close(u)
        stop
      endif

      ioutput = 1
      call fd3d()

      open(newunit=u, file='result/rates.bin', form='unformatted', access='stream', status='replace')
      write(u) MSRX, MSRZ, MomentRate
      close(u)
      open(newunit=u, file='result/final.bin', form='unformatted', access='stream', status='replace')
      write(u) ruptime, rise, sliptime, slipX, slipZ, schangeX, schangeZ
      close(u)
      END PROGRAM
