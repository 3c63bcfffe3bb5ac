! fd3d_tsn_driver: the FD3D_TSN program on the MI355X time-stepping core.
!
! Replaces src/dynamicsolver.f90 + SUBROUTINE fd3d() of the reference.  It
! reads the reference's input files from the working directory
!   inputfd3d.dat  (grid, dt, PML, damping, stations; fd3d_init.f90:145-160)
!   crustal.dat    (layered medium; readcrustalmodel, fd3d_init.f90:211-279)
! and the fault setup chosen on the command line, then runs the time loop
! through the C ABI (include/fd3d_tsn.h) and writes the reference's outputs
! into result/:
!   per step   sliprateX.res sliprateZ.res shearstressX.res shearstressZ.res
!              [psi.res]  (stream files of default REAL, x fastest;
!              fd3d_theo.f90:546-553), stan0.txt.. (stations, :555-557)
!   at the end ruptime.res risetime.res slipX.res slipZ.res stressdropX.res
!              stressdropZ.res contour.res czone.res rvel.res (:687-748)
!   at start   vmodel.inp friction.inp (:93-116)
!
! usage: fd3d_tsn_driver [sw|fvw] [fspace|free] [file|tpv5|tpv104|inversion] [dump]
!   sw/fvw      friction law (the reference's -DFVW)
!   fspace/free full space or free surface (the reference's -DFSPACE)
!   file        per-node fault fields from fault_in.bin (see below)
!   tpv5        scecmodel.dat, SCEC TPV5 layout (forwardspecialTPV5, fd3d_init.f90:535-617)
!   tpv104      scecmodel.dat, SCEC TPV104 layout (forwardspecialTPV104, :393-501)
!   inversion   inputinv.dat + forwardmodel.dat (readinversionresult, :786-854)
!   dump        write the set-up arrays to setup_dump.bin and stop
! fault_in.bin: slip weakening, float32 striniX striniZ peak_xz dyn_xz Dc coh;
! rate-and-state, float64 Dc striniX striniZ a b psi vw aX bX psiX vwX and the
! scalars Sn uini wini f0 v0 fw hx0 hz0 perturb RR2 TT2 strinixI.
program fd3d_tsn_driver
  use iso_c_binding
  use fd3d_tsn_mod
  implicit none

  ! ---- inputfd3d.dat
  integer :: nxtT, nytT, nztT, ntfd, nabc, nfs, nstations
  integer :: nxt, nyt, nzt, nysc
  real :: dh, dt, dip, pml_vp, pml_fact, damp_s, waveT
  integer, allocatable, target :: sta(:, :)
  ! ---- medium
  real(c_float), allocatable :: lam1(:, :, :), mu1(:, :, :), d1(:, :, :)
  ! ---- slip weakening fault (Z-nodes)
  real(c_float), allocatable :: striniX(:, :), striniZ(:, :), peak_xz(:, :), dyn_xz(:, :), Dc(:, :), coh(:, :)
  ! ---- rate-and-state fault
  real(c_double), allocatable :: dDc(:, :), dstrX(:, :), dstrZ(:, :), fa(:, :), fb(:, :), fpsi(:, :), fvw(:, :)
  real(c_double), allocatable :: aX(:, :), bX(:, :), psiX(:, :), vwX(:, :)
  real(c_double) :: Sn, uini, wini, f0, v0, fw, hx0, hz0, perturb, RR2, TT2, strinixI
  ! ---- run
  logical :: is_fvw, fspace, dump
  character(len=32) :: arg, setup
  type(fd3d_config) :: cfg
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer :: i, k, n, it, chunk, nxs, nzs, ring, nout, nbegun, inflight
  integer(c_int) :: nw
  type(c_ptr) :: pout, pseis
  real(c_float), pointer :: outbuf(:), seisbuf(:)
  real(c_float), allocatable :: final7(:)
  real, allocatable :: T0X(:, :), T0Z(:, :)
  real :: maxX, maxZ, time
  real :: cput1, cput2
  integer(8) :: c1, c2, crate

  is_fvw = .false.
  fspace = .true.
  dump = .false.
  setup = 'file'
  do n = 1, command_argument_count()
    call get_command_argument(n, arg)
    select case (trim(arg))
    case ('sw'); is_fvw = .false.
    case ('fvw'); is_fvw = .true.
    case ('fspace'); fspace = .true.
    case ('free'); fspace = .false.
    case ('file', 'tpv5', 'tpv104', 'inversion'); setup = arg
    case ('dump'); dump = .true.
    case default
      print *, 'usage: fd3d_tsn_driver [sw|fvw] [fspace|free] [file|tpv5|tpv104|inversion] [dump]'
      stop 1
    end select
  end do

  call read_inputfd3d()
  call read_crustal_model()
  if (is_fvw) then
    call setup_fvw()
  else
    call setup_sw()
  end if
  if (dump) then
    call write_dump()
    stop
  end if

  ! ---- device problem
  cfg%nxt = nxt; cfg%nyt = nyt; cfg%nzt = nzt; cfg%nabc = nabc
  cfg%fspace = merge(1, 0, fspace)
  cfg%friction = merge(FD3D_FVW, FD3D_SW, is_fvw)
  cfg%nucleation = merge(1, 0, is_fvw)
  cfg%nstations = nstations
  if (nstations > 0) then
    cfg%stations = c_loc(sta)
  else
    cfg%stations = c_null_ptr
  end if
  cfg%dh = dh; cfg%dt = dt; cfg%damp_s = damp_s; cfg%pml_vp = pml_vp; cfg%pml_fact = pml_fact
  cfg%Sn = 0; cfg%uini = 0; cfg%wini = 0; cfg%f0 = 0; cfg%v0 = 0; cfg%fw = 0
  cfg%hx0 = 0; cfg%hz0 = 0; cfg%perturb = 0; cfg%RR2 = 0; cfg%TT2 = 0; cfg%strinixI = 0
  if (is_fvw) then
    cfg%Sn = Sn; cfg%uini = uini; cfg%wini = wini; cfg%f0 = f0; cfg%v0 = v0; cfg%fw = fw
    cfg%hx0 = hx0; cfg%hz0 = hz0; cfg%perturb = perturb; cfg%RR2 = RR2; cfg%TT2 = TT2
    cfg%strinixI = strinixI
  end if
  cfg%newton_max = 0
  cfg%ring_steps = 0
  h = c_null_ptr
  rc = fd3d_create(cfg, h)
  call fd3d_check(rc, h, 'fd3d_create')
  call fd3d_check(fd3d_upload_medium(h, lam1, mu1, d1), h, 'fd3d_upload_medium')
  if (is_fvw) then
    call fd3d_check(fd3d_upload_fault_fvw(h, dDc, dstrX, dstrZ, fa, fb, fpsi, fvw, aX, bX, psiX, vwX), &
                    h, 'fd3d_upload_fault_fvw')
  else
    call fd3d_check(fd3d_upload_fault_sw(h, striniX, striniZ, peak_xz, dyn_xz, Dc, coh), h, &
                    'fd3d_upload_fault_sw')
  end if
  call fd3d_check(fd3d_init_state(h), h, 'fd3d_init_state')

  call interp_t0()
  call write_inputs_echo()

  ! ---- time loop (fd3d_theo.f90:187-602).  Outputs stream out chunk by
  ! chunk: while the files of one chunk are written, the GPU computes the
  ! next (fd3d_step_begin / fd3d_step_wait, pinned buffers of the library)
  nxs = nxt - 2 * nabc
  nzs = nzt - nfs - nabc
  nout = merge(5, 4, is_fvw)
  ring = 64
  chunk = ring / 2
  call open_outputs()
  call cpu_time(cput1)
  call system_clock(c1, crate)
  it = 0
  nbegun = 0
  inflight = 0
  do while (nbegun < ntfd .and. inflight < 2)
    call fd3d_check(fd3d_step_begin(h, min(chunk, ntfd - nbegun)), h, 'fd3d_step_begin')
    nbegun = nbegun + min(chunk, ntfd - nbegun)
    inflight = inflight + 1
  end do
  do while (inflight > 0)
    call fd3d_check(fd3d_step_wait(h, nw, pout, pseis), h, 'fd3d_step_wait')
    inflight = inflight - 1
    call c_f_pointer(pout, outbuf, [int(nw, 8) * FD3D_NOUT * nxs * nzs])
    call c_f_pointer(pseis, seisbuf, [max(1, nw * nstations * 3)])
    do n = 1, nw
      call write_step(n, it + n)
    end do
    it = it + nw
    ! the buffer just written is the one the chunk after next reuses
    if (nbegun < ntfd) then
      call fd3d_check(fd3d_step_begin(h, min(chunk, ntfd - nbegun)), h, 'fd3d_step_begin')
      nbegun = nbegun + min(chunk, ntfd - nbegun)
      inflight = inflight + 1
    end if
  end do
  call system_clock(c2)
  call cpu_time(cput2)
  print *, 'CPU TIME OF TIME LOOP: ', cput2 - cput1
  print *, 'WALL TIME OF TIME LOOP: ', real(c2 - c1) / real(crate)
  call close_outputs()

  allocate (final7(FD3D_NFINAL * nxt * nzt))
  call fd3d_check(fd3d_download_final(h, final7), h, 'fd3d_download_final')
  call fd3d_destroy(h)
  call write_results()

contains

  ! ------------------------------------------------------------------ input
  subroutine read_inputfd3d()
    integer :: u
    open (newunit=u, file='inputfd3d.dat', status='old')
    read (u, *) nxtT, nytT, nztT
    read (u, *) dh
    read (u, *) ntfd
    read (u, *) dt
    read (u, *) dip
    read (u, *) nabc, pml_vp, pml_fact
    read (u, *) damp_s
    read (u, *) nstations
    allocate (sta(3, max(1, nstations)))
    do n = 1, nstations
      read (u, *) sta(1, n), sta(2, n), sta(3, n)
    end do
    read (u, *) waveT
    close (u)
    nfs = merge(nabc, 2, fspace)          ! fd3d_init.f90:161-165
    nxt = nxtT + 2 * nabc
    nyt = nytT + nabc
    nzt = nztT + nabc + nfs
    nysc = nyt
    print *, nxt, nyt, nzt, nfs, ntfd
  end subroutine

  ! Layered medium, as readcrustalmodel (fd3d_init.f90:211-279) fills it.
  subroutine read_crustal_model()
    real(8), parameter :: PI = 3.1415926535
    real :: vpe(2), dum, dd, vpp, vss, cfl
    real, allocatable :: vp(:), vs(:), depth(:), rho(:)
    integer :: ndepth, j, kk, u
    allocate (lam1(nxt, nyt, nzt), mu1(nxt, nyt, nzt), d1(nxt, nyt, nzt))
    vpe(2) = 0.
    vpe(1) = 1.0E+10
    open (newunit=u, file='crustal.dat', status='old')
    read (u, *)
    read (u, *)
    read (u, *) ndepth
    allocate (depth(ndepth), vp(ndepth), vs(ndepth), rho(ndepth))
    read (u, *)
    read (u, *)
    do kk = 1, ndepth
      read (u, *) depth(kk), vp(kk), vs(kk), rho(kk)
    end do
    close (u)
    depth = depth * 1.e3; vp = vp * 1.e3; vs = vs * 1.e3; rho = rho * 1.e3
    do kk = nzt, 1, -1
      dum = (dh * real(nzt - nfs - kk) + dh / 2.) * sin(dip / 180.d0 * PI)
      if (dum > depth(ndepth)) then
        vpp = vp(ndepth); vss = vs(ndepth); dd = rho(ndepth)
      else
        do j = 2, ndepth
          if (dum < depth(j)) exit
        end do
        vpp = vp(j - 1); vss = vs(j - 1); dd = rho(j - 1)
      end if
      if (vpp > vpe(2)) vpe(2) = vpp
      if (vpp < vpe(1)) vpe(1) = vpp
      lam1(:, :, kk) = dd * (vpp**2 - 2. * vss**2)
      mu1(:, :, kk) = dd * vss**2
      d1(:, :, kk) = dd
    end do
    cfl = vpe(2) * dt / dh
    if (cfl > 0.25) then
      print *, 'Your simulation is numerically unstable', cfl
      stop 3
    end if
  end subroutine

  real function normstress(kk)
    integer, intent(in) :: kk
    real, parameter :: pi = 3.1415926535
    normstress = max(1.e5, 16200. * dh * real(nzt - nfs - kk) * sin(dip / 180. * pi))
  end function

  ! ---------------------------------------------------- slip-weakening setup
  subroutine setup_sw()
    integer :: u
    allocate (striniX(nxt, nzt), striniZ(nxt, nzt), peak_xz(nxt, nzt), dyn_xz(nxt, nzt), Dc(nxt, nzt), coh(nxt, nzt))
    striniX = 0.; striniZ = 0.; peak_xz = 0.; Dc = 0.; dyn_xz = 0.; coh = 0.
    select case (trim(setup))
    case ('file')
      open (newunit=u, file='fault_in.bin', form='unformatted', access='stream', status='old')
      read (u) striniX, striniZ, peak_xz, dyn_xz, Dc, coh
      close (u)
    case ('tpv5')
      call setup_tpv5()
    case ('inversion')
      call setup_inversion()
    case default
      print *, 'setup ', trim(setup), ' is not a slip-weakening setup'
      stop 1
    end select
  end subroutine

  ! TPV5 (forwardspecialTPV5, fd3d_init.f90:535-617): prestress patches and
  ! a strength barrier outside the fault.
  subroutine setup_tpv5()
    real :: hx0r, hz0r, h1x0, h1z0, h2x0, h2z0, hdelta, T0, T0n, T0h1, T0h2, sn, mus, muso, mud, d0h
    integer :: no0, u, ii, kk
    open (newunit=u, file='scecmodel.dat', status='old')
    read (u, *) no0, hx0r, hz0r, h1x0, h1z0, h2x0, h2z0
    read (u, *) hdelta
    read (u, *) T0, T0n, T0h1, T0h2
    read (u, *) sn
    read (u, *) mus, muso
    read (u, *) mud
    read (u, *) d0h
    close (u)
    striniZ = 0.
    coh = 0.
    do ii = 1, nxt
      do kk = 1, nzt - 2
        striniX(ii, kk) = T0
        if (inpatch(ii, kk, hx0r, hz0r)) striniX(ii, kk) = T0n
        if (inpatch(ii, kk, h1x0, h1z0)) striniX(ii, kk) = T0h1
        if (inpatch(ii, kk, h2x0, h2z0)) striniX(ii, kk) = T0h2
        peak_xz(ii, kk) = sn * muso
        dyn_xz(ii, kk) = sn * mud
        Dc(ii, kk) = d0h
      end do
    end do
    do kk = no0 + 1, nzt - 2
      do ii = no0 + 1, nxt - no0
        peak_xz(ii, kk) = sn * mus
      end do
    end do
    do ii = 1, nxt
      peak_xz(ii, nzt - 1) = peak_xz(ii, nzt - 2)
      dyn_xz(ii, nzt - 1) = dyn_xz(ii, nzt - 2)
      Dc(ii, nzt - 1) = Dc(ii, nzt - 2)
      striniX(ii, nzt - 1) = striniX(ii, nzt - 2)
      peak_xz(ii, nzt) = peak_xz(ii, nzt - 3)
      dyn_xz(ii, nzt) = dyn_xz(ii, nzt - 3)
      Dc(ii, nzt) = Dc(ii, nzt - 3)
      striniX(ii, nzt) = striniX(ii, nzt - 3)
    end do
  contains
    logical function inpatch(ii, kk, cx, cz)
      integer, intent(in) :: ii, kk
      real, intent(in) :: cx, cz
      inpatch = (((real(ii) - 1.) * dh - cx >= -hdelta / 2.0) .and. ((real(ii) - 1.) * dh - cx <= hdelta / 2.0)) &
                .and. (((real(kk) - 1.) * dh - cz >= -hdelta / 2.0) .and. ((real(kk) - 1.) * dh - cz <= hdelta / 2.0))
    end function
  end subroutine

  ! Inverted model on coarse control points (inversion_init,
  ! readinversionresult, inversion_modeltofd3d; fd3d_init.f90:786-854):
  ! bilinear interpolation of T0, Ts (strength over normal stress) and Dc.
  subroutine setup_inversion()
    integer :: runi, nli, nwi, u, ii, kk
    real, allocatable :: T0I(:, :), TsI(:, :), DcI(:, :)
    real :: dum, DL, DW, XS, ZS, t, uu
    open (newunit=u, file='inputinv.dat', status='old')
    read (u, *) runi
    read (u, *) nli, nwi
    close (u)
    allocate (T0I(nli, nwi), TsI(nli, nwi), DcI(nli, nwi))
    open (newunit=u, file='forwardmodel.dat', status='old')
    read (u, *) dum, dum, T0I(:, :), TsI(:, :), DcI(:, :)
    close (u)
    DL = dh * (nxt - 2 * nabc) / real(nli - 1)
    DW = dh * (nzt - nfs - nabc) / real(nwi - 1)
    do k = nabc + 1, nzt - nfs
      ZS = dh * (k - 1 - nabc)
      kk = int(ZS / DW) + 1
      uu = (ZS - DW * (kk - 1)) / DW
      do i = nabc + 1, nxt - nabc
        XS = dh * (i - 1 - nabc)
        ii = int(XS / DL) + 1
        t = (XS - DL * (ii - 1)) / DL
        striniX(i, k) = (1. - t) * (1. - uu) * T0I(ii, kk) + t * (1. - uu) * T0I(ii + 1, kk) + &
                        t * uu * T0I(ii + 1, kk + 1) + (1. - t) * uu * T0I(ii, kk + 1)
        striniZ(i, k) = 0.
        peak_xz(i, k) = ((1. - t) * (1. - uu) * TsI(ii, kk) + t * (1. - uu) * TsI(ii + 1, kk) + &
                         t * uu * TsI(ii + 1, kk + 1) + (1. - t) * uu * TsI(ii, kk + 1)) * normstress(k)
        Dc(i, k) = (1. - t) * (1. - uu) * DcI(ii, kk) + t * (1. - uu) * DcI(ii + 1, kk) + &
                   t * uu * DcI(ii + 1, kk + 1) + (1. - t) * uu * DcI(ii, kk + 1)
      end do
    end do
    dyn_xz = 0.
    coh = 0.e6
  end subroutine

  ! ---------------------------------------------------- rate-and-state setup
  subroutine setup_fvw()
    integer :: u
    real(8) :: scal(12)
    allocate (dDc(nxt, nzt), dstrX(nxt, nzt), dstrZ(nxt, nzt), fa(nxt, nzt), fb(nxt, nzt), fpsi(nxt, nzt), &
              fvw(nxt, nzt), aX(nxt, nzt), bX(nxt, nzt), psiX(nxt, nzt), vwX(nxt, nzt))
    dDc = 0; dstrX = 0; dstrZ = 0; fa = 0; fb = 0; fpsi = 0; fvw = 0; aX = 0; bX = 0; psiX = 0; vwX = 0
    select case (trim(setup))
    case ('file')
      open (newunit=u, file='fault_in.bin', form='unformatted', access='stream', status='old')
      read (u) dDc, dstrX, dstrZ, fa, fb, fpsi, fvw, aX, bX, psiX, vwX
      read (u) scal
      close (u)
      Sn = scal(1); uini = scal(2); wini = scal(3); f0 = scal(4); v0 = scal(5); fw = scal(6)
      hx0 = scal(7); hz0 = scal(8); perturb = scal(9); RR2 = scal(10); TT2 = scal(11); strinixI = scal(12)
    case ('tpv104')
      call setup_tpv104()
    case default
      print *, 'setup ', trim(setup), ' is not a rate-and-state setup'
      stop 1
    end select
  end subroutine

  ! TPV104 (forwardspecialTPV104, fd3d_init.f90:393-501), in double
  ! precision like the reference built with -fdefault-real-8.
  subroutine setup_tpv104()
    real(8) :: wfx, wfz, trans, fringe, f0ini, v0ini, a0ini, bini, d0h, fwini, vw0ini, vwdeltaini, vini, RRini
    real(8) :: Tini, Snini, BX1, BZ1, hx, hz, hx0t, hz0t, dhd
    integer :: u, ii, kk
    open (newunit=u, file='scecmodel.dat', status='old')
    read (u, *) wfx, wfz, hx0t, hz0t, trans, fringe
    read (u, *) f0ini
    read (u, *) v0ini
    read (u, *) a0ini
    read (u, *) bini
    read (u, *) d0h
    read (u, *) fwini
    read (u, *) vw0ini
    read (u, *) vwdeltaini
    read (u, *) vini
    read (u, *) Tini
    read (u, *) Snini
    read (u, *) perturb
    read (u, *) RRini
    read (u, *) TT2
    close (u)
    dhd = dble(dh)
    RR2 = RRini * RRini
    Sn = Snini
    uini = vini
    wini = 0.d0
    f0 = f0ini
    v0 = v0ini
    fb = bini
    strinixI = Tini
    fw = fwini
    hx0 = hx0t + fringe + dble(nabc) * dhd
    hz0 = hz0t + fringe + dble(nabc) * dhd
    dstrZ = 0.d0
    do kk = 1, nzt
      do ii = 1, nxt
        dstrX(ii, kk) = Tini
        dDc(ii, kk) = d0h
        hx = dble(ii) * dhd
        hz = dble(kk) * dhd
        BX1 = taper(abs(hx - hx0), abs(hx - hx0), wfx, trans)
        BZ1 = taper(-(hz - hz0), abs(hz - hz0), wfz, trans)
        fa(ii, kk) = a0ini + a0ini * (1.d0 - BX1 * BZ1)
        fvw(ii, kk) = vw0ini + vwdeltaini * (1.d0 - BX1 * BZ1)
        fpsi(ii, kk) = fa(ii, kk) * (log((2 * v0ini / (2 * vini))) + log(sinh(dstrX(ii, kk) / (fa(ii, kk) * Sn))))
      end do
    end do
    do kk = 1, nzt
      do ii = 1, nxt
        hx = dble(ii) * dhd + dhd / 2.d0
        hz = dble(kk) * dhd + dhd / 2.d0
        BX1 = taper(abs(hx - hx0), abs(hx - hx0), wfx, trans)
        BZ1 = taper(-(hz - hz0), abs(hz - hz0), wfz, trans)
        aX(ii, kk) = a0ini + a0ini * (1.d0 - BX1 * BZ1)
        vwX(ii, kk) = vw0ini + vwdeltaini * (1.d0 - BX1 * BZ1)
        psiX(ii, kk) = aX(ii, kk) * (log((2 * v0 / (2 * uini))) + log(sinh(dstrX(ii, kk) / (aX(ii, kk) * Sn))))
        bX(ii, kk) = bini
      end do
    end do
  contains
    ! the boxcar-with-tanh taper: the branch test uses dsel, the ramp abs value
    real(8) function taper(dsel, dabs, half, tr)
      real(8), intent(in) :: dsel, dabs, half, tr
      if (dsel >= half + tr) then
        taper = 0.d0
      else if ((dsel > half) .and. (dsel < half + tr)) then
        taper = 0.5d0 * (1 + tanh(tr / (dabs - half - tr) + tr / (dabs - half)))
      else
        taper = 1.d0
      end if
    end function
  end subroutine

  subroutine write_dump()
    integer :: u
    open (newunit=u, file='setup_dump.bin', form='unformatted', access='stream', status='replace')
    if (is_fvw) then
      write (u) dDc, dstrX, dstrZ, fa, fb, fpsi, fvw, aX, bX, psiX, vwX
      write (u) Sn, uini, wini, f0, v0, fw, hx0, hz0, perturb, RR2, TT2, strinixI
    else
      write (u) striniX, striniZ, peak_xz, dyn_xz, Dc, coh
    end if
    write (u) lam1, mu1, d1
    close (u)
  end subroutine

  ! T0X/T0Z as interp_fric forms them (fd3d_deriv.f90:3164-3186), for friction.inp
  subroutine interp_t0()
    real, allocatable :: sx(:, :), sz(:, :)
    allocate (T0X(nxt, nzt), T0Z(nxt, nzt), sx(nxt, nzt), sz(nxt, nzt))
    if (is_fvw) then
      sx = real(dstrX); sz = real(dstrZ)
    else
      sx = striniX; sz = striniZ
    end if
    T0X = 0.; T0Z = sz
    do i = 2, nxt - 1
      do k = 2, nzt - 1
        T0X(i, k) = (sx(i, k) + sx(i - 1, k) + sx(i, k - 1) + sx(i - 1, k - 1)) / 4.
      end do
    end do
  end subroutine

  ! vmodel.inp and friction.inp (fd3d_theo.f90:93-116)
  subroutine write_inputs_echo()
    integer :: u95, u96
    call execute_command_line('mkdir -p result')
    open (newunit=u95, file='result/vmodel.inp')
    open (newunit=u96, file='result/friction.inp')
    do k = nabc + 1, nzt - nfs
      do i = nabc + 1, nxt - nabc
        write (u95, *) mu1(i, nysc, k)
        if (is_fvw) then
          write (u96, '(5E13.5)') T0X(i, k), T0Z(i, k), real(fa(i, k) - fb(i, k)), real(dstrX(i, k))
        else
          write (u96, '(5E13.5)') T0X(i, k), T0Z(i, k), peak_xz(i, k), Dc(i, k), peak_xz(i, k) / normstress(k)
        end if
      end do
    end do
    close (u95)
    close (u96)
  end subroutine

  ! ----------------------------------------------------------------- output
  subroutine open_outputs()
    integer :: s
    character(len=32) :: fname
    if (is_fvw) open (24, file='result/psi.res', form='unformatted', access='stream', status='replace')
    open (25, file='result/sliprateZ.res', form='unformatted', access='stream', status='replace')
    open (27, file='result/sliprateX.res', form='unformatted', access='stream', status='replace')
    open (26, file='result/shearstressZ.res', form='unformatted', access='stream', status='replace')
    open (28, file='result/shearstressX.res', form='unformatted', access='stream', status='replace')
    do s = 1, nstations
      write (fname, '(a,i0,a)') 'result/stan', s - 1, '.txt'
      open (1000 + s, file=trim(fname))
    end do
  end subroutine

  subroutine close_outputs()
    integer :: s
    if (is_fvw) close (24)
    close (25); close (26); close (27); close (28)
    do s = 1, nstations
      close (1000 + s)
    end do
  end subroutine

  ! one step's slabs from the chunk buffer: order sliprateX, sliprateZ,
  ! shearstressX, shearstressZ, psi (include/fd3d_tsn.h, fd3d_output)
  subroutine write_step(n, itg)
    integer, intent(in) :: n, itg
    integer(8) :: slab, base
    integer :: s
    slab = int(nxs, 8) * nzs
    base = int(n - 1, 8) * FD3D_NOUT * slab
    if (is_fvw) write (24) outbuf(base + 4 * slab + 1:base + 5 * slab)
    write (25) outbuf(base + slab + 1:base + 2 * slab)
    write (27) outbuf(base + 1:base + slab)
    write (26) outbuf(base + 3 * slab + 1:base + 4 * slab)
    write (28) outbuf(base + 2 * slab + 1:base + 3 * slab)
    do s = 1, nstations
      write (1000 + s, *) seisbuf((int(n - 1) * nstations + s - 1) * 3 + 1), &
        seisbuf((int(n - 1) * nstations + s - 1) * 3 + 2), seisbuf((int(n - 1) * nstations + s - 1) * 3 + 3)
    end do
    ! progress line every simulated second (fd3d_theo.f90:580-583)
    if (mod(itg, max(1, int(1. / dt))) == 0) then
      time = itg * dt
      maxZ = maxval(outbuf(base + slab + 1:base + 2 * slab))
      maxX = maxval(outbuf(base + 1:base + slab))
      write (*, *) 'Time: ', time, 'Slip rate max: ', maxX, maxZ
    end if
  end subroutine

  ! End-of-run files and summaries (fd3d_theo.f90:658-749)
  subroutine write_results()
    real, allocatable :: ruptime(:, :), rise(:, :), sliptime(:, :), slipX(:, :), slipZ(:, :), sdX(:, :), sdZ(:, :)
    real :: output_param(6), numer, denom, rv, cz, tint, tint2
    integer :: u94, u95, u96, u97, u98, u99, u501, u502, u503
    integer(8) :: n2
    n2 = int(nxt, 8) * nzt
    allocate (ruptime(nxt, nzt), rise(nxt, nzt), sliptime(nxt, nzt), slipX(nxt, nzt), slipZ(nxt, nzt), &
              sdX(nxt, nzt), sdZ(nxt, nzt))
    ruptime = reshape(final7(1:n2), [nxt, nzt])
    rise = reshape(final7(n2 + 1:2 * n2), [nxt, nzt])
    sliptime = reshape(final7(2 * n2 + 1:3 * n2), [nxt, nzt])
    slipX = reshape(final7(3 * n2 + 1:4 * n2), [nxt, nzt])
    slipZ = reshape(final7(4 * n2 + 1:5 * n2), [nxt, nzt])
    sdX = reshape(final7(5 * n2 + 1:6 * n2), [nxt, nzt])
    sdZ = reshape(final7(6 * n2 + 1:7 * n2), [nxt, nzt])
    output_param = 0.
    numer = 0.
    denom = 0.
    do k = nabc + 1, nzt - nfs
      do i = nabc + 1, nxt - nabc
        output_param(2) = output_param(2) + sqrt(slipX(i, k)**2 + slipZ(i, k)**2) * mu1(i, nysc, k) * (dh * dh)
        numer = numer + sdZ(i, k) * slipZ(i, k) + sdX(i, k) * slipX(i, k)
        denom = denom + sqrt(slipX(i, k)**2 + slipZ(i, k)**2)
        if (denom /= 0.0) then
          output_param(4) = -(numer / denom)
        else
          output_param(4) = 0.0
        end if
      end do
    end do
    open (newunit=u96, file='result/risetime.res')
    open (newunit=u97, file='result/ruptime.res')
    open (newunit=u98, file='result/slipX.res')
    open (newunit=u99, file='result/stressdropX.res')
    open (newunit=u95, file='result/stressdropZ.res')
    open (newunit=u94, file='result/slipZ.res')
    do k = nabc + 1, nzt - nfs
      do i = nabc + 1, nxt - nabc
        write (u96, *) rise(i, k)
        write (u97, *) ruptime(i, k)
        write (u98, *) slipX(i, k)
        write (u94, *) slipZ(i, k)
        write (u99, *) sdX(i, k)
        write (u95, *) sdZ(i, k)
      end do
    end do
    close (u96); close (u97); close (u98); close (u99); close (u95); close (u94)
    open (newunit=u501, file='result/contour.res')
    write (u501, *) 'j k t'
    do k = nabc + 1, nzt - nfs
      do i = nabc + 1, nxt - nabc
        write (u501, *) (real(i - nabc - 1) - real(nxt - 2 * nabc) / 2.) * dh, real(nzt - k - nabc - 1) * dh, &
          ruptime(i, k)
      end do
    end do
    close (u501)
    open (newunit=u502, file='result/czone.res')
    open (newunit=u503, file='result/rvel.res')
    tint = 0.
    tint2 = 0.
    do k = nabc + 1, nzt - nfs
      do i = nabc + 1, nxt - nabc
        cz = 0.
        rv = 0.
        if ((ruptime(i, k) /= 1.e4) .and. (ruptime(i + 1, k) /= 1.e4) .and. (ruptime(i, k + 1) /= 1.e4) &
            .and. (ruptime(i - 1, k) /= 1.e4) .and. (ruptime(i, k - 1) /= 1.e4)) then
          if (sliptime(i, k) /= 1.e4) then
            rv = (sqrt((ruptime(i + 1, k) - ruptime(i, k))**2 + (ruptime(i, k + 1) - ruptime(i, k))**2) &
                  + sqrt((ruptime(i, k) - ruptime(i - 1, k))**2 + (ruptime(i, k) - ruptime(i, k - 1))**2))
            if (rv /= 0.) then
              rv = 2 * dh / (sqrt((ruptime(i + 1, k) - ruptime(i, k))**2 + (ruptime(i, k + 1) - ruptime(i, k))**2) &
                             + sqrt((ruptime(i, k) - ruptime(i - 1, k))**2 + (ruptime(i, k) - ruptime(i, k - 1))**2))
              tint = tint + rv * sqrt(slipX(i, k)**2 + slipZ(i, k)**2)
              tint2 = tint2 + sqrt(slipX(i, k)**2 + slipZ(i, k)**2)
            else
              rv = 0.
            end if
            cz = rv * (sliptime(i, k) - ruptime(i, k))
          end if
        end if
        write (u502, *) cz
        write (u503, *) rv
      end do
    end do
    output_param(1) = tint / tint2
    close (u502)
    close (u503)
    print *, '-----------------------------------------------------'
    print *, 'Average speed:       ', output_param(1)
    print *, 'Seismic moment:      ', output_param(2)
    print *, 'Surface of rupture:  ', output_param(3)
    print *, 'Average stress drop: ', output_param(4)
    print *, '-----------------------------------------------------'
  end subroutine

end program fd3d_tsn_driver
