!=============================================================================
! Global state of the channel driver: run control read from a namelist, the
! physical and MRT constants (computed in channel_para exactly as
! Channel-Flow/para.f90:59-174 computes them), the device handle, and the
! host copies of f and the macroscopic fields that the I/O routines use.
!
! Axes follow the reference: x wall-normal (lx), y streamwise (ly),
! z spanwise (lz); host arrays are f(0:18,lx,ly,lz) and field(lx,ly,lz).
!=============================================================================
      module channel_vars
      use iso_c_binding
      use lbm_capi
      implicit none

      integer, parameter :: npop = 19
      real(8), parameter :: rho0 = 1.0d0
      real(8), parameter :: pi = 4.0d0*atan(1.0d0)

      ! run control (namelist /channel/)
      integer :: lx = 32, ly = 64, lz = 32
      integer :: nsteps = 1000, ndiag = 250, nflowout = 0, nstat = 0
      integer :: npforcing = 0   ! last step of the perturbation forcing (para.f90:72-73)
      integer :: precision = 0, device = 0, mrttype_in = 0
      logical :: laminarflow = .false., newrun = .true., newinitflow = .true.
      logical :: savecntd = .true.
      integer :: istpload = 0
      character(len=200) :: dirgenr = './'
      ! checkpoint sets of an nprocY x nprocZ run of the reference
      ! (para.f90:218-262): loadcntdflow reads nprocy_in x nprocz_in files,
      ! savecntdflow writes nprocy_out x nprocz_out files
      integer :: nprocy_in = 1, nprocz_in = 1, nprocy_out = 1, nprocz_out = 1
      ! dissipation_calc cadence (saveload.f90:4644-4882; 0 = never)
      integer :: ndiss = 0
      ! particles (para.f90:332-379): positions from partfile, one "x y z"
      ! line per sphere in node coordinates (node i at i - 0.5)
      logical :: ipart = .false.
      integer :: npart = 0
      real(8) :: rad = 5.0d0, rhopart = 1.0d0, mingap = 3.0d0, mingap_w = 3.0d0
      character(len=200) :: partfile = 'partpos.dat'
      ! learned closure (no reference counterpart): the LBM_MLP_TOTAL FP32
      ! parameters of lbm_closure_set as raw little-endian floats in
      ! closurefile ('' = plain MRT), FP32/FP16 storage only; visc_in > 0
      ! replaces the turbulent branch's nu (a 2x-coarse run of the
      ! Re_tau = 180 channel: lx = 96, visc_in = 0.0018)
      character(len=200) :: closurefile = ''
      real(8) :: visc_in = 0.0d0

      ! time stepping
      integer :: istep, istep0, istat, imovie

      ! physical and MRT constants (para.f90:59-174)
      integer :: mrttype
      real(8) :: visc, rstar, ustar, force_in_y, ystar, force_mag, tau
      real(8) :: s1, s2, s4, s9, s10, s13, s16
      real(8) :: omegepsl, omegepslj, omegxx
      real(8) :: coef1, coef2, coef3, coef4, coef5, coef3i, coef4i
      real(8) :: val1, val2, val3, val4, val5, val6, val7, val8, val9
      real(8) :: val1i, val2i, val3i, val4i, val5i, val6i, val7i, val8i, val9i
      real(8) :: ww0, ww1, ww2
      real(8) :: rhoepsl = 1.0d-5
      real(8) :: rhoerrmax = 0.0d0

      ! device lattice
      type(c_ptr) :: lbm = c_null_ptr

      ! host copies (allocated by the routines that need them)
      real(8), allocatable, target :: f(:,:,:,:)
      real(8), allocatable, target :: rho(:,:,:), ux(:,:,:), uy(:,:,:), uz(:,:,:)
      real(8), allocatable, target :: ox(:,:,:), oy(:,:,:), oz(:,:,:)
      ! particle state (var_inc.f90:142-150): centres, velocities, angular
      ! velocities, hydrodynamic force and torque, lubrication force
      real(8), allocatable, target :: ypglb(:,:), wp(:,:), omgp(:,:), fHIp(:,:), torqp(:,:), flubp(:,:)

      namelist /channel/ lx, ly, lz, nsteps, ndiag, nflowout, nstat, npforcing, precision, &
                         device, mrttype_in, laminarflow, newrun, newinitflow,  &
                         savecntd, istpload, dirgenr, rhoepsl, nprocy_in, nprocz_in, &
                         nprocy_out, nprocz_out, ndiss, ipart, npart, rad, rhopart, &
                         mingap, mingap_w, partfile, closurefile, visc_in

      contains

      ! Physical and MRT constants, Channel-Flow/para.f90:59-174.
      subroutine channel_para
        if (.not. laminarflow) then
          visc = 0.0036d0
          if (visc_in > 0.0d0) visc = visc_in
          rstar = 180.0d0
          ustar = 2.0d0*rstar*visc/dble(lx)
          force_in_y = 2.d0*rho0*ustar*ustar/dble(lx)
          ystar = visc/ustar
          force_mag = 1.0d0
          mrttype = 1
        else
          rstar = 20.0d0
          ustar = 0.05d0
          visc = 2.d0*ustar*dble(lx)/rstar
          force_in_y = 8.d0*visc*ustar/(dble(lx)*dble(lx))
          ystar = visc/ustar
          force_mag = 1.0d0
          mrttype = 2
        end if
        if (mrttype_in > 0) mrttype = mrttype_in

        tau = 3.0d0*visc + 0.5d0
        s9 = 1.0d0/tau
        s13 = s9
        select case (mrttype)
        case (1)
          s1 = 1.5d0
          s2 = 1.4d0
          s4 = 1.2d0
          s10 = 1.4d0
          s16 = 1.98d0
          omegepsl = 0.0d0
          omegepslj = -475.0d0/63.0d0
          omegxx = 0.0d0
        case (2)
          s1 = s9
          s2 = s9
          s4 = s9
          s10 = s9
          s16 = s9
          omegepsl = 3.0d0
          omegepslj = -11.0d0/2.0d0
          omegxx = -1.0d0/2.0d0
        case default
          s1 = 1.8d0
          s2 = s1
          s4 = s9
          s10 = s1
          s16 = s1
          omegepsl = 3.0d0
          omegepslj = -11.0d0/2.0d0
          omegxx = -1.0d0/2.0d0
        end select

        coef1 = -2.0d0/3.0d0
        coef2 = -11.0d0
        coef3 = 8.0d0
        coef4 = -4.0d0
        coef5 = 2.0d0
        coef3i = 1.0d0/coef3
        coef4i = 1.0d0/coef4
        val1 = 19.0d0
        val2 = 2394.0d0
        val3 = 252.0d0
        val4 = 10.0d0
        val5 = 40.0d0
        val6 = 36.0d0
        val7 = 72.0d0
        val8 = 12.0d0
        val9 = 24.0d0
        val1i = 1.0d0/val1
        val2i = 1.0d0/val2
        val3i = 1.0d0/val3
        val4i = 1.0d0/val4
        val5i = 1.0d0/val5
        val6i = 1.0d0/val6
        val7i = 1.0d0/val7
        val8i = 1.0d0/val8
        val9i = 1.0d0/val9
        ww0 = 1.0d0/3.0d0
        ww1 = 1.0d0/18.0d0
        ww2 = 1.0d0/36.0d0
      end subroutine channel_para

      ! Device allocation: replaces allocarray (para.f90:418-447) for f,
      ! the macroscopic fields and the force.
      subroutine channel_alloc
        type(lbm_config) :: cfg
        cfg%lx = lx
        cfg%ly = ly
        cfg%lz = lz
        cfg%precision = precision
        cfg%device = device
        cfg%reserved = 0
        cfg%s1 = s1
        cfg%s2 = s2
        cfg%s4 = s4
        cfg%s9 = s9
        cfg%s10 = s10
        cfg%s13 = s13
        cfg%s16 = s16
        cfg%omegepsl = omegepsl
        cfg%omegepslj = omegepslj
        cfg%omegxx = omegxx
        cfg%coef1 = coef1
        cfg%coef2 = coef2
        cfg%coef3 = coef3
        cfg%coef4 = coef4
        cfg%coef5 = coef5
        cfg%coef3i = coef3i
        cfg%coef4i = coef4i
        cfg%val8 = val8
        cfg%val1i = val1i
        cfg%val2i = val2i
        cfg%val3i = val3i
        cfg%val4i = val4i
        cfg%val5i = val5i
        cfg%val6i = val6i
        cfg%val7i = val7i
        cfg%val8i = val8i
        cfg%val9i = val9i
        cfg%ww0 = ww0
        cfg%ww1 = ww1
        cfg%ww2 = ww2
        call lbm_check(lbm_create(cfg, lbm), 'lbm_create')
      end subroutine channel_alloc

      subroutine ensure_host_fields
        if (.not. allocated(rho)) then
          allocate(rho(lx,ly,lz), ux(lx,ly,lz), uy(lx,ly,lz), uz(lx,ly,lz))
        end if
      end subroutine ensure_host_fields

      subroutine ensure_host_f
        if (.not. allocated(f)) allocate(f(0:npop-1,lx,ly,lz))
      end subroutine ensure_host_f

      end module channel_vars
