!=============================================================================
! Checkpoints and monitors in the reference's file formats.
!
! Files are Fortran unformatted sequential, one per rank (suffix .NNN):
!   initflow/finit.000                      istat ; f          (saveload.f90:102-124)
!   cntdflow/endrunflow2D16x8.SSSSSSS.NNN   istep, istat, imovie ; f
!                                                              (saveload.f90:196-231)
! f is the reference-layout f(0:18,lx,ly,lz) in FP64, moved to and from the
! device with lbm_download_f / lbm_upload_f.  Checkpoints are read from and
! written as the set of files of an nprocY x nprocZ run (each rank its y/z
! block of f, para.f90:218-262), one file when both are 1.
! diag.dat, diag2.dat, profiles.dat, profiles2.dat and dissipation.dat
! follow saveload.f90:1507-1676, :1202-1342, :1348-1502 and :4644-4882; flowout/uy_profile and
! flowout/pressCut follow outputflow (:696-714, :845-947, :1025-1120);
! probe/probeN.dat follows probe (:4059-4100).
!=============================================================================
      subroutine saveinitflow
      use channel_vars
      implicit none
      call ensure_host_f
      call lbm_check(lbm_download_f(lbm, f), 'saveinitflow')
      open(10, file=trim(dirgenr)//'initflow/finit.000', status='unknown', form='unformatted')
      write(10) istat
      write(10) f
      close(10)
      end subroutine saveinitflow

      subroutine loadinitflow
      use channel_vars
      implicit none
      call ensure_host_f
      open(10, file=trim(dirgenr)//'initflow/finit.000', status='old', form='unformatted')
      read(10) istat
      read(10) f
      close(10)
      call lbm_check(lbm_upload_f(lbm, f), 'loadinitflow')
      end subroutine loadinitflow

      function cntd_name(step, myid) result(fnm)
      use channel_vars
      implicit none
      integer, intent(in) :: step, myid
      character(len=300) :: fnm
      write(fnm, '(a,a,i7.7,a,i3.3)') trim(dirgenr), 'cntdflow/endrunflow2D16x8.', step, '.', myid
      end function cntd_name

      ! The y/z block of rank myid of an npy x npz run (para.f90:218-262):
      ! indy = mod(myid, npy), indz = myid/npy; the first mod(ly, npy)
      ! columns (mod(lz, npz) rows) hold one node more.
      subroutine rank_block(myid, npy, npz, y0, nyr, z0, nzr)
      use channel_vars
      implicit none
      integer, intent(in) :: myid, npy, npz
      integer, intent(out) :: y0, nyr, z0, nzr
      integer :: indy, indz, i
      indy = mod(myid, npy)
      indz = myid/npy
      y0 = 0
      do i = 0, indy - 1
        y0 = y0 + ly/npy
        if (i < mod(ly, npy)) y0 = y0 + 1
      end do
      nyr = ly/npy
      if (indy < mod(ly, npy)) nyr = nyr + 1
      z0 = 0
      do i = 0, indz - 1
        z0 = z0 + lz/npz
        if (i < mod(lz, npz)) z0 = z0 + 1
      end do
      nzr = lz/npz
      if (indz < mod(lz, npz)) nzr = nzr + 1
      end subroutine rank_block

      ! savecntdflow (saveload.f90:196-231): one file per rank of an
      ! nprocy_out x nprocz_out run, so that run can resume from it.
      subroutine savecntdflow
      use channel_vars
      implicit none
      character(len=300) :: cntd_name
      integer :: myid, y0, nyr, z0, nzr
      call ensure_host_f
      call lbm_check(lbm_download_f(lbm, f), 'savecntdflow')
      do myid = 0, nprocy_out*nprocz_out - 1
        call rank_block(myid, nprocy_out, nprocz_out, y0, nyr, z0, nzr)
        open(12, file=trim(cntd_name(istep, myid)), status='unknown', form='unformatted')
        write(12) istep, istat, imovie
        write(12) f(:, :, y0+1:y0+nyr, z0+1:z0+nzr)
        close(12)
      end do
      end subroutine savecntdflow

      ! loadcntdflow (saveload.f90:296-332) from the checkpoint set of an
      ! nprocy_in x nprocz_in run, assembled into the one device lattice (the
      ! re-decomposition of loadcntdflow_frmmore, :336-381, for any y/z split).
      subroutine loadcntdflow
      use channel_vars
      implicit none
      character(len=300) :: cntd_name
      integer :: myid, y0, nyr, z0, nzr, st0, st1, st2
      real(8), allocatable :: blk(:,:,:,:)
      call ensure_host_f
      do myid = 0, nprocy_in*nprocz_in - 1
        call rank_block(myid, nprocy_in, nprocz_in, y0, nyr, z0, nzr)
        allocate(blk(0:npop-1, lx, nyr, nzr))
        open(12, file=trim(cntd_name(istpload, myid)), status='old', form='unformatted')
        read(12) st0, st1, st2
        read(12) blk
        close(12)
        if (myid == 0) then
          istep0 = st0
          istat = st1
          imovie = st2
        else if (st0 /= istep0) then
          write(*,*) 'loadcntdflow: rank', myid, ' holds step', st0, ', rank 0 step', istep0
          error stop 3
        end if
        f(:, :, y0+1:y0+nyr, z0+1:z0+nzr) = blk
        deallocate(blk)
      end do
      call lbm_check(lbm_upload_f(lbm, f), 'loadcntdflow')
      end subroutine loadcntdflow

      ! Run monitor, saveload.f90:1507-1676, from device reductions over the
      ! fluid nodes (no field download): diag.dat, and diag2.dat with the
      ! fastest and fastest-spinning particle when there are particles.
      subroutine diag
      use channel_vars
      implicit none
      real(8) :: out(13), wpmax, omgpmax, mag
      integer :: i, idwp, idomgp
      call lbm_check(lbm_diag(lbm, ustar, out), 'diag')
      open(26, file=trim(dirgenr)//'diag/diag.dat', status='unknown', form='formatted', position='append')
      write(26, 260) dble(istep), out(1), nint(out(2)), nint(out(3)), nint(out(4)), out(5:13)
      close(26)
      write(*,*) istep, out(12), out(13)
      if (ipart) then
        call fetch_particles
        wpmax = 0.0d0
        omgpmax = 0.0d0
        idwp = 0
        idomgp = 0
        do i = 1, npart
          mag = sqrt(wp(1,i)**2 + wp(2,i)**2 + wp(3,i)**2)
          if (mag > wpmax) then
            wpmax = mag
            idwp = i
          end if
          mag = sqrt(omgp(1,i)**2 + omgp(2,i)**2 + omgp(3,i)**2)
          if (mag > omgpmax) then
            omgpmax = mag
            idomgp = i
          end if
        end do
        open(58, file=trim(dirgenr)//'diag/diag2.dat', status='unknown', form='formatted', position='append')
        if (idwp > 0 .and. idomgp > 0) then
          write(58, 370) dble(istep), wpmax, idwp, ypglb(:, idwp), omgpmax, idomgp, ypglb(:, idomgp)
        else
          write(58, 370) dble(istep), wpmax, idwp, 0.0d0, 0.0d0, 0.0d0, omgpmax, idomgp, 0.0d0, 0.0d0, 0.0d0
        end if
        close(58)
      end if
260   format(2x, 2(1pe16.6), 3I6, 9(1pe16.6))
370   format(2x, 2(1pe16.6), 1I6, 3(1pe16.6), 1(1pe16.6), 1I6, 3(1pe16.6))
      end subroutine diag

      ! dissipation_calc, saveload.f90:4644-4882, on the device: plane means
      ! of S_ij S_ij and S_xy by the moment and population methods.
      subroutine dissipation_calc
      use channel_vars
      implicit none
      real(8), allocatable :: prof(:,:)
      integer :: i
      allocate(prof(6, lx))
      call lbm_check(lbm_dissipation(lbm, tau, ystar, prof), 'dissipation_calc')
      open(127, file=trim(dirgenr)//'stat/dissipation.dat', status='unknown', form='formatted', position='append')
      write(127,*) istep
      do i = 1, lx
        write(127, 985) prof(:, i)
      end do
      close(127)
      deallocate(prof)
985   format(2x, 6(1pe16.6))
      end subroutine dissipation_calc

      ! vortcalc, saveload.f90:3929-4023, on the device into ox, oy, oz.
      subroutine vortcalc
      use channel_vars
      implicit none
      if (.not. allocated(ox)) allocate(ox(lx,ly,lz), oy(lx,ly,lz), oz(lx,ly,lz))
      call lbm_check(lbm_vortcalc(lbm, c_loc(ox), c_loc(oy), c_loc(oz)), 'vortcalc')
      end subroutine vortcalc

      ! outputflow, saveload.f90:696-714: outputuy and outputpress from the
      ! fields of the current step (downloaded, as the reference gathers them
      ! on rank 0).
      subroutine outputflow
      use channel_vars
      implicit none
      call fetch_fields
      call outputuy
      call outputpress
      end subroutine outputflow

      ! outputuy, saveload.f90:845-947: flowout/uy_profileSSSSSSS.dat holds,
      ! for each wall-normal node of the lower half, its index, its distance
      ! from the centre plane, the time in viscous units of the half width,
      ! the streamwise velocity / u* along y at z = lz/2, and the laminar
      ! start-up and steady solutions / u* (:920-936).
      subroutine outputuy
      use channel_vars
      implicit none
      character(len=300) :: fnm
      integer :: i, k, n
      real(8) :: time1, xx0, uuss, uut
      write(fnm, '(a,a,i7.7,a)') trim(dirgenr), 'flowout/uy_profile', istep, '.dat'
      open(17, file=trim(fnm), status='unknown', form='formatted')
      time1 = istep*visc/((real(lx, 8)/2.d0)**2)
      k = lz/2
      do i = 1, lx/2
        xx0 = abs(real(i, 8) - 0.5d0 - real(lx, 8)/2.d0)
        uuss = (4*ustar)/(real(lx, 8)**2)*((real(lx, 8)/2)**2 - xx0**2)
        uut = 0.0d0
        do n = 0, 25
          uut = uut + (4*(-1)**n)/(pi*(n + 0.5d0))**3*exp(-((real(n, 8) + 0.5d0)**2)*(pi**2)*time1)* &
                dcos((real(n, 8) + 0.5d0)*pi*(xx0/(real(lx, 8)/2.d0)))
        end do
        uut = (1 - (xx0/(real(lx, 8)/2.d0))**2) - uut
        write(17, 170) real(i, 8), xx0, time1, uy(i, :, k)/ustar, uut, uuss/ustar
      end do
      close(17)
170   format(2x, 8(1pe16.6))
      end subroutine outputuy

      ! outputpress, saveload.f90:1025-1120: flowout/pressCutSSSSSSS.dat holds
      ! (y, z, rho/3) over every x and y node of the plane z = lx/2 (the
      ! reference's k = nx/2), x outermost; nothing is written when lx/2 > lz,
      ! where that plane does not exist.
      subroutine outputpress
      use channel_vars
      implicit none
      character(len=300) :: fnm
      integer :: i, j, k
      real(8) :: yy0, zz0
      k = lx/2
      if (k < 1 .or. k > lz) then
        write(*,*) 'outputpress: no plane z = lx/2 =', k, ' in lz =', lz, '; pressCut not written'
        return
      end if
      write(fnm, '(a,a,i7.7,a)') trim(dirgenr), 'flowout/pressCut', istep, '.dat'
      open(20, file=trim(fnm), status='unknown', form='formatted')
      zz0 = dble(k) - 0.5d0
      do i = 1, lx
        do j = 1, ly
          yy0 = dble(j) - 0.5d0
          write(20, 120) yy0, zz0, rho(i, j, k)/3.0d0
        end do
      end do
      close(20)
120   format(2x, 3(1pe16.6))
      end subroutine outputpress

      ! The output directories under dirgenr, as para.f90:310-327 makes them
      ! (libc mkdir; an existing directory is fine).
      subroutine make_dirs
      use channel_vars
      implicit none
      interface
        integer(c_int) function c_mkdir(path, mode) bind(C, name="mkdir")
          import :: c_int, c_char
          character(kind=c_char), intent(in) :: path(*)
          integer(c_int), value :: mode
        end function
      end interface
      character(len=8), parameter :: names(6) = [character(len=8) :: 'diag', 'stat', 'probe', 'initflow', &
                                                 'cntdflow', 'flowout']
      integer :: k, rc
      do k = 1, size(names)
        rc = c_mkdir(trim(dirgenr)//trim(names(k))//c_null_char, int(o'755', c_int))
      end do
      end subroutine make_dirs

      ! probe, saveload.f90:4059-4100: probe/probeN.dat holds the velocity at
      ! the centre node (lx/2, ny_r/2, nz_r/2 of its block) of every rank of
      ! the nprocy_out x nprocz_out decomposition, one line per rank (one
      ! line for a single rank), 0 where that node lies inside a particle.
      subroutine probe
      use channel_vars
      implicit none
      character(len=300) :: fnm
      character(len=25) :: filenum
      integer(c_int), allocatable, target :: solid(:)
      integer :: id, y0, nyr, z0, nzr, jy, kz, n
      real(8) :: up, vp, wp9
      call fetch_fields
      if (ipart) then
        allocate(solid(int(lx, 8)*ly*lz))
        call lbm_check(lbm_particles_solid(lbm, c_loc(solid)), 'particles_solid')
      end if
      write(filenum, '(i10)') istep
      fnm = trim(dirgenr)//'probe/probe'//trim(adjustl(filenum))//'.dat'
      open(60, file=trim(fnm), status='unknown', form='formatted')
      write(60,*) 'Center probe at istep ', istep
      write(60,*) 'Processor    x-velocity    y-velocity    z-velocity'
      do id = 0, nprocy_out*nprocz_out - 1
        call rank_block(id, nprocy_out, nprocz_out, y0, nyr, z0, nzr)
        jy = y0 + nyr/2
        kz = z0 + nzr/2
        up = ux(lx/2, jy, kz)
        vp = uy(lx/2, jy, kz)
        wp9 = uz(lx/2, jy, kz)
        if (ipart) then
          n = lx/2 + lx*((jy - 1) + ly*(kz - 1))
          if (solid(n) >= 0) then
            up = 0.0d0
            vp = 0.0d0
            wp9 = 0.0d0
          end if
        end if
        write(60, 600) id, up, vp, wp9
      end do
      close(60)
      if (allocated(solid)) deallocate(solid)
600   format(2x, 1i5, 3(1pe16.6))
      end subroutine probe

      ! statistc, saveload.f90:1202-1342: plane means of u / u*, the velocity
      ! variances and shear stresses / u*^2, and the mean and variance of the
      ! density fluctuation (/ u*^2, / u*^4) into stat/profiles.dat, from the
      ! fields macrovar forms of the current lattice (the reference calls it
      ! once, at the start).
      subroutine statistc
      use channel_vars
      implicit none
      real(8) :: vxave(lx), vyave(lx), vzave(lx), vxsq(lx), vysq(lx), vzsq(lx)
      real(8) :: sxz(lx), sxy(lx), syz(lx), prave(lx), prsq(lx), yplus, area
      integer :: i
      call fetch_fields
      do i = 1, lx
        vxave(i) = sum(ux(i,:,:))
        vyave(i) = sum(uy(i,:,:))
        vzave(i) = sum(uz(i,:,:))
        sxz(i) = sum(ux(i,:,:)*uz(i,:,:))
        sxy(i) = sum(ux(i,:,:)*uy(i,:,:))
        syz(i) = sum(uy(i,:,:)*uz(i,:,:))
        vxsq(i) = sum(ux(i,:,:)**2)
        vysq(i) = sum(uy(i,:,:)**2)
        vzsq(i) = sum(uz(i,:,:)**2)
        prave(i) = sum(rho(i,:,:))
        prsq(i) = sum(rho(i,:,:)*rho(i,:,:))
      end do
      area = dble(ly)*dble(lz)
      vxave = vxave/area/ustar
      vyave = vyave/area/ustar
      vzave = vzave/area/ustar
      vxsq = vxsq/area/ustar**2 - vxave**2
      vysq = vysq/area/ustar**2 - vyave**2
      vzsq = vzsq/area/ustar**2 - vzave**2
      sxz = sxz/area/ustar**2 - vxave*vzave
      sxy = sxy/area/ustar**2 - vxave*vyave
      syz = syz/area/ustar**2 - vyave*vzave
      prave = prave/area/ustar**2
      prsq = prsq/area/ustar**4 - prave*prave
      open(27, file=trim(dirgenr)//'stat/profiles.dat', status='unknown', form='formatted', position='append')
      write(27,*) istep
      do i = 1, lx
        yplus = (dble(i) - 0.5d0)/ystar
        if (i > lx/2) yplus = (dble(lx - i) + 0.5d0)/ystar
        write(27, 460) dble(i) - 0.5d0, yplus, vxave(i), vyave(i), vzave(i), vxsq(i), vysq(i), vzsq(i), &
                       sxz(i), sxy(i), syz(i), prave(i), prsq(i)
      end do
      close(27)
460   format(2x, 13(1pe15.6))
      end subroutine statistc

      ! Plane profiles, saveload.f90:1348-1502, computed on the device.
      subroutine statistc2
      use channel_vars
      implicit none
      real(8), allocatable :: prof(:,:)
      integer :: i
      allocate(prof(15, lx))
      call lbm_check(lbm_statistc2(lbm, ustar, ystar, prof), 'statistc2')
      open(27, file=trim(dirgenr)//'stat/profiles2.dat', status='unknown', form='formatted', position='append')
      write(27,*) istep
      do i = 1, lx
        write(27, 460) prof(:, i)
      end do
      close(27)
      deallocate(prof)
460   format(2x, 15(1pe15.6))
      end subroutine statistc2
