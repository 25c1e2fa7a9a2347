!=============================================================================
! Checkpoints and monitors in the reference's file formats.
!
! Files are Fortran unformatted sequential, one per rank (suffix .NNN):
!   initflow/finit.000                      istat ; f          (saveload.f90:102-124)
!   cntdflow/endrunflow2D16x8.SSSSSSS.000   istep, istat, imovie ; f
!                                                              (saveload.f90:196-231)
! f is the reference-layout f(0:18,lx,ly,lz) in FP64, moved to and from the
! device with lbm_download_f / lbm_upload_f.
! diag.dat and profiles2.dat follow saveload.f90:1507-1676 and :1348-1502.
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

      function cntd_name(step) result(fnm)
      use channel_vars
      implicit none
      integer, intent(in) :: step
      character(len=300) :: fnm
      write(fnm, '(a,a,i7.7,a)') trim(dirgenr), 'cntdflow/endrunflow2D16x8.', step, '.000'
      end function cntd_name

      subroutine savecntdflow
      use channel_vars
      implicit none
      character(len=300) :: cntd_name
      call ensure_host_f
      call lbm_check(lbm_download_f(lbm, f), 'savecntdflow')
      open(12, file=trim(cntd_name(istep)), status='unknown', form='unformatted')
      write(12) istep, istat, imovie
      write(12) f
      close(12)
      end subroutine savecntdflow

      subroutine loadcntdflow
      use channel_vars
      implicit none
      character(len=300) :: cntd_name
      call ensure_host_f
      open(12, file=trim(cntd_name(istpload)), status='old', form='unformatted')
      read(12) istep0, istat, imovie
      read(12) f
      close(12)
      call lbm_check(lbm_upload_f(lbm, f), 'loadcntdflow')
      end subroutine loadcntdflow

      ! Run monitor, saveload.f90:1535-1668 (all nodes fluid, one rank).
      subroutine diag
      use channel_vars
      implicit none
      real(8) :: umean, vmean, wmean, urms, vrms, wrms, vmax, vel, rmax, rmin, rn
      integer :: i, j, k, im, jm, km
      call fetch_fields
      rn = dble(lx)*dble(ly)*dble(lz)
      umean = sum(ux)/rn
      vmean = sum(uy)/rn
      wmean = sum(uz)/rn
      urms = sqrt(sum(ux*ux)/rn - umean**2)
      vrms = sqrt(sum(uy*uy)/rn - vmean**2)
      wrms = sqrt(sum(uz*uz)/rn - wmean**2)
      vmax = 0.0d0
      im = 0
      jm = 0
      km = 0
      do k = 1, lz
        do j = 1, ly
          do i = 1, lx
            vel = sqrt(ux(i,j,k)*ux(i,j,k) + uy(i,j,k)*uy(i,j,k) + uz(i,j,k)*uz(i,j,k))
            if (vel > vmax) then
              vmax = vel
              im = i
              jm = j
              km = k
            end if
          end do
        end do
      end do
      rmax = maxval(rho)
      rmin = minval(rho)
      open(26, file=trim(dirgenr)//'diag/diag.dat', status='unknown', form='formatted', position='append')
      write(26, 260) dble(istep), vmax, im, jm, km, umean/ustar, vmean/ustar, wmean/ustar,  &
                     urms/ustar, vrms/ustar, wrms/ustar, 0.0d0, rmax, rmin
      close(26)
      write(*,*) istep, rmax, rmin
260   format(2x, 2(1pe16.6), 3I6, 9(1pe16.6))
      end subroutine diag

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
