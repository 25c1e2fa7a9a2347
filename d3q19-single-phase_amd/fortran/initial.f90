!=============================================================================
! Initial conditions.
!   initvel  Channel-Flow/initial.f90:75-147: law-of-the-wall streamwise
!            profile (linear below y+ = 10.8, ln(y+)/0.41 + 5 above), mirrored
!            about the centre plane; the reference's perturbation amplitude is
!            zero (initial.f90:84), so none is added.  Laminar runs start at rest.
!   initpop  Channel-Flow/initial.f90:19-46: rho = 0 and f = equilibrium of u,
!            evaluated on the device.
! The law-of-the-wall loop restates initial.f90:104-115 (GPL-3) closely, as
! bitwise-equal initial fields require; it keeps that file's licence.
!=============================================================================
      subroutine initvel
      use channel_vars
      implicit none
      integer :: i
      real(8) :: yplus, u9
      call ensure_host_fields
      ux = 0.0d0
      uy = 0.0d0
      uz = 0.0d0
      if (.not. laminarflow) then
        do i = 1, lx/2
          yplus = (dble(i) - 0.5d0)/ystar
          if (yplus < 10.8d0) then
            uy(i,:,:) = yplus*ustar
            uy(lx+1-i,:,:) = yplus*ustar
          else
            u9 = log(yplus)/0.41d0 + 5.0d0
            u9 = u9*ustar
            uy(i,:,:) = u9
            uy(lx+1-i,:,:) = u9
          end if
        end do
      end if
      call lbm_check(lbm_set_macros(lbm, c_null_ptr, c_loc(ux), c_loc(uy), c_loc(uz)), 'initvel')
      end subroutine initvel

      subroutine initpop
      use channel_vars
      implicit none
      call lbm_check(lbm_initpop(lbm), 'initpop')
      end subroutine initpop
