!=============================================================================
! Hot-path entry points under the reference's names.
!
! Channel-Flow/collision.f90 defines argument-free subroutines acting on the
! module arrays; these keep the names, the absence of arguments and the call
! order, and forward to liblbm on the device lattice held in channel_vars.
!   collision_MRT  -> lbm_collision_mrt   (collision.f90:24-273, exchange :281-372)
!   macrovar       -> lbm_macrovar        (collision.f90:378-463)
!   rhoupdat       -> lbm_rhoupdat        (collision.f90:469-480)
!   avedensity     -> lbm_avedensity      (collision.f90:487-513)
!   FORCING        -> lbm_set_force       (collision.f90:515-527)
!   FORCINGP       -> lbm_forcingp        (collision.f90:529-602)
!   run_steps(n)   -> lbm_run             (n x collision_MRT + macrovar)
! macrovar leaves the fields on the device; fetch_fields copies them to the
! host arrays when an I/O routine needs them.
!=============================================================================
      subroutine collision_MRT
      use channel_vars
      implicit none
      call lbm_check(lbm_collision_mrt(lbm), 'collision_MRT')
      end subroutine collision_MRT

      subroutine macrovar
      use channel_vars
      implicit none
      call lbm_check(lbm_macrovar(lbm, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'macrovar')
      end subroutine macrovar

      ! n iterations of collision_MRT followed by macrovar (main.f90:157-161)
      ! in one library call.
      subroutine run_steps(n)
      use channel_vars
      implicit none
      integer, intent(in) :: n
      call lbm_check(lbm_run(lbm, int(n, c_int)), 'run_steps')
      end subroutine run_steps

      ! n main-loop steps from istep with the perturbation forcing switched on
      ! (FORCINGP before each collision while istep < npforcing, FORCING at
      ! npforcing; main.f90:148-153), as one library call.
      subroutine run_forcingp_steps(n)
      use channel_vars
      implicit none
      integer, intent(in) :: n
      call lbm_check(lbm_run_forcingp(lbm, int(istep - 1, c_int), int(n, c_int), int(npforcing, c_int), &
                     force_in_y, force_in_y*force_mag), 'run_forcingp_steps')
      end subroutine run_forcingp_steps

      ! macrovar with the fields copied into the host arrays rho, ux, uy, uz.
      subroutine fetch_fields
      use channel_vars
      implicit none
      call ensure_host_fields
      call lbm_check(lbm_macrovar(lbm, c_loc(rho), c_loc(ux), c_loc(uy), c_loc(uz)), 'macrovar')
      end subroutine fetch_fields

      ! rhoupdat; also leaves rhoerrmax = max|rho_new - rho_old|, the
      ! convergence measure of the pre-relaxation loop (main.f90:79-81).
      subroutine rhoupdat
      use channel_vars
      implicit none
      call lbm_check(lbm_rhoupdat(lbm, rhoerrmax), 'rhoupdat')
      end subroutine rhoupdat

      subroutine avedensity
      use channel_vars
      implicit none
      real(8) :: rhomean
      call lbm_check(lbm_avedensity(lbm, rhomean), 'avedensity')
      end subroutine avedensity

      subroutine FORCING
      use channel_vars
      implicit none
      call lbm_check(lbm_set_force(lbm, 0.0d0, force_in_y*force_mag, 0.0d0), 'FORCING')
      end subroutine FORCING

      ! Perturbation forcing at the current istep (mean force_in_y plus the
      ! banded divergence-free perturbation); the device evaluates it per node.
      subroutine FORCINGP
      use channel_vars
      implicit none
      call lbm_check(lbm_forcingp(lbm, int(istep, c_int), force_in_y), 'FORCINGP')
      end subroutine FORCINGP
