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
!   closure_load   -> lbm_closure_set     (the learned closure; no reference
!                                          counterpart)
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

      ! The learned closure inside every following collision_MRT: the
      ! LBM_MLP_TOTAL FP32 parameters read from closurefile (stream access,
      ! raw floats), ystar of the run.
      subroutine closure_load
      use channel_vars
      implicit none
      real(c_float) :: params(LBM_MLP_TOTAL)
      integer :: u, ios
      open(newunit=u, file=trim(closurefile), access='stream', form='unformatted', status='old', &
           action='read', iostat=ios)
      if (ios /= 0) then
        write(*,*) 'cannot open closure file ', trim(closurefile)
        error stop 2
      end if
      read(u, iostat=ios) params
      close(u)
      if (ios /= 0) then
        write(*,*) 'closure file ', trim(closurefile), ' holds fewer than ', LBM_MLP_TOTAL, ' values'
        error stop 2
      end if
      call lbm_check(lbm_closure_set(lbm, params, ystar), 'lbm_closure_set')
      end subroutine closure_load

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

      ! Particles (the beads_* routines of var_inc.f90:166-168): create npart
      ! spheres at the centres listed in partfile, at rest, with the contact
      ! constants of para.f90:357-360, the lubrication correction at the
      ! fluid viscosity, and avedensity after every 100th step
      ! (main.f90:163-167).
      subroutine beads_init
      use channel_vars
      implicit none
      type(lbm_particle_config) :: pc
      integer :: u, i
      allocate(ypglb(3,npart), wp(3,npart), omgp(3,npart), fHIp(3,npart), torqp(3,npart), flubp(3,npart))
      open(newunit=u, file=trim(partfile), status='old', action='read')
      do i = 1, npart
        read(u, *) ypglb(:, i)
      end do
      close(u)
      wp = 0.0d0
      omgp = 0.0d0
      pc = lbm_particle_config(npart, 0, rad, rhopart, mingap, mingap_w, 0.025d0, 0.002d0, 0.025d0, 0.002d0, &
                               visc, 2.0d0/3.0d0, 0.01d0, 1, 100)
      call lbm_check(lbm_particles_create(lbm, pc, ypglb, wp, omgp), 'beads_init')
      end subroutine beads_init

      ! n particle-laden main-loop steps from istep (beads_collision,
      ! collision_MRT, macrovar, beads_move, beads_links, beads_filling).
      subroutine run_particle_steps(n)
      use channel_vars
      implicit none
      integer, intent(in) :: n
      call lbm_check(lbm_run_particles(lbm, int(istep - 1, c_int), int(n, c_int)), 'run_particle_steps')
      end subroutine run_particle_steps

      ! The particle state into ypglb, wp, omgp, fHIp, torqp, flubp.
      subroutine fetch_particles
      use channel_vars
      implicit none
      call lbm_check(lbm_particles_get(lbm, c_loc(ypglb), c_loc(wp), c_loc(omgp), c_loc(fHIp), c_loc(torqp), &
                     c_null_ptr, c_loc(flubp)), 'fetch_particles')
      end subroutine fetch_particles
