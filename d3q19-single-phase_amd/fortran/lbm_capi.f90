!=============================================================================
! ISO_C_BINDING interface to liblbm (include/lbm.h).
!
! Each interface binds one C entry point; argument order and meaning are those
! of the header.  Status returns are 0 on success.
!=============================================================================
      module lbm_capi
      use iso_c_binding
      implicit none

      integer(c_int), parameter :: LBM_OK = 0
      integer(c_int), parameter :: LBM_FP64 = 0, LBM_FP32 = 1, LBM_FP16 = 2, LBM_FP16C = 3

      ! struct lbm_config
      type, bind(C) :: lbm_config
        integer(c_int) :: lx, ly, lz, precision, device, reserved
        real(c_double) :: s1, s2, s4, s9, s10, s13, s16
        real(c_double) :: omegepsl, omegepslj, omegxx
        real(c_double) :: coef1, coef2, coef3, coef4, coef5, coef3i, coef4i
        real(c_double) :: val8
        real(c_double) :: val1i, val2i, val3i, val4i, val5i, val6i,      &
                          val7i, val8i, val9i
        real(c_double) :: ww0, ww1, ww2
      end type lbm_config

      ! struct lbm_particle_config
      type, bind(C) :: lbm_particle_config
        integer(c_int) :: npart, maxlink
        real(c_double) :: rad, rhopart, mingap, mingap_w
        real(c_double) :: stf0, stf1, stf0_w, stf1_w
        real(c_double) :: visc, lub_cut, lub_min
        integer(c_int) :: pressure_force, nave
      end type lbm_particle_config

      integer(c_int), parameter :: LBM_MLP_TOTAL = 1453

      interface
        function lbm_last_error() bind(C, name="lbm_last_error")
          import :: c_ptr
          type(c_ptr) :: lbm_last_error
        end function

        integer(c_int) function lbm_create(cfg, h) bind(C, name="lbm_create")
          import :: c_int, c_ptr, lbm_config
          type(lbm_config), intent(in) :: cfg
          type(c_ptr), intent(out) :: h
        end function

        integer(c_int) function lbm_destroy(h) bind(C, name="lbm_destroy")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_upload_f(h, f) bind(C, name="lbm_upload_f")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), intent(in) :: f(*)
        end function

        integer(c_int) function lbm_download_f(h, f) bind(C, name="lbm_download_f")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), intent(out) :: f(*)
        end function

        integer(c_int) function lbm_set_force(h, fx, fy, fz) bind(C, name="lbm_set_force")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), value :: fx, fy, fz
        end function

        integer(c_int) function lbm_forcingp(h, istep, force_in_y) bind(C, name="lbm_forcingp")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          integer(c_int), value :: istep
          real(c_double), value :: force_in_y
        end function

        integer(c_int) function lbm_run_forcingp(h, istep0, nsteps, npforcing, force_in_y, fy_forcing) &
            bind(C, name="lbm_run_forcingp")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          integer(c_int), value :: istep0, nsteps, npforcing
          real(c_double), value :: force_in_y, fy_forcing
        end function

        integer(c_int) function lbm_set_macros(h, rho, ux, uy, uz) bind(C, name="lbm_set_macros")
          import :: c_int, c_ptr
          type(c_ptr), value :: h, rho, ux, uy, uz
        end function

        integer(c_int) function lbm_initpop(h) bind(C, name="lbm_initpop")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_collision_mrt(h) bind(C, name="lbm_collision_mrt")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_macrovar(h, rho, ux, uy, uz) bind(C, name="lbm_macrovar")
          import :: c_int, c_ptr
          type(c_ptr), value :: h, rho, ux, uy, uz
        end function

        integer(c_int) function lbm_rhoupdat(h, rhoerrmax) bind(C, name="lbm_rhoupdat")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), intent(out) :: rhoerrmax
        end function

        integer(c_int) function lbm_avedensity(h, rhomean) bind(C, name="lbm_avedensity")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), intent(out) :: rhomean
        end function

        integer(c_int) function lbm_run(h, nsteps) bind(C, name="lbm_run")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
          integer(c_int), value :: nsteps
        end function

        integer(c_int) function lbm_mass_sum(h, mass) bind(C, name="lbm_mass_sum")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), intent(out) :: mass
        end function

        integer(c_int) function lbm_statistc2(h, ustar, ystar, out) bind(C, name="lbm_statistc2")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), value :: ustar, ystar
          real(c_double), intent(out) :: out(*)
        end function

        integer(c_int) function lbm_sync(h) bind(C, name="lbm_sync")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        ! run monitor, vorticity, dissipation on the device (saveload.f90:1507-1676,
        ! :3929-4023, :4644-4882)
        integer(c_int) function lbm_diag(h, ustar, out) bind(C, name="lbm_diag")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), value :: ustar
          real(c_double), intent(out) :: out(13)
        end function

        integer(c_int) function lbm_vortcalc(h, ox, oy, oz) bind(C, name="lbm_vortcalc")
          import :: c_int, c_ptr
          type(c_ptr), value :: h, ox, oy, oz
        end function

        integer(c_int) function lbm_dissipation(h, tau, ystar, out) bind(C, name="lbm_dissipation")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), value :: tau, ystar
          real(c_double), intent(out) :: out(*)
        end function

        ! time-averaged statistics (running device sums)
        integer(c_int) function lbm_stats_accumulate(h) bind(C, name="lbm_stats_accumulate")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_stats_read(h, ustar, ystar, out, nsamples) &
            bind(C, name="lbm_stats_read")
          import :: c_int, c_ptr, c_double, c_long_long
          type(c_ptr), value :: h
          real(c_double), value :: ustar, ystar
          real(c_double), intent(out) :: out(*)
          integer(c_long_long), intent(out) :: nsamples
        end function

        integer(c_int) function lbm_stats_reset(h) bind(C, name="lbm_stats_reset")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        ! rigid particles (beads_* of var_inc.f90:166-168)
        integer(c_int) function lbm_particles_create(h, pc, pos, vel, omg) &
            bind(C, name="lbm_particles_create")
          import :: c_int, c_ptr, c_double, lbm_particle_config
          type(c_ptr), value :: h
          type(lbm_particle_config), intent(in) :: pc
          real(c_double), intent(in) :: pos(*), vel(*), omg(*)
        end function

        integer(c_int) function lbm_beads_links(h) bind(C, name="lbm_beads_links")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_beads_filling(h) bind(C, name="lbm_beads_filling")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_beads_collision(h) bind(C, name="lbm_beads_collision")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_beads_move(h) bind(C, name="lbm_beads_move")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_run_particles(h, istep0, nsteps) bind(C, name="lbm_run_particles")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
          integer(c_int), value :: istep0, nsteps
        end function

        ! any of the seven outputs may be c_null_ptr
        integer(c_int) function lbm_particles_get(h, pos, vel, omg, force, torque, contact, lub) &
            bind(C, name="lbm_particles_get")
          import :: c_int, c_ptr
          type(c_ptr), value :: h, pos, vel, omg, force, torque, contact, lub
        end function

        ! node flags (N ints): -1 fluid, p >= 0 inside particle p
        integer(c_int) function lbm_particles_solid(h, solid) bind(C, name="lbm_particles_solid")
          import :: c_int, c_ptr
          type(c_ptr), value :: h, solid
        end function

        ! learned closure (FP32/FP16 storage only)
        integer(c_int) function lbm_closure_set(h, params, ystar) bind(C, name="lbm_closure_set")
          import :: c_int, c_ptr, c_float, c_double
          type(c_ptr), value :: h
          real(c_float), intent(in) :: params(*)
          real(c_double), value :: ystar
        end function

        integer(c_int) function lbm_closure_clear(h) bind(C, name="lbm_closure_clear")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function
      end interface

      contains

      ! Message of the last failing liblbm call as a Fortran string.
      function lbm_error_string() result(msg)
        character(len=512) :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: cs(:)
        integer :: i
        msg = ' '
        p = lbm_last_error()
        if (.not. c_associated(p)) return
        call c_f_pointer(p, cs, [512])
        do i = 1, 512
          if (cs(i) == c_null_char) exit
          msg(i:i) = cs(i)
        end do
      end function lbm_error_string

      ! Stop the run with the library's message when a call failed.
      subroutine lbm_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc /= LBM_OK) then
          write(*,*) 'liblbm error in ', trim(where), ': ', trim(lbm_error_string())
          error stop 1
        end if
      end subroutine lbm_check

      end module lbm_capi
