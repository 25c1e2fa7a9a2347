!=============================================================================
! Channel driver: the call sequence of Channel-Flow/main.f90:41-231 on one
! GPU.  Run control comes from the namelist file named on the command line
! (default channel.nml), e.g.
!   &channel lx=32, ly=64, lz=32, nsteps=1000, newinitflow=.false. /
! The throughput of the main loop is printed as MLUPS (time_max of
! main.f90:214-231, measured around the same loop).
!=============================================================================
      program channel_main
      use channel_vars
      implicit none
      character(len=300) :: nml
      integer :: u, ios, nrun
      integer(8) :: t0, t1, rate
      real(8) :: time_max, mass

      nml = 'channel.nml'
      if (command_argument_count() >= 1) call get_command_argument(1, nml)
      open(newunit=u, file=trim(nml), status='old', action='read', iostat=ios)
      if (ios /= 0) then
        write(*,*) 'cannot open namelist file ', trim(nml)
        error stop 2
      end if
      read(u, nml=channel)
      close(u)

      call channel_para
      call make_dirs
      call channel_alloc
      istat = 0
      imovie = 0
      istep0 = 0

      ! main.f90:52-121
      if (newrun) then
        if (newinitflow) then
          call initvel
          call FORCING
          call initpop
          istep = 0
          do
            call rhoupdat
            call collision_MRT
            write(*,*) istep, rhoerrmax
            if (rhoerrmax <= rhoepsl .or. istep > 15000) then
              write(*,*) 'final relaxation => ', istep, rhoerrmax
              exit
            end if
            istep = istep + 1
          end do
          call saveinitflow
          call macrovar
          istep0 = 0
          istep = istep0
          call statistc
          call statistc2
        else
          call loadinitflow
          istep0 = 0
          istep = istep0
          call statistc
          call statistc2
        end if
      else
        call loadcntdflow
      end if

      ! main.f90:132-137
      call FORCING
      call macrovar
      if (len_trim(closurefile) > 0) call closure_load
      if (ipart) then
        if (npforcing > 0) then
          write(*,*) 'the perturbation forcing is not used with particles (main.f90:148)'
          error stop 2
        end if
        call beads_init
      end if
      call lbm_check(lbm_sync(lbm), 'sync')
      call system_clock(t0, rate)

      ! main.f90:142-208.  Steps with nothing between them but collision_MRT
      ! and macrovar run as one lbm_run call up to the next diag/statistc2
      ! step (the library runs pairs of them as one blocked sweep; the
      ! results are those of single steps).
      istep = istep0 + 1
      do while (istep <= istep0 + nsteps)
        if (istep <= npforcing) then
          ! perturbation forcing until npforcing (the block at main.f90:148-153):
          ! FORCINGP or FORCING, collision_MRT and macrovar per step, batched
          ! up to npforcing and the next output step
          nrun = min(npforcing, istep0 + nsteps) - istep + 1
          if (ndiag > 0) nrun = min(nrun, ndiag - mod(istep - 1, ndiag))
          if (nstat > 0) nrun = min(nrun, nstat - mod(istep - 1, nstat))
          if (ndiss > 0) nrun = min(nrun, ndiss - mod(istep - 1, ndiss))
          if (nflowout > 0) nrun = min(nrun, nflowout - mod(istep - 1, nflowout))
          call run_forcingp_steps(nrun)
        else
          nrun = istep0 + nsteps - istep + 1
          if (ndiag > 0) nrun = min(nrun, ndiag - mod(istep - 1, ndiag))
          if (nstat > 0) nrun = min(nrun, nstat - mod(istep - 1, nstat))
          if (ndiss > 0) nrun = min(nrun, ndiss - mod(istep - 1, ndiss))
          if (nflowout > 0) nrun = min(nrun, nflowout - mod(istep - 1, nflowout))
          if (ipart) then
            ! the particle-laden loop body, avedensity after every 100th
            ! step included (main.f90:157-167)
            call run_particle_steps(nrun)
          else
            call run_steps(nrun)
          end if
        end if
        istep = istep + nrun - 1
        if (ndiag > 0) then
          if (mod(istep, ndiag) == 0) call diag
        end if
        if (nstat > 0) then
          if (mod(istep, nstat) == 0) call statistc2
        end if
        if (ndiss > 0) then
          if (mod(istep, ndiss) == 0) call dissipation_calc
        end if
        ! main.f90:184
        if (nflowout > 0) then
          if (mod(istep, nflowout) == 0) call outputflow
        end if
        istep = istep + 1
      end do
      istep = istep0 + nsteps

      call lbm_check(lbm_sync(lbm), 'sync')
      call system_clock(t1)
      time_max = dble(t1 - t0)/dble(rate)
      call lbm_check(lbm_mass_sum(lbm, mass), 'mass')
      call probe
      if (savecntd) call savecntdflow
      if (ipart) then
        call fetch_particles
        open(newunit=u, file=trim(dirgenr)//'partfinal.dat', status='unknown', form='formatted')
        do ios = 1, npart
          write(u, '(9(1pe24.16))') ypglb(:, ios), wp(:, ios), omgp(:, ios)
        end do
        close(u)
      end if
      write(*,*) 'time_max = ', time_max
      write(*,'(a,f12.2)') ' MLUPS = ', dble(lx)*dble(ly)*dble(lz)*dble(nsteps)/time_max/1.0d6
      write(*,*) 'mass = ', mass
      call lbm_check(lbm_destroy(lbm), 'destroy')
      end program channel_main
