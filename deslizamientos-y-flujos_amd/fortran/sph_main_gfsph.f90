!  SPH_MAIN_GFSPH -- drop-in replacement of the reference's
!  time_integration_sph ("7 SPH_MAIN_2019.f90":220-283) that advances the
!  Geoflow-SPH module state on the GPU through the C ABI (gfsph_c).
!
!  It is compiled together with the reference's modules: it reads the
!  particle arrays of SPH_Main_Vars_2019 ("5 SPH_Main_Vars_2019.f90":29-80),
!  the material vector const(:) and pore-pressure switches of
!  SPH_SW_Vars_2019 ("4 SPH_SW_Vars_2019.f90":17-27, 196), the DEM of
!  SPH_FEM_topo_2019 ("2 SPH_FEM_topo_2019.f90":94-110) and dt_sph of the
!  driver time variables, exactly as the reference step does, so the
!  driver's call site keeps its no-argument form:
!
!      call time_integration_sph_gfsph          ! instead of time_integration_sph
!      time_sph = time_sph + dt_sph
!
!  The first call uploads the state (already initialised by Init_sph, so
!  init_input = 0) and the DEM with the reference's own nodal slopes
!  topol(1:9,:); every call advances one step.  The particle state stays on
!  the GPU between steps and is copied back into the module arrays only when
!  the driver is about to read them: when its print, save or plot counter
!  fires after this step (t_print_reset >= time_print, t_save_reset >=
!  time_save, t_plot_reset >= time_plot -- Driver_Flow and Driver_Flow_23
!  advance the counters before the step, "8 SPH_FEM_Driver_2019.f90":1578-1580,
!  1488-1492, and test them after it, 1663-1686, 1500-1525), in a case-table
!  run (stop_case_SPH reads vx and If_Out_Domain after every step), and on
!  the step after which the driver's loop ends (time reaching time_end,
!  maxtimestep).  A caller that reads the arrays at other times calls
!  gfsph_sync_state first, or sets gfsph_pull_every_step.  The build must use
!  the same real kind as the reference modules (irk = double):
!  -fdefault-real-8.
!
!  With ic_adapt_dt_sph /= 0 the context performs Adaptive_dt_sph_SW
!  ("6 SPH_SW_2019.f90":6521-6664) on the device at the start of every step
!  and each call returns the dt_sph that step used, so the driver drops its
!  own "call adaptive_dt_sph" (8:1590) and keeps "time_sph = time_sph + dt_sph".
!
!  With a pore-pressure column (icpwp = 2) the whole Uaux(1:nUaux,:) goes to
!  the context at the first call and comes back with the state.
!
!  With erosion (ic_erosion = 1) the context starts from the terrain's
!  erosion state topol(1), topol(14), topol(15) as Input_SW left it (restart
!  included), and the masses and that state come back with the particle
!  state as ExtForces_SW(1) updates them ("6 SPH_SW_2019.f90":8724-8806).

module SPH_MAIN_GFSPH

use, intrinsic :: iso_c_binding
use gfsph_c
use SPH_FEM_variable_types_2019
use SPH_FEM_Driver_time_vars_2019
use SPH_Main_Vars_2019
use SPH_SW_Vars_2019
use SPH_FEM_topo_2019

implicit none
private

type(c_ptr), save :: ctx = c_null_ptr
integer, save, public :: gfsph_integrator = 4     ! 4: RK4 as RK45 (7:746); 2: RK2
integer, save, public :: gfsph_status = 0
logical, save, public :: gfsph_pull_every_step = .false.
integer, save, public :: gfsph_pulls = 0          ! copies back so far (diagnostics)
logical, save :: stale = .false.                  ! module arrays behind the GPU state

public :: time_integration_sph_gfsph, gfsph_release, gfsph_sync_state

contains

subroutine time_integration_sph_gfsph

   integer(c_int) rc
   real(c_double) tnow(1), dtnow(1)

   if (.not. c_associated(ctx)) call create_context
   rc = gfsph_step(ctx, 1_c_int32_t)
   call check(rc, 'gfsph_step')
   if (ic_adapt_dt_sph .ne. 0) then
      rc = gfsph_time(ctx, tnow, dtnow)
      call check(rc, 'gfsph_time')
      dt_sph = dtnow(1)
   endif
   stale = .true.
   if (driver_reads_state()) call gfsph_sync_state

end subroutine time_integration_sph_gfsph


! Whether the driver reads the particle arrays after the step now ending:
! an output counter fires, a case table's stop test runs, or its loop ends.
logical function driver_reads_state()
   driver_reads_state = gfsph_pull_every_step                  &
        .or. t_print_reset .ge. time_print                     &
        .or. t_save_reset  .ge. time_save                      &
        .or. t_plot_reset  .ge. time_plot                      &
        .or. ic_cases_mat  .ge. 1                              &
        .or. time + dt_sph .ge. time_end                       &
        .or. itimestep     .ge. nstart + maxtimestep
end function driver_reads_state


! Copies the GPU state back into the module arrays if they are behind it.
subroutine gfsph_sync_state
   if (stale .and. c_associated(ctx)) then
      call pull_state
      gfsph_pulls = gfsph_pulls + 1
   endif
   stale = .false.
end subroutine gfsph_sync_state


subroutine create_context

   type(gfsph_params_t) :: p
   type(gfsph_dem_t)    :: d
   real(c_double),  allocatable, target :: t9(:,:), zz(:)
   integer(c_int32_t), allocatable, target :: inm(:), ity(:)
   type(gfsph_zone_t), allocatable, target :: zn(:)
   integer(c_int) rc
   integer k, np16

   if (ntotal .ne. npoin) then
      write(*,*) ' gfsph: virtual, absorbing or vn0 particles are not supported'
      stop
   endif

   p%integrator        = gfsph_integrator
   p%pa_sph            = pa_sph
   p%nnps              = nnps
   p%sle               = sle
   p%skf               = skf
   p%summation_density = merge(1, 0, summation_density)
   p%average_velocity  = merge(1, 0, average_velocity)
   p%nor_density       = merge(1, 0, nor_density)
   p%icpwp             = icpwp
   p%npoin0            = npoin0
   p%init_input        = 0          ! Init_sph has already run (6:1428-1452)
   p%check_finite      = 0
   p%dt                = dt_sph
   p%h_inf             = h_inf_SW
   p%pwp0rel           = pwp0rel
   p%cst               = 0.0_c_double
   do k = 1, min(110, size(const))
      p%cst(k) = real(const(k), c_double)      ! REAL(4) values, widened
   enddo
   p%ic_k0     = ic_K0                          ! read for frictional laws only (6:1373-1390)
   p%k0_active = K0_active
   p%k0_pasive = K0_pasive
   if (icpwp .eq. 2) then                       ! the column Initialize_Pwp set up (6:1743-1768)
      p%nuaux          = nUaux
      p%pwp_bc_type(:) = BC_pwp_Type(:)
      p%pwp_bc_val(:)  = BC_pwp_Val(:)
   endif

   allocate (t9(9, npoig), zz(npoig), inm(npoig), ity(npoin))
   t9(1:9, :) = topol(1:9, 1:npoig)
   zz(:)      = topol(1, 1:npoig)
   inm(:)     = poig_in_model(1:npoig)
   ity(:)     = itype(1:npoin)
   d%nx = npoigx;  d%ny = npoigy
   d%xmin = xming; d%ymin = yming; d%dx = deltxg; d%dy = deltyg
   d%xmax = xmaxg; d%ymax = ymaxg
   d%Z        = c_loc(zz)
   d%in_model = c_loc(inm)
   d%topol    = c_loc(t9)
   if (ic_topo_props .eq. 10 .and. ntopo_zones .gt. 0) then   ! terrain zones (2:516-576)
      allocate (zn(ntopo_zones))
      np16 = min(16, ntopo_props)
      do k = 1, ntopo_zones
         zn(k)%type = basal_type(k)
         zn(k)%law  = basal_law(k)
         zn(k)%props(1:np16) = real(basal_props(1:np16, k), c_double)
         if (int(basal_props(1, k) + 0.1) .eq. 101) then
            write(*,*) ' gfsph: terrain zones of geometry 101 (node lists of the terrain module) are not passed'
            stop
         endif
      enddo
      d%nzones = int(ntopo_zones, c_int32_t)
      d%zones  = c_loc(zn)
   endif

   rc = gfsph_create(p, d, int(npoin, c_int32_t), x, vx, rho, mass, hsml, u, ity, ctx)
   call check(rc, 'gfsph_create')
   if (ic_adapt_dt_sph .ne. 0) then
      rc = gfsph_set_adaptive_dt(ctx, int(ic_adapt_dt_sph, c_int32_t))
      call check(rc, 'gfsph_set_adaptive_dt')
   endif
   if (ic_erosion .eq. 1) call push_erosion
   if (int(const(5) + 0.01) .eq. 3) then       ! Bingham erosion reads mass0 (6:9451)
      rc = gfsph_set_mass0(ctx, mass0)
      call check(rc, 'gfsph_set_mass0')
   endif
   if (icpwp .eq. 2) then                       ! the columns Init_sph produced
      rc = gfsph_set_pwp_column(ctx, Uaux)
      call check(rc, 'gfsph_set_pwp_column')
   endif
   deallocate (t9, zz, inm, ity)
   if (allocated(zn)) deallocate (zn)

end subroutine create_context


subroutine pull_state

   real(c_double), allocatable :: ua(:)
   integer(c_int32_t), allocatable :: fl(:)
   integer(c_int) rc

   allocate (ua(npoin), fl(npoin))
   rc = gfsph_get_state(ctx, x, vx, rho, u, ua, hsml, fl)
   call check(rc, 'gfsph_get_state')
   if (icpwp .eq. 2) then
      rc = gfsph_get_pwp_column(ctx, Uaux)
      call check(rc, 'gfsph_get_pwp_column')
   elseif (allocated(Uaux)) then
      Uaux(1, 1:npoin) = ua(:)
   endif
   If_Out_Domain(1:npoin) = fl(:)
   deallocate (ua, fl)
   if (ic_erosion .eq. 1) call pull_erosion

end subroutine pull_state


subroutine push_erosion

   real(c_double), allocatable :: z(:), e(:), a(:)
   integer(c_int) rc

   allocate (z(npoig), e(npoig), a(npoig))
   z(:) = topol(1, 1:npoig)
   e(:) = topol(14, 1:npoig)
   a(:) = topol(15, 1:npoig)
   rc = gfsph_erosion_set(ctx, 0_c_int32_t, z, e, a)
   call check(rc, 'gfsph_erosion_set')
   deallocate (z, e, a)

end subroutine push_erosion


subroutine pull_erosion

   real(c_double), allocatable :: m(:), z(:), e(:), a(:)
   integer(c_int) rc

   allocate (m(npoin), z(npoig), e(npoig), a(npoig))
   rc = gfsph_get_mass(ctx, m)
   call check(rc, 'gfsph_get_mass')
   mass(1:npoin) = m(:)
   rc = gfsph_erosion_read(ctx, 0_c_int32_t, z, e, a)
   call check(rc, 'gfsph_erosion_read')
   topol(1, 1:npoig)  = z(:)
   topol(14, 1:npoig) = e(:)
   topol(15, 1:npoig) = a(:)
   deallocate (m, z, e, a)

end subroutine pull_erosion


subroutine gfsph_release
   call gfsph_sync_state
   if (c_associated(ctx)) call gfsph_destroy(ctx)
   ctx = c_null_ptr
end subroutine gfsph_release


subroutine check(rc, what)
   integer(c_int), intent(in) :: rc
   character(*), intent(in)   :: what
   gfsph_status = rc
   if (rc .ne. GFSPH_OK) then
      write(*,*) ' gfsph: ', what, ' failed with status ', rc, ': ', gfsph_last_error_f(ctx)
      stop 2
   endif
end subroutine check

end module SPH_MAIN_GFSPH
