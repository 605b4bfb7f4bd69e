!  gfsph_c -- ISO_C_BINDING interface to the C ABI of include/gfsph.h.
!
!  Plain interop declarations only (no dependency on the Geoflow-SPH
!  modules): a Fortran driver USEs this module to call the MI355X
!  propagation step.  Two-component particle arrays are passed as the
!  Fortran x(2,n) / vx(2,n) arrays, whose memory layout is the [n][2] layout
!  the C ABI expects.

module gfsph_c
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: GFSPH_OK = 0

  type, bind(C), public :: gfsph_params_t
     integer(c_int32_t) :: integrator        = 4
     integer(c_int32_t) :: pa_sph            = 2
     integer(c_int32_t) :: nnps              = 4
     integer(c_int32_t) :: sle               = 0
     integer(c_int32_t) :: skf               = 1
     integer(c_int32_t) :: summation_density = 1
     integer(c_int32_t) :: average_velocity  = 1
     integer(c_int32_t) :: nor_density       = 0
     integer(c_int32_t) :: icpwp             = 0
     integer(c_int32_t) :: npoin0            = 0
     integer(c_int32_t) :: init_input        = 0
     integer(c_int32_t) :: check_finite      = 0
     real(c_double)     :: dt                = 0.0_c_double
     real(c_double)     :: h_inf             = 0.0_c_double
     real(c_double)     :: pwp0rel           = 0.0_c_double
     real(c_double)     :: cst(0:110)        = 0.0_c_double
     integer(c_int32_t) :: nuaux             = 0          ! icpwp = 2: nodes of the column
     integer(c_int32_t) :: pwp_bc_type(2)    = [2, 1]     ! BC_pwp_Type
     integer(c_int32_t) :: pwp0_type         = 0
     real(c_double)     :: pwp_bc_val(2)     = 0.0_c_double  ! BC_pwp_Val
     real(c_double)     :: xirel             = 0.0_c_double
     integer(c_int32_t) :: ic_k0             = 0          ! K0 active / passive (frictional laws)
     integer(c_int32_t) :: pad_k0            = 0
     real(c_double)     :: k0_active         = 1.0_c_double
     real(c_double)     :: k0_pasive         = 1.0_c_double
  end type gfsph_params_t

  type, bind(C), public :: gfsph_zone_t                ! basal_type, basal_law, basal_props(1:16)
     integer(c_int32_t) :: type = 0, law = 0
     real(c_double)     :: props(16) = 0.0_c_double
  end type gfsph_zone_t

  type, bind(C), public :: gfsph_dem_t
     integer(c_int32_t) :: nx = 0, ny = 0
     real(c_double)     :: xmin = 0, ymin = 0, dx = 0, dy = 0, xmax = 0, ymax = 0
     type(c_ptr)        :: Z        = c_null_ptr
     type(c_ptr)        :: in_model = c_null_ptr
     type(c_ptr)        :: topol    = c_null_ptr
     integer(c_int32_t) :: nzones = 0, pad_z = 0          ! terrain zones (ic_topo_props = 10)
     type(c_ptr)        :: zones      = c_null_ptr       ! gfsph_zone_t(nzones)
     type(c_ptr)        :: zone_nodes = c_null_ptr       ! node-list bits of geometry-101 zones
     type(c_ptr)        :: eros_avail = c_null_ptr       ! topol(15,:) of the terrain
  end type gfsph_dem_t

  public :: gfsph_create, gfsph_step, gfsph_sync, gfsph_get_state, gfsph_destroy, gfsph_last_error_f
  public :: gfsph_raster_idw, gfsph_mc_pixel_accumulate, gfsph_mc_pixel_read, gfsph_counts
  public :: gfsph_set_adaptive_dt, gfsph_time
  public :: gfsph_get_mass, gfsph_set_mass0, gfsph_erosion_read, gfsph_erosion_set
  public :: gfsph_get_pwp_column, gfsph_set_pwp_column

  interface
     integer(c_int) function gfsph_create(p, dem, n, x, v, h, mass, hsml, u, itype, ctx) bind(C, name="gfsph_create")
       import :: c_int, c_int32_t, c_double, c_ptr, gfsph_params_t, gfsph_dem_t
       type(gfsph_params_t), intent(in) :: p
       type(gfsph_dem_t), intent(in)    :: dem
       integer(c_int32_t), value        :: n
       real(c_double), intent(in)       :: x(*), v(*), h(*), mass(*), hsml(*), u(*)
       integer(c_int32_t), intent(in)   :: itype(*)
       type(c_ptr), intent(out)         :: ctx
     end function gfsph_create

     integer(c_int) function gfsph_step(ctx, nsteps) bind(C, name="gfsph_step")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value        :: ctx
       integer(c_int32_t), value :: nsteps
     end function gfsph_step

     integer(c_int) function gfsph_sync(ctx) bind(C, name="gfsph_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function gfsph_sync

     integer(c_int) function gfsph_get_state(ctx, x, v, h, u, uaux, hsml, flag) bind(C, name="gfsph_get_state")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value                :: ctx
       real(c_double), intent(out)       :: x(*), v(*), h(*), u(*), uaux(*), hsml(*)
       integer(c_int32_t), intent(out)   :: flag(*)
     end function gfsph_get_state

     integer(c_int) function gfsph_raster_idw(ctx, out, wsum) bind(C, name="gfsph_raster_idw")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: out(*), wsum(*)
     end function gfsph_raster_idw

     integer(c_int) function gfsph_mc_pixel_accumulate(ctx) bind(C, name="gfsph_mc_pixel_accumulate")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function gfsph_mc_pixel_accumulate

     integer(c_int) function gfsph_mc_pixel_read(ctx, acc, cnt, reset) bind(C, name="gfsph_mc_pixel_read")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value              :: ctx
       real(c_double), intent(out)     :: acc(*)
       integer(c_int32_t), intent(out) :: cnt(*)
       integer(c_int32_t), value       :: reset
     end function gfsph_mc_pixel_read

     integer(c_int) function gfsph_counts(ctx, n, nactive) bind(C, name="gfsph_counts")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value              :: ctx
       integer(c_int32_t), intent(out) :: n, nactive
     end function gfsph_counts

     integer(c_int) function gfsph_set_adaptive_dt(ctx, mode) bind(C, name="gfsph_set_adaptive_dt")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value        :: ctx
       integer(c_int32_t), value :: mode
     end function gfsph_set_adaptive_dt

     integer(c_int) function gfsph_get_pwp_column(ctx, col) bind(C, name="gfsph_get_pwp_column")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: col(*)
     end function gfsph_get_pwp_column

     integer(c_int) function gfsph_set_pwp_column(ctx, col) bind(C, name="gfsph_set_pwp_column")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value         :: ctx
       real(c_double), intent(in) :: col(*)
     end function gfsph_set_pwp_column

     integer(c_int) function gfsph_get_mass(ctx, mass) bind(C, name="gfsph_get_mass")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: mass(*)
     end function gfsph_get_mass

     integer(c_int) function gfsph_set_mass0(ctx, mass0) bind(C, name="gfsph_set_mass0")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value         :: ctx
       real(c_double), intent(in) :: mass0(*)
     end function gfsph_set_mass0

     integer(c_int) function gfsph_erosion_read(ctx, r, z, eroded, avail) bind(C, name="gfsph_erosion_read")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value          :: ctx
       integer(c_int32_t), value   :: r
       real(c_double), intent(out) :: z(*), eroded(*), avail(*)
     end function gfsph_erosion_read

     integer(c_int) function gfsph_erosion_set(ctx, r, z, eroded, avail) bind(C, name="gfsph_erosion_set")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value         :: ctx
       integer(c_int32_t), value  :: r
       real(c_double), intent(in) :: z(*), eroded(*), avail(*)
     end function gfsph_erosion_set

     integer(c_int) function gfsph_time(ctx, time, dt) bind(C, name="gfsph_time")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: time(*), dt(*)
     end function gfsph_time

     type(c_ptr) function gfsph_last_error(ctx) bind(C, name="gfsph_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function gfsph_last_error

     subroutine gfsph_destroy(ctx) bind(C, name="gfsph_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine gfsph_destroy

     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! Message of the last failure on ctx, as a Fortran string.
  function gfsph_last_error_f(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: k, n
    p = gfsph_last_error(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate (character(len=n) :: msg)
    do k = 1, n
       msg(k:k) = chars(k)
    enddo
  end function gfsph_last_error_f

end module gfsph_c
