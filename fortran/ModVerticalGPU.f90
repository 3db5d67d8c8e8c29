! ModVerticalGPU.f90 -- the ISO_C_BINDING layer between GITM's Fortran and
! the HIP library of the vertical step (include/gitm_vert.h).
!
! The bind(C) types mirror the C structs field for field (tests/
! test_fortran_binding.py compares their sizes and offsets with the C
! header's); the interfaces bind the entry points the drop-in calls.  The
! descriptor helpers turn one of GITM's module arrays into a gitm_vert_field:
! the address of its element (iLon=1, iLat=1, first level, component 1,
! block 1) and its strides in elements, taken from the array's own shape, so
! the library reads GITM's (-1:nLons+2,-1:nLats+2,-1:nAlts+2[,n],nBlocks)
! arrays in place.
module ModVerticalGPU

  use iso_c_binding

  implicit none

  integer(c_int), parameter :: GITM_VERT_RUSANOV = 0, GITM_VERT_AUSM = 1
  integer(c_int), parameter :: GITM_VERT_OK = 0, GITM_VERT_EUNSUPPORTED = 5

  type, bind(C) :: gitm_vert_field
     type(c_ptr)        :: ptr = c_null_ptr
     integer(c_int64_t) :: s_lon = 0, s_lat = 0, s_blk = 0, s_lev = 0, s_comp = 0
     integer(c_int32_t) :: first_level = -1, reserved = 0
  end type gitm_vert_field

  type, bind(C) :: gitm_vert_columns
     integer(c_int32_t) :: n_lons = 0, n_lats = 0, n_blocks = 0, n_alts = 0
     type(gitm_vert_field) :: ndensity_s, vertical_velocity, velocity, &
          temperature, rho, log_ns, ndensity, temp_unit, mean_major_mass, &
          gamma, cmax, eddy, altitude, gravity, inv_radial_distance, dalt, &
          cos_latitude, idensity_s, ivelocity, sza, mlat, msis_bc, &
          tides_east, tides_north, tides_temp
  end type gitm_vert_columns

  type, bind(C) :: gitm_vert_params
     real(c_double) :: dt, omega_body, max_vertical_velocity, beta_limiter, &
          min_tec, vert_tau, mass(6), thermal_diff_coef(6)
     integer(c_int32_t) :: use_neutral_friction, use_coriolis, &
          use_ion_advection, use_improved_ion_advection, use_nighttime_ion_bcs, &
          use_msis, use_gswm_tides, use_waccm_tides, use_bc_perturbation, &
          during_perturb, use_damping, reserved
  end type gitm_vert_params

  type, bind(C) :: gitm_vert_error
     integer(c_int32_t) :: code, column, level, reserved
  end type gitm_vert_error

  interface
     subroutine gitm_vert_default_params(p) bind(C, name="gitm_vert_default_params")
       import :: gitm_vert_params
       type(gitm_vert_params), intent(out) :: p
     end subroutine gitm_vert_default_params

     integer(c_int) function gitm_vert_create(h, device) bind(C, name="gitm_vert_create")
       import :: c_ptr, c_int
       type(c_ptr), intent(out) :: h
       integer(c_int), value :: device
     end function gitm_vert_create

     integer(c_int) function gitm_vert_destroy(h) bind(C, name="gitm_vert_destroy")
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function gitm_vert_destroy

     integer(c_int) function gitm_vert_advance_host(h, p, c, solver, n_steps) &
          bind(C, name="gitm_vert_advance_host")
       import :: c_ptr, c_int, gitm_vert_params, gitm_vert_columns
       type(c_ptr), value :: h
       type(gitm_vert_params), intent(in) :: p
       type(gitm_vert_columns), intent(in) :: c
       integer(c_int), value :: solver, n_steps
     end function gitm_vert_advance_host

     integer(c_int) function gitm_vert_get_error(h, e) bind(C, name="gitm_vert_get_error")
       import :: c_ptr, c_int, gitm_vert_error
       type(c_ptr), value :: h
       type(gitm_vert_error), intent(out) :: e
     end function gitm_vert_get_error
  end interface

  ! the handle the drop-in keeps for the run (created on first use)
  type(c_ptr), save :: VertHandle = c_null_ptr

  ! TempUnit and MeanMajorMass are per-block scratch in GITM (calc_rates
  ! fills them for one block at a time, ModGITM.f90:93,170); the drop-in
  ! keeps every block's copy here for the one call over all blocks
  real(c_double), allocatable, target, save :: TempUnitB(:,:,:,:), MeanMajorMassB(:,:,:,:)

  ! UseMSIS: per column, the level-0 values msis_bcs gives (LogNS(1:6),
  ! Temp, east and north winds), filled by the caller
  real(c_double), allocatable, target, save :: MsisBC(:,:,:,:)

contains

  ! Descriptor of a (lon, lat, level[, component], block) array whose index
  ! lower bounds are lo(1:3) for lon, lat and level, and whose extents are
  ! ext(1:nd); the element (1, 1, first level, first component, 1) is at
  ! `first`.  nd = 4 (no component index) or 5.
  function field_of(first, ext, nd, lev_lo) result(f)
    type(c_ptr), intent(in) :: first
    integer, intent(in) :: ext(5), nd, lev_lo
    type(gitm_vert_field) :: f
    f%ptr = first
    f%s_lon = 1
    f%s_lat = ext(1)
    f%s_lev = int(ext(1), c_int64_t) * ext(2)
    f%s_comp = f%s_lev * ext(3)
    if (nd == 5) then
       f%s_blk = f%s_comp * ext(4)
    else
       f%s_blk = f%s_comp
       f%s_comp = 0
    end if
    f%first_level = lev_lo
  end function field_of

  ! Descriptor of a per-column array A(lat, block) or A(lon, lat, block).
  function column_field_of(first, ext_lon, ext_lat) result(f)
    type(c_ptr), intent(in) :: first
    integer, intent(in) :: ext_lon, ext_lat
    type(gitm_vert_field) :: f
    f%ptr = first
    if (ext_lon == 0) then  ! (lat, block)
       f%s_lon = 0
       f%s_lat = 1
       f%s_blk = ext_lat
    else
       f%s_lon = 1
       f%s_lat = ext_lon
       f%s_blk = int(ext_lon, c_int64_t) * ext_lat
    end if
    f%first_level = 0
  end function column_field_of

end module ModVerticalGPU
