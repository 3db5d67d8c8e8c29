! vertical_gpu.f90 -- GITM's vertical step over every column of every block
! in one call of the HIP library (gitm_vert_advance_host), on GITM's own
! module arrays.  advance_vertical_all_gpu is the body of the drop-in
! advance_vertical_all (advance_vertical_all.f90): it does what the column
! loop of src/advance_vertical.f90:20-29 does, advance_vertical(iLon,iLat,
! iBlock) for every column (:45-219: gather, advance_vertical_1d_ausm or
! _rusanov, scatter), with TempUnit and MeanMajorMass taken from the
! per-block copies in ModVerticalGPU.
!
! iError: 0, or the library's status (GITM_VERT_EUNSUPPORTED for an option
! the library does not implement: the caller then runs the Fortran path).
! The first failing column and level are printed, as the reference's
! stop_gitm calls do before stopping.
subroutine advance_vertical_all_gpu(iError)

  use ModSizeGitm, only: nLons, nLats, nAlts
  use ModGITM
  use ModInputs
  use ModPlanet, only: nSpecies, nSpeciesTotal, nSpeciesAll, nIons, Mass, VertTau
  use ModSources, only: KappaEddyDiffusion
  use ModEUV, only: SZA
  use ModTides, only: TidesEast, TidesNorth, TidesTemp
  use ModVerticalGPU

  implicit none

  integer, intent(out) :: iError
  type(gitm_vert_params) :: p
  type(gitm_vert_error)  :: e
  integer(c_int) :: rc
  integer :: iBlock

  if (.not. c_associated(VertHandle)) then
     rc = gitm_vert_create(VertHandle, -1_c_int)
     if (rc /= 0) then
        iError = rc
        return
     end if
  end if

  ! the run parameters the path reads (ModInputs, ModPlanet)
  call gitm_vert_default_params(p)
  p%dt = Dt
  p%omega_body = OmegaBodyInput
  p%max_vertical_velocity = MaximumVerticalVelocity
  p%beta_limiter = BetaLimiter
  p%min_tec = MinTEC
  p%vert_tau = VertTau(1)
  p%mass = Mass(1:nSpecies)
  p%thermal_diff_coef = ThermalDiffCoefS(1:nSpecies)
  p%use_neutral_friction = merge(1, 0, UseNeutralFriction)
  p%use_coriolis = merge(1, 0, UseCoriolis)
  p%use_ion_advection = merge(1, 0, UseIonAdvection)
  p%use_improved_ion_advection = merge(1, 0, UseImprovedIonAdvection)
  p%use_nighttime_ion_bcs = merge(1, 0, UseNighttimeIonBCs)
  p%use_msis = merge(1, 0, UseMSIS)
  p%use_gswm_tides = merge(1, 0, UseGSWMTides)
  p%use_waccm_tides = merge(1, 0, UseWACCMTides)
  p%use_bc_perturbation = merge(1, 0, UseBcPerturbation)
  p%during_perturb = merge(1, 0, DuringPerturb)
  p%use_damping = merge(1, 0, UseDamping)

  ! SpeciesDensityOld before the step (advance_vertical.f90:152-155)
  do iBlock = 1, nBlocks
     SpeciesDensityOld(:,:,:,1:nSpeciesTotal,iBlock) = NDensityS(:,:,:,1:nSpeciesTotal,iBlock)
     SpeciesDensityOld(:,:,:,nSpeciesTotal+1:nSpeciesAll,iBlock) = &
          IDensityS(:,:,:,1:nIons-1,iBlock)
  end do

  call gpu_step(NDensityS, VerticalVelocity, Velocity, Temperature, Rho, LogNS, NDensity, &
       TempUnitB, MeanMajorMassB, Gamma, cMax_GDB, KappaEddyDiffusion, Altitude_GB, &
       Gravity_GB, InvRadialDistance_GB, dAlt_GB, CosLatitude, IDensityS, IVelocity, &
       SZA, MLatitude, TidesEast, TidesNorth, TidesTemp, rc)
  if (rc /= 0) then
     iError = rc
     if (rc /= GITM_VERT_EUNSUPPORTED) then
        rc = gitm_vert_get_error(VertHandle, e)
        write(*,*) "vertical step failed: code", e%code, " column", e%column, &
             " level", e%level
     end if
     return
  end if

  ! SpeciesDensity after the step (advance_vertical.f90:214-217)
  do iBlock = 1, nBlocks
     SpeciesDensity(:,:,:,1:nSpeciesTotal,iBlock) = NDensityS(:,:,:,1:nSpeciesTotal,iBlock)
     SpeciesDensity(:,:,:,nSpeciesTotal+1:nSpeciesAll,iBlock) = &
          IDensityS(:,:,:,1:nIons-1,iBlock)
  end do
  iError = 0

contains

  ! Describes the arrays to the library and runs the step.  The arrays are
  ! assumed-size TARGET dummies, so that c_loc of their elements is valid for
  ! the call; each descriptor points at the array's element (iLon=1, iLat=1,
  ! first level, first component, block 1).
  subroutine gpu_step(nds, vvel, vel, temp, rho_, logns_, nden, tu, mmm, gam, cmax, eddy, &
       alt, grav, invr, dalt, coslat, ions, ivel, sza_, mlat, teast, tnorth, ttemp, rc)
    real(c_double), target, intent(inout) :: nds(*), vvel(*), vel(*), temp(*), rho_(*), &
         logns_(*), nden(*), ions(*)
    real(c_double), target, intent(in) :: tu(*), mmm(*), gam(*), cmax(*), eddy(*), alt(*), &
         grav(*), invr(*), dalt(*), coslat(*), ivel(*), sza_(*), mlat(*), teast(*), &
         tnorth(*), ttemp(*)
    integer(c_int), intent(out) :: rc
    type(gitm_vert_columns) :: c
    integer :: nx, ny, nz, i0, ext(5)

    nx = nLons + 4
    ny = nLats + 4
    nz = nAlts + 4
    ! (-1:nLons+2,-1:nLats+2,-1:nAlts+2[,n],nBlocks): element (1,1,-1)
    i0 = 1 + 2 + 2 * nx
    c%n_lons = nLons
    c%n_lats = nLats
    c%n_blocks = nBlocks
    c%n_alts = nAlts
    ext = (/ nx, ny, nz, nSpeciesTotal, 0 /)
    c%ndensity_s = field_of(c_loc(nds(i0)), ext, 5, -1)
    ext(4) = nSpecies
    c%vertical_velocity = field_of(c_loc(vvel(i0)), ext, 5, -1)
    c%log_ns = field_of(c_loc(logns_(i0)), ext, 5, -1)
    ext(4) = 3
    c%velocity = field_of(c_loc(vel(i0)), ext, 5, -1)
    c%ivelocity = field_of(c_loc(ivel(i0)), ext, 5, -1)
    ext(4) = nIons
    c%idensity_s = field_of(c_loc(ions(i0)), ext, 5, -1)
    c%temperature = field_of(c_loc(temp(i0)), ext, 4, -1)
    c%rho = field_of(c_loc(rho_(i0)), ext, 4, -1)
    c%ndensity = field_of(c_loc(nden(i0)), ext, 4, -1)
    c%temp_unit = field_of(c_loc(tu(i0)), ext, 4, -1)
    c%mean_major_mass = field_of(c_loc(mmm(i0)), ext, 4, -1)
    c%gamma = field_of(c_loc(gam(i0)), ext, 4, -1)
    c%altitude = field_of(c_loc(alt(i0)), ext, 4, -1)
    c%gravity = field_of(c_loc(grav(i0)), ext, 4, -1)
    c%inv_radial_distance = field_of(c_loc(invr(i0)), ext, 4, -1)
    c%dalt = field_of(c_loc(dalt(i0)), ext, 4, -1)
    ! cMax_GDB(0:nLons+1,0:nLats+1,0:nAlts+1,3,nBlocks): element (1,1,0,iUp_,1)
    ext = (/ nLons + 2, nLats + 2, nAlts + 2, 3, 0 /)
    c%cmax = field_of(c_loc(cmax(1 + 1 + (nLons + 2) + &
         (iUp_ - 1) * (nLons + 2) * (nLats + 2) * (nAlts + 2))), ext, 5, 0)
    c%cmax%s_comp = 0
    ! KappaEddyDiffusion(nLons,nLats,-1:nAlts+2,nBlocks): its first element
    ext = (/ nLons, nLats, nz, 0, 0 /)
    c%eddy = field_of(c_loc(eddy(1)), ext, 4, -1)
    ! CosLatitude(-1:nLats+2,nBlocksMax): element (1,1)
    c%cos_latitude = column_field_of(c_loc(coslat(3)), 0, ny)
    ! SZA(nLons,nLats,nBlocks); MLatitude(iLon,iLat,nAlts,iBlock)
    c%sza = column_field_of(c_loc(sza_(1)), nLons, nLats)
    c%mlat = column_field_of(c_loc(mlat(i0 + (nAlts + 1) * nx * ny)), nx, ny)
    c%mlat%s_blk = int(nx, c_int64_t) * ny * nz
    ! UseMSIS: MsisBC(9,nLons,nLats,nBlocks), components fastest
    if (allocated(MsisBC)) then
       c%msis_bc%ptr = c_loc(MsisBC(1,1,1,1))
       c%msis_bc%s_comp = 1
       c%msis_bc%s_lon = 9
       c%msis_bc%s_lat = 9 * nLons
       c%msis_bc%s_blk = 9 * nLons * nLats
       c%msis_bc%first_level = 0
    end if
    ! Tides: TidesEast/North/Temp(-1:nLons+2,-1:nLats+2,2,nBlocksMax), levels
    ! -1 and 0 (ModTides.f90:29-31): element (1,1,level -1,block 1)
    if (UseGSWMTides .or. UseWACCMTides) then
       ext = (/ nx, ny, 2, 0, 0 /)
       c%tides_east = field_of(c_loc(teast(i0)), ext, 4, -1)
       c%tides_north = field_of(c_loc(tnorth(i0)), ext, 4, -1)
       c%tides_temp = field_of(c_loc(ttemp(i0)), ext, 4, -1)
    end if
    rc = gitm_vert_advance_host(VertHandle, p, c, &
         merge(GITM_VERT_AUSM, GITM_VERT_RUSANOV, UseAUSMSolver), 1_c_int)
  end subroutine gpu_step

end subroutine advance_vertical_all_gpu
