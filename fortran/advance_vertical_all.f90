! advance_vertical_all.f90 -- the drop-in for GITM's advance_vertical_all
! (src/advance_vertical.f90:8-39, called from src/advance.f90:41): same name,
! same empty argument list.  The block loop's rate and viscosity updates
! stay in Fortran, as in the reference (:20-24); the column loop runs on the
! GPU in one call (vertical_gpu.f90).  When the library does not implement a
! selected option (GITM_VERT_EUNSUPPORTED: BC perturbation, UseDamping
! with the AUSM solver) the original routine runs, renamed advance_vertical_all_fortran
! by the maintainer; any other failure stops the run, as the reference's
! stop_gitm calls on this path do.
subroutine advance_vertical_all

  use ModSizeGitm, only: nLons, nLats, nAlts
  use ModGITM, only: nBlocks, TempUnit, MeanMajorMass
  use ModInputs, only: UseMSIS
  use ModVerticalGPU

  implicit none

  integer :: iBlock, iError

  if (.not. allocated(TempUnitB)) then
     allocate(TempUnitB(-1:nLons+2,-1:nLats+2,-1:nAlts+2,nBlocks))
     allocate(MeanMajorMassB(-1:nLons+2,-1:nLats+2,-1:nAlts+2,nBlocks))
  endif

  call report("advance_vertical_all", 2)
  call start_timing("vertical_all")

  do iBlock = 1, nBlocks
     call calc_rates(iBlock)
     call calc_viscosity(iBlock)
     TempUnitB(:,:,:,iBlock) = TempUnit
     MeanMajorMassB(:,:,:,iBlock) = MeanMajorMass
  enddo

  ! UseMSIS: the caller's level-0 MSIS values, one column at a time
  ! (msis_at_level0.f90: the reference's msis_bcs, called as
  ! set_vertical_bcs.Earth.f90:116-131 calls it for iAlt = 0)
  if (UseMSIS) call fill_msis_bc

  call advance_vertical_all_gpu(iError)
  if (iError == GITM_VERT_EUNSUPPORTED) then
     call advance_vertical_all_fortran
  else if (iError /= 0) then
     call stop_gitm("Can't Continue")
  endif

  call end_timing("vertical_all")

end subroutine advance_vertical_all
