! ref_horiz_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Driver that runs the reference GITM horizontal step (compiled from the
! reference sources in place, see oracle/Makefile) on blocks read from a
! binary file and writes the updated blocks back: the block arrays are read
! straight into the reference's module globals, the reference subroutine
! advance_horizontal(iBlock) (src/advance_horizontal.f90:58-924) is called
! for every block, and the globals are written out.  The halo cells are the
! caller's (the reference fills them by exchange_messages_sphere before the
! call, src/advance_horizontal.f90:20).
!
! File formats (stream, native little-endian; every array in the reference's
! own declared shape and order, blocks 1..nB):
!   input : int32 hdr(16) = magic, nB, nLons, nLats, nAlts, UseCoriolis,
!           UseIonAdvection, UseImprovedIonAdvection, UseTopography, 0...;
!           real64 par(8) =
!           Dt, OmegaBodyInput, BetaLimiter, 0...; then Rho, Temperature,
!           Velocity(3),
!           VerticalVelocity(6), NDensityS(1:6), Gamma, InvRadialDistance_GB,
!           RadialDistance_GB, InvDLonDist_GB, InvDLonDist_FB, dLonDist_FB,
!           InvDLatDist_GB, InvDLatDist_FB, dLatDist_FB, Latitude,
!           TanLatitude, CosLatitude, cMax_GDB(3), and with ions
!           IDensityS(1:nIons), IVelocity(3); with UseImprovedIonAdvection
!           (hdr(8)) also IVelocityPar(3), GradLonM_CB, GradLon0_CB,
!           GradLonP_CB, GradLatM_CB, GradLat0_CB, GradLatP_CB; with
!           UseTopography (hdr(9)) also dAlt_GB, dAltDLon_CB, dAltDLat_CB
!   output: int32 status(4), then Rho, Temperature, Velocity,
!           VerticalVelocity, NDensityS(1:6), HorizontalTempSource per block
!           (nLons, nLats, nAlts), and with ions IDensityS(1:nIons)
program gitm_horiz_ref

  use ModSizeGitm
  use ModPlanet
  use ModGITM
  use ModInputs
  use ModSources, only: HorizontalTempSource, init_mod_sources

  implicit none

  integer :: hdr(16), status(4)
  real :: par(8)
  integer :: nB, iBlock
  logical :: doIons
  real, allocatable :: tsrc(:,:,:,:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) par
  nB = hdr(2)
  if (hdr(1) /= 1196840531 .or. hdr(3) /= nLons .or. hdr(4) /= nLats .or. &
       hdr(5) /= nAlts .or. nB < 1 .or. nB > nBlocksMax) then
     write(*,*) 'ref_horiz_harness: bad header or grid mismatch', hdr(1:5)
     stop 2
  endif
  doIons = hdr(7) /= 0

  nBlocks = nB
  call init_planet
  call init_mod_gitm
  call init_mod_sources

  UseCoriolis             = (hdr(6) /= 0)
  UseIonAdvection         = doIons
  UseImprovedIonAdvection = doIons .and. hdr(8) /= 0
  UseTopography           = (hdr(9) /= 0)
  DoCheckForNans          = .false.
  Dt                      = par(1)
  OmegaBodyInput          = par(2)
  BetaLimiter             = par(3)

  read(10) Rho(:,:,:,1:nB)
  read(10) Temperature(:,:,:,1:nB)
  read(10) Velocity(:,:,:,:,1:nB)
  read(10) VerticalVelocity(:,:,:,1:nSpecies,1:nB)
  read(10) NDensityS(:,:,:,1:nSpecies,1:nB)
  read(10) Gamma(:,:,:,1:nB)
  read(10) InvRadialDistance_GB(:,:,:,1:nB)
  read(10) RadialDistance_GB(:,:,:,1:nB)
  read(10) InvDLonDist_GB(:,:,:,1:nB)
  read(10) InvDLonDist_FB(:,:,:,1:nB)
  read(10) dLonDist_FB(:,:,:,1:nB)
  read(10) InvDLatDist_GB(:,:,:,1:nB)
  read(10) InvDLatDist_FB(:,:,:,1:nB)
  read(10) dLatDist_FB(:,:,:,1:nB)
  read(10) Latitude(:,1:nB)
  read(10) TanLatitude(:,1:nB)
  read(10) CosLatitude(:,1:nB)
  read(10) cMax_GDB(:,:,:,:,1:nB)
  if (doIons) then
     read(10) IDensityS(:,:,:,1:nIons,1:nB)
     read(10) IVelocity(:,:,:,:,1:nB)
     if (UseImprovedIonAdvection) then
        read(10) IVelocityPar(:,:,:,:,1:nB)
        read(10) GradLonM_CB(:,1:nB)
        read(10) GradLon0_CB(:,1:nB)
        read(10) GradLonP_CB(:,1:nB)
        read(10) GradLatM_CB(:,1:nB)
        read(10) GradLat0_CB(:,1:nB)
        read(10) GradLatP_CB(:,1:nB)
     endif
  else
     IDensityS = 1.0e6
     IVelocity = 0.0
  endif
  if (UseTopography) then
     read(10) dAlt_GB(:,:,:,1:nB)
     read(10) dAltDLon_CB(:,:,:,1:nB)
     read(10) dAltDLat_CB(:,:,:,1:nB)
  endif
  close(10)

  allocate(tsrc(nLons, nLats, nAlts, nB))
  status = 0
  do iBlock = 1, nB
     call advance_horizontal(iBlock)
     tsrc(:,:,:,iBlock) = HorizontalTempSource
  enddo

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) status
  write(11) Rho(:,:,:,1:nB)
  write(11) Temperature(:,:,:,1:nB)
  write(11) Velocity(:,:,:,:,1:nB)
  write(11) VerticalVelocity(:,:,:,1:nSpecies,1:nB)
  write(11) NDensityS(:,:,:,1:nSpecies,1:nB)
  write(11) tsrc
  if (doIons) write(11) IDensityS(:,:,:,1:nIons,1:nB)
  close(11)

end program gitm_horiz_ref
