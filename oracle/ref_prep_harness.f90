! ref_prep_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Driver that runs the reference GITM column prep and vertical time step
! (src/calc_pressure.f90 calc_pressure, src/calc_rates.Earth.f90
! calc_rates, src/Earth.f90 calc_eddy_diffusion_coefficient and
! src/calc_timestep.f90 calc_timestep_vertical, compiled from the reference
! sources in place, see oracle/Makefile) on columns read from a binary file:
! the columns fill the interior cells (iLon, iLat, iBlock) of whole 9x9
! blocks; calc_pressure runs over all blocks, then calc_rates and
! calc_eddy_diffusion_coefficient block by block (their outputs are per-block
! scratch, read back after each call), then calc_timestep_vertical with Dt
! starting at FixedDt, as main.f90 sets it before the time-step routines.
!
! File formats (stream, native little-endian):
!   input : int32 hdr(16) = (magic 1196840534, nCols (a multiple of 81),
!           nAlts, UseIonAdvection, ...), real64 par(8) = (Cfl, FixedDt,
!           DtEnd, EddyDiffusionCoef, EddyDiffusionPressure0,
!           EddyDiffusionPressure1, ...), then [col][component][level]
!           arrays of nAlts+4 levels: nds(6) temp (normalised) rho vvel(6)
!           dalt ivel(3)
!   output: int32 status(4), then [col][level] arrays on levels
!           -1..nAlts+2: NDensity Pressure cp Gamma MeanMajorMass TempUnit
!           KappaEddyDiffusion cMax(up) (cMax on levels 0..nAlts+1, the
!           others zero), then real64 Dt, Mass(1:6), Vibration(1:6)
program gitm_prep_ref

  use ModSizeGitm
  use ModPlanet
  use ModGITM
  use ModInputs
  use ModTime, only: CurrentTime, EndTime
  use ModSources, only: KappaEddyDiffusion, init_mod_sources

  implicit none

  integer :: hdr(16), status(4)
  real :: par(8)
  integer :: nCols, nA, L, iCol, iLon, iLat, iB, s, d, nB
  real, allocatable :: nds(:,:,:), temp(:,:), rhoH(:,:), vvel(:,:,:), dalt(:,:), ivel(:,:,:)
  real, allocatable :: nden_o(:,:), pres_o(:,:), cp_o(:,:), gam_o(:,:), mmm_o(:,:), &
       tu_o(:,:), eddy_o(:,:), cmax_o(:,:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) par
  nCols = hdr(2)
  nA = hdr(3)
  if (hdr(1) /= 1196840534 .or. nA /= nAlts .or. mod(nCols, nLons*nLats) /= 0 .or. &
       nCols / (nLons*nLats) > nBlocksMax) then
     write(*,*) 'ref_prep_harness: bad header, nAlts or column count', hdr(1), nA, nCols
     stop 2
  endif
  L = nA + 4
  allocate(nds(L,nSpecies,nCols), temp(L,nCols), rhoH(L,nCols), vvel(L,nSpecies,nCols))
  allocate(dalt(L,nCols), ivel(L,3,nCols))
  read(10) nds
  read(10) temp
  read(10) rhoH
  read(10) vvel
  read(10) dalt
  read(10) ivel
  close(10)

  nB = nCols / (nLons*nLats)
  nBlocks = nB
  call init_planet
  call init_mod_gitm
  call init_mod_sources

  Cfl = par(1)
  Dt = par(2)
  CurrentTime = 0.0
  EndTime = par(3)
  EddyDiffusionCoef = par(4)
  EddyDiffusionPressure0 = par(5)
  EddyDiffusionPressure1 = par(6)
  UseIonAdvection = (hdr(4) /= 0)
  UseBarriers = .false.

  ! every cell holds a valid state: halo cells too
  NDensityS = 1.0
  Temperature = 1.0
  Rho = 1.0
  VerticalVelocity = 0.0
  dAlt_GB = 1.0
  IVelocity = 0.0
  IDensityS = 1.0
  ITemperature = 1.0
  eTemperature = 1.0

  do iCol = 1, nCols
     iLon = mod(iCol-1, nLons) + 1
     iLat = mod((iCol-1)/nLons, nLats) + 1
     iB = (iCol-1)/(nLons*nLats) + 1
     do s = 1, nSpecies
        NDensityS(iLon,iLat,-1:nA+2,s,iB) = nds(:,s,iCol)
        VerticalVelocity(iLon,iLat,-1:nA+2,s,iB) = vvel(:,s,iCol)
     enddo
     Temperature(iLon,iLat,-1:nA+2,iB) = temp(:,iCol)
     Rho(iLon,iLat,-1:nA+2,iB) = rhoH(:,iCol)
     dAlt_GB(iLon,iLat,-1:nA+2,iB) = dalt(:,iCol)
     do d = 1, 3
        IVelocity(iLon,iLat,-1:nA+2,d,iB) = ivel(:,d,iCol)
     enddo
  enddo

  allocate(nden_o(L,nCols), pres_o(L,nCols), cp_o(L,nCols), gam_o(L,nCols), &
       mmm_o(L,nCols), tu_o(L,nCols), eddy_o(L,nCols), cmax_o(L,nCols))
  cmax_o = 0.0

  call calc_pressure
  do iB = 1, nB
     call calc_rates(iB)
     call calc_eddy_diffusion_coefficient(iB)
     do iCol = (iB-1)*nLons*nLats + 1, iB*nLons*nLats
        iLon = mod(iCol-1, nLons) + 1
        iLat = mod((iCol-1)/nLons, nLats) + 1
        mmm_o(:,iCol) = MeanMajorMass(iLon,iLat,-1:nA+2)
        tu_o(:,iCol) = TempUnit(iLon,iLat,-1:nA+2)
        eddy_o(:,iCol) = KappaEddyDiffusion(iLon,iLat,-1:nA+2,iB)
     enddo
  enddo
  call calc_timestep_vertical

  do iCol = 1, nCols
     iLon = mod(iCol-1, nLons) + 1
     iLat = mod((iCol-1)/nLons, nLats) + 1
     iB = (iCol-1)/(nLons*nLats) + 1
     nden_o(:,iCol) = NDensity(iLon,iLat,-1:nA+2,iB)
     pres_o(:,iCol) = Pressure(iLon,iLat,-1:nA+2,iB)
     cp_o(:,iCol) = cp(iLon,iLat,-1:nA+2,iB)
     gam_o(:,iCol) = Gamma(iLon,iLat,-1:nA+2,iB)
     cmax_o(2:L-1,iCol) = cMax_GDB(iLon,iLat,0:nA+1,iUp_,iB)
  enddo

  status = 0
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) status
  write(11) nden_o
  write(11) pres_o
  write(11) cp_o
  write(11) gam_o
  write(11) mmm_o
  write(11) tu_o
  write(11) eddy_o
  write(11) cmax_o
  write(11) Dt
  write(11) Mass(1:nSpecies)
  write(11) Vibration(1:nSpecies)
  close(11)

end program gitm_prep_ref
