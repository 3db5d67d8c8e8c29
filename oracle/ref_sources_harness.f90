! ref_sources_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Driver that runs the reference GITM thermal conduction and viscosity
! (src/calc_conduction.f90 calc_thermal_conduction and calc_conduction_1d,
! src/calc_viscosity.f90 calc_viscosity with src/calc_rates.Earth.f90
! calc_viscosity_coef, compiled from the reference sources in place, see
! oracle/Makefile) on columns read from a binary file, block by block: the
! columns fill the interior cells (iLon, iLat, iBlock) of whole 9x9 blocks,
! each block's per-block scratch (TempUnit, MeanMajorMass) is set, then
! calc_thermal_conduction(iBlock) and calc_viscosity(iBlock) run, and their
! outputs are read back from the module arrays.
!
! File formats (stream, native little-endian):
!   input : int32 hdr(16) = (magic 1196840533, nCols (a multiple of 81),
!           nAlts, UseTurbulentCond, ...), real64 par(8) = (Dt, ...), then
!           [col][component][level] arrays of nAlts+4 levels: temp
!           (normalised) tempunit rho cp eddy nds(6) kappa_temp vel(3)
!           vvel(6) alt raddist mmm
!   output: int32 status(4), then on levels -1..nAlts+2 (levels -1 and
!           nAlts+2 zero): conduction viscosity(3) vert_viscosity(6)
!           visc_coef visc_coef_s(6)
program gitm_sources_ref

  use ModSizeGitm
  use ModPlanet
  use ModGITM
  use ModInputs
  use ModSources, only: KappaEddyDiffusion, Conduction, init_mod_sources

  implicit none

  integer :: hdr(16), status(4)
  real :: par(8)
  integer :: nCols, nA, L, iCol, iLon, iLat, iB, s, d, nB
  real, allocatable :: temp(:,:), tu(:,:), rhoH(:,:), cpH(:,:), eddy(:,:)
  real, allocatable :: nds(:,:,:), kt(:,:), vel(:,:,:), vvel(:,:,:), alt(:,:), rd(:,:)
  real, allocatable :: mmm(:,:)
  real, allocatable :: cond_o(:,:), visc_o(:,:,:), vvis_o(:,:,:), vc_o(:,:), vcs_o(:,:,:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) par
  nCols = hdr(2)
  nA = hdr(3)
  if (hdr(1) /= 1196840533 .or. nA /= nAlts .or. mod(nCols, nLons*nLats) /= 0 .or. &
       nCols / (nLons*nLats) > nBlocksMax) then
     write(*,*) 'ref_sources_harness: bad header, nAlts or column count', hdr(1), nA, nCols
     stop 2
  endif
  L = nA + 4
  allocate(temp(L,nCols), tu(L,nCols), rhoH(L,nCols), cpH(L,nCols), eddy(L,nCols))
  allocate(nds(L,nSpecies,nCols), kt(L,nCols), vel(L,3,nCols), vvel(L,nSpecies,nCols))
  allocate(alt(L,nCols), rd(L,nCols), mmm(L,nCols))
  read(10) temp
  read(10) tu
  read(10) rhoH
  read(10) cpH
  read(10) eddy
  read(10) nds
  read(10) kt
  read(10) vel
  read(10) vvel
  read(10) alt
  read(10) rd
  read(10) mmm
  close(10)

  nB = nCols / (nLons*nLats)
  nBlocks = nB
  call init_planet
  call init_mod_gitm
  call init_mod_sources

  Dt = par(1)
  UseTurbulentCond = (hdr(4) /= 0)
  TestViscosityFactor = 1.0

  ! every cell holds a valid state: halo cells too
  Temperature = 1.0
  Rho = 1.0
  cp = 1.0
  KappaEddyDiffusion = 0.0
  NDensityS = 1.0
  KappaTemp = 0.0
  Velocity = 0.0
  VerticalVelocity = 0.0
  Altitude_GB = 0.0
  RadialDistance_GB = 1.0

  do iCol = 1, nCols
     iLon = mod(iCol-1, nLons) + 1
     iLat = mod((iCol-1)/nLons, nLats) + 1
     iB = (iCol-1)/(nLons*nLats) + 1
     Temperature(iLon,iLat,-1:nA+2,iB) = temp(:,iCol)
     Rho(iLon,iLat,-1:nA+2,iB) = rhoH(:,iCol)
     cp(iLon,iLat,0:nA+1,iB) = cpH(2:L-1,iCol)
     KappaEddyDiffusion(iLon,iLat,-1:nA+2,iB) = eddy(:,iCol)
     do s = 1, nSpecies
        NDensityS(iLon,iLat,-1:nA+2,s,iB) = nds(:,s,iCol)
        VerticalVelocity(iLon,iLat,-1:nA+2,s,iB) = vvel(:,s,iCol)
     enddo
     KappaTemp(iLon,iLat,0:nA+1,iB) = kt(2:L-1,iCol)
     do d = 1, 3
        Velocity(iLon,iLat,-1:nA+2,d,iB) = vel(:,d,iCol)
     enddo
     Altitude_GB(iLon,iLat,-1:nA+2,iB) = alt(:,iCol)
     RadialDistance_GB(iLon,iLat,-1:nA+2,iB) = rd(:,iCol)
  enddo

  allocate(cond_o(L,nCols), visc_o(L,3,nCols), vvis_o(L,nSpecies,nCols), vc_o(L,nCols), &
       vcs_o(L,nSpecies,nCols))
  cond_o = 0.0
  visc_o = 0.0
  vvis_o = 0.0
  vc_o = 0.0
  vcs_o = 0.0

  do iB = 1, nB
     ! the per-block scratch calc_rates fills for this block
     TempUnit = 1.0
     MeanMajorMass = 1.0
     do iCol = (iB-1)*nLons*nLats + 1, iB*nLons*nLats
        iLon = mod(iCol-1, nLons) + 1
        iLat = mod((iCol-1)/nLons, nLats) + 1
        TempUnit(iLon,iLat,-1:nA+2) = tu(:,iCol)
        MeanMajorMass(iLon,iLat,-1:nA+2) = mmm(:,iCol)
     enddo
     call calc_thermal_conduction(iB)
     call calc_viscosity(iB)
     do iCol = (iB-1)*nLons*nLats + 1, iB*nLons*nLats
        iLon = mod(iCol-1, nLons) + 1
        iLat = mod((iCol-1)/nLons, nLats) + 1
        cond_o(2:L-1,iCol) = Conduction(iLon,iLat,0:nA+1)
        do d = 1, 3
           visc_o(2:L-1,d,iCol) = Viscosity(iLon,iLat,0:nA+1,d)
        enddo
        do s = 1, nSpecies
           vvis_o(2:L-1,s,iCol) = VerticalViscosityS(iLon,iLat,0:nA+1,s)
           vcs_o(2:L-1,s,iCol) = ViscCoefS(iLon,iLat,0:nA+1,s)
        enddo
        vc_o(2:L-1,iCol) = ViscCoef(iLon,iLat,0:nA+1)
     enddo
  enddo

  status = 0
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) status
  write(11) cond_o
  write(11) visc_o
  write(11) vvis_o
  write(11) vc_o
  write(11) vcs_o
  close(11)

end program gitm_sources_ref
