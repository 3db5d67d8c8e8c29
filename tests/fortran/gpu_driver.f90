! gpu_driver.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Runs the Fortran drop-in (fortran/vertical_gpu.f90: advance_vertical_all_gpu
! over ISO_C_BINDING into the HIP library) on GITM's own module arrays, built
! against the reference's module objects (oracle/Makefile, target ref).
! Input and output files have the format of oracle/ref_harness.f90; column c
! (0-based) of the file is grid cell (iLon, iLat, iBlock) = (mod(c,nLons)+1,
! mod(c/nLons,nLats)+1, c/(nLons*nLats)+1) of the reference's 9x9-cell
! blocks, so up to nBlocksMax = 4 blocks (324 columns) run in one call.
program gitm_vert_gpu_driver

  use ModSizeGitm
  use ModPlanet
  use ModGITM
  use ModInputs
  use ModSources, only: KappaEddyDiffusion, EuvHeating, init_mod_sources
  use ModEUV, only: SZA
  use ModTides, only: TidesEast, TidesNorth, TidesTemp
  use ModTime, only: UTime, iJulianDay, CurrentTime
  use EUA_ModMsis00, only: meters, tselec
  use ModVerticalGPU

  implicit none

  integer :: hdr(16), status(4)
  real :: par(8)
  integer :: nCols, nA, solver, nStepsH, iCol, iStep, L, s, d, iLon, iLat, iB, iError
  logical :: doIons, doTides, doMsis
  real, allocatable :: latr(:), lonr(:), msis_o(:,:)
  real, allocatable :: nds(:,:,:), vvel(:,:,:), vel(:,:,:), temp(:,:), rhoH(:,:)
  real, allocatable :: tempunit_(:,:), mmm(:,:), gam(:,:), cmax(:,:), eddy(:,:)
  real, allocatable :: alt(:,:), grav(:,:), invr(:,:), dalt(:,:), coslat(:)
  real, allocatable :: ions(:,:,:), ivel(:,:,:), sza_(:), mlat(:)
  real, allocatable :: logns_o(:,:,:), nden_o(:,:), tides(:,:,:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) par
  nCols = hdr(2)
  nA = hdr(3)
  solver = hdr(4)
  nStepsH = hdr(5)
  if (hdr(1) /= 1196840530 .or. nA /= nAlts .or. mod(nCols, nLons*nLats) /= 0 .or. &
       nCols > nLons*nLats*nBlocksMax) then
     write(*,*) 'gpu_driver: bad header, nAlts or column count', hdr(1), nA, nCols
     stop 2
  endif
  L = nA + 4
  doIons = hdr(8) /= 0

  allocate(nds(L,nSpecies,nCols), vvel(L,nSpecies,nCols), vel(L,3,nCols))
  allocate(temp(L,nCols), rhoH(L,nCols), tempunit_(L,nCols), mmm(L,nCols))
  allocate(gam(L,nCols), cmax(L,nCols), eddy(L,nCols), alt(L,nCols))
  allocate(grav(L,nCols), invr(L,nCols), dalt(L,nCols), coslat(nCols))
  allocate(logns_o(L,nSpecies,nCols), nden_o(L,nCols))
  read(10) nds
  read(10) vvel
  read(10) vel
  read(10) temp
  read(10) rhoH
  read(10) tempunit_
  read(10) mmm
  read(10) gam
  read(10) cmax
  read(10) eddy
  read(10) alt
  read(10) grav
  read(10) invr
  read(10) dalt
  read(10) coslat
  if (doIons) then
     allocate(ions(L,nIons,nCols), ivel(L,3,nCols), sza_(nCols), mlat(nCols))
     read(10) ions
     read(10) ivel
     read(10) sza_
     read(10) mlat
  endif
  doTides = hdr(11) /= 0 .or. hdr(12) /= 0
  if (doTides) then
     allocate(tides(2,3,nCols))
     read(10) tides
  endif
  doMsis = hdr(14) /= 0
  if (doMsis) then
     allocate(latr(nCols), lonr(nCols), msis_o(9,nCols))
     read(10) latr
     read(10) lonr
  endif
  close(10)

  nBlocks = nCols / (nLons*nLats)
  call init_planet
  call init_mod_gitm
  call init_mod_sources
  allocate(TempUnitB(-1:nLons+2,-1:nLats+2,-1:nAlts+2,nBlocks))
  allocate(MeanMajorMassB(-1:nLons+2,-1:nLats+2,-1:nAlts+2,nBlocks))

  UseAUSMSolver              = (solver == 1)
  UseNeutralFriction         = (hdr(6) /= 0)
  UseCoriolis                = (hdr(7) /= 0)
  UseIonAdvection            = doIons
  UseImprovedIonAdvection    = (hdr(9) /= 0)
  UseNighttimeIonBCs         = (hdr(10) /= 0)
  UseMSIS                    = doMsis
  UseGSWMTides               = (hdr(11) /= 0)
  UseWACCMTides              = (hdr(12) /= 0)
  UseBcPerturbation          = .false.
  UseDamping                 = (hdr(13) /= 0)
  Dt                         = par(1)
  OmegaBodyInput             = par(2)
  MaximumVerticalVelocity    = par(3)
  BetaLimiter                = par(4)
  MinTEC                     = par(5)
  if (doMsis) then
     ! time, solar inputs and MSIS-00 set-up as oracle/ref_harness.f90 does
     iJulianDay = hdr(15)
     UTime = par(6)
     CurrentTime = par(6)
     f107 = par(7)
     f107a = par(8)
     UseMSISTides = .false.
     call meters(.true.)
     sw_msis = 0
     sw_msis(1) = 1
     call tselec(sw_msis)
  endif
  ThermalDiffCoefS(1:nSpecies) = 0.0
  ThermalDiffCoefS(iHe_) = -0.38

  ! every cell holds a valid state: halo cells too (the library reads only
  ! the interior columns, but the halos must not hold garbage)
  NDensityS = 1.0
  IDensityS = 1.0e6
  VerticalVelocity = 0.0
  Velocity = 0.0
  Temperature = 1.0
  Rho = 1.0
  IVelocity = 0.0
  MLatitude = 0.0
  SZA = 0.0
  CosLatitude = 1.0
  cMax_GDB = 0.0
  KappaEddyDiffusion = 0.0
  TempUnitB = 1.0
  MeanMajorMassB = 1.0
  Gamma = 1.5
  Altitude_GB = 0.0
  Gravity_GB = 0.0
  InvRadialDistance_GB = 1.0
  dAlt_GB = 1.0

  do iCol = 1, nCols
     iLon = mod(iCol-1, nLons) + 1
     iLat = mod((iCol-1)/nLons, nLats) + 1
     iB = (iCol-1)/(nLons*nLats) + 1
     do s = 1, nSpecies
        NDensityS(iLon,iLat,-1:nA+2,s,iB) = nds(:,s,iCol)
        VerticalVelocity(iLon,iLat,-1:nA+2,s,iB) = vvel(:,s,iCol)
     enddo
     do d = 1, 3
        Velocity(iLon,iLat,-1:nA+2,d,iB) = vel(:,d,iCol)
     enddo
     Temperature(iLon,iLat,-1:nA+2,iB) = temp(:,iCol)
     Rho(iLon,iLat,-1:nA+2,iB) = rhoH(:,iCol)
     TempUnitB(iLon,iLat,-1:nA+2,iB) = tempunit_(:,iCol)
     MeanMajorMassB(iLon,iLat,-1:nA+2,iB) = mmm(:,iCol)
     Gamma(iLon,iLat,-1:nA+2,iB) = gam(:,iCol)
     cMax_GDB(iLon,iLat,0:nA+1,iUp_,iB) = cmax(2:L-1,iCol)
     KappaEddyDiffusion(iLon,iLat,-1:nA+2,iB) = eddy(:,iCol)
     Altitude_GB(iLon,iLat,-1:nA+2,iB) = alt(:,iCol)
     Gravity_GB(iLon,iLat,-1:nA+2,iB) = grav(:,iCol)
     InvRadialDistance_GB(iLon,iLat,-1:nA+2,iB) = invr(:,iCol)
     dAlt_GB(iLon,iLat,-1:nA+2,iB) = dalt(:,iCol)
     ! CosLatitude is per (iLat, iBlock): the file's columns of one latitude
     ! row of a block share it
     CosLatitude(iLat,iB) = coslat(iCol)
     if (doIons) then
        do s = 1, nIons
           IDensityS(iLon,iLat,-1:nA+2,s,iB) = ions(:,s,iCol)
        enddo
        do d = 1, 3
           IVelocity(iLon,iLat,-1:nA+2,d,iB) = ivel(:,d,iCol)
        enddo
        SZA(iLon,iLat,iB) = sza_(iCol)
        MLatitude(iLon,iLat,nA,iB) = mlat(iCol)
     endif
     if (doMsis) then
        ! per latitude row and longitude of a block, as GITM holds them
        Latitude(iLat,iB) = latr(iCol)
        Longitude(iLon,iB) = lonr(iCol)
     endif
     if (doTides) then
        TidesEast(iLon,iLat,1:2,iB) = tides(:,1,iCol)
        TidesNorth(iLon,iLat,1:2,iB) = tides(:,2,iCol)
        TidesTemp(iLon,iLat,1:2,iB) = tides(:,3,iCol)
     endif
  enddo

  status = 0
  do iStep = 1, nStepsH
     if (doMsis) call fill_msis_bc
     call advance_vertical_all_gpu(iError)
     if (iError /= 0) then
        status(1) = iError
        exit
     endif
  enddo

  do iCol = 1, nCols
     iLon = mod(iCol-1, nLons) + 1
     iLat = mod((iCol-1)/nLons, nLats) + 1
     iB = (iCol-1)/(nLons*nLats) + 1
     do s = 1, nSpecies
        nds(:,s,iCol) = NDensityS(iLon,iLat,-1:nA+2,s,iB)
        vvel(:,s,iCol) = VerticalVelocity(iLon,iLat,-1:nA+2,s,iB)
        logns_o(:,s,iCol) = LogNS(iLon,iLat,-1:nA+2,s,iB)
     enddo
     do d = 1, 3
        vel(:,d,iCol) = Velocity(iLon,iLat,-1:nA+2,d,iB)
     enddo
     temp(:,iCol) = Temperature(iLon,iLat,-1:nA+2,iB)
     rhoH(:,iCol) = Rho(iLon,iLat,-1:nA+2,iB)
     nden_o(:,iCol) = NDensity(iLon,iLat,-1:nA+2,iB)
     if (doIons) then
        do s = 1, nIons
           ions(:,s,iCol) = IDensityS(iLon,iLat,-1:nA+2,s,iB)
        enddo
     endif
  enddo

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) status
  write(11) nds
  write(11) vvel
  write(11) vel
  write(11) temp
  write(11) rhoH
  write(11) logns_o
  write(11) nden_o
  if (doIons) write(11) ions
  if (doMsis) then
     ! the level-0 MSIS values the drop-in passed, per column
     do iCol = 1, nCols
        iLon = mod(iCol-1, nLons) + 1
        iLat = mod((iCol-1)/nLons, nLats) + 1
        iB = (iCol-1)/(nLons*nLats) + 1
        msis_o(:,iCol) = MsisBC(:,iLon,iLat,iB)
     enddo
     write(11) msis_o
  endif
  close(11)

end program gitm_vert_gpu_driver
