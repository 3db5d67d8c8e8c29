! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Driver that runs the reference GITM vertical step (compiled from the
! reference sources in place, see oracle/Makefile) on columns read from a
! binary file and writes the updated columns back.  Each column is loaded
! into grid cell (iLon,iLat,iBlock) = (1,1,1) of the reference's module
! globals, the reference subroutine advance_vertical(1,1,1)
! (src/advance_vertical.f90:45-219) is called n_steps times, and the result
! is read back from the same globals.  Columns are independent in the
! reference, so this is the same computation advance_vertical_all performs
! over a block (src/advance_vertical.f90:20-29).
!
! File formats (stream, native little-endian):
!   input : int32 hdr(16), real64 par(8), then the canonical column arrays
!           (level fastest, [col][component][level]) in the order
!           nds(6) vvel(6) vel(3) temp rho tempunit mmm gamma cmax eddy alt
!           grav invr dalt, coslat(col), and if ions: ions(10) ivel(3),
!           sza(col), mlat(col), and if tides (hdr(11) GSWM or hdr(12) WACCM):
!           TidesEast, TidesNorth, TidesTemp at levels -1, 0 ([col][3][2]);
!           hdr(13): UseDamping; hdr(14): UseMSIS, with hdr(15) = iJulianDay,
!           par(6:8) = UTime, F107, F107A and then Latitude(col),
!           Longitude(col) in radians
!   output: int32 status(4), then nds(6) vvel(6) vel(3) temp rho logns(6)
!           nden, and if ions: ions(10), and if UseMSIS: what msis_bcs
!           returns for each column's level 0 ([col][9]: LogNS(1:6), Temp,
!           east and north wind), computed as set_vertical_bcs computes it
program gitm_vert_ref

  use ModSizeGitm
  use ModPlanet
  use ModGITM
  use ModInputs
  use ModSources, only: KappaEddyDiffusion, EuvHeating, init_mod_sources
  use ModEUV, only: SZA
  use ModTides, only: TidesEast, TidesNorth, TidesTemp
  use ModTime, only: UTime, iJulianDay, CurrentTime
  use ModIndicesInterfaces, only: get_HPI
  use EUA_ModMsis00, only: meters, tselec

  implicit none

  integer :: hdr(16), status(4)
  real :: par(8)
  integer :: nCols, nA, solver, nStepsH, iCol, iStep, L, k, s, d, nIonsIO
  logical :: doIons, doTides, doMsis
  real, allocatable :: latr(:), lonr(:), msis_o(:,:)
  real :: LatD, LonD, AltM, Lst, HP, Ap, temptemp, LogRhoM, v(2), logNS_Species(6)
  integer :: iError
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
  nCols  = hdr(2)
  nA     = hdr(3)
  solver = hdr(4)
  nStepsH = hdr(5)
  if (hdr(1) /= 1196840530 .or. nA /= nAlts) then
     write(*,*) 'ref_harness: bad header or nAlts mismatch', hdr(1), nA, nAlts
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
  nIonsIO = nIons
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

  ! Model set-up the reference performs before any step
  nBlocks = 1
  call init_planet
  call init_mod_gitm
  call init_mod_sources

  ! Run flags (ModInputs) for this path
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
     ! the time and solar inputs msis_bcs reads (ModTime, ModInputs)
     iJulianDay = hdr(15)
     UTime = par(6)
     CurrentTime = par(6)
     f107 = par(7)
     f107a = par(8)
     ! no HWM-14 winds (their data files are not here): msis_bcs gives zero
     ! winds (init_msis.Earth.f90:448-456)
     UseMSISTides = .false.
     ! MSIS-00 set up as init_msis sets it (init_msis.Earth.f90:165-186)
     call meters(.true.)
     if (UseMsisTides) then
        sw_msis = 1
     else if (UseMSISOnly) then
        sw_msis = 1
        sw_msis(7) = 0
        sw_msis(8) = 0
        sw_msis(14) = 0
     else
        sw_msis = 0
        sw_msis(1) = 1
        sw_msis(9) = 1
     endif
     sw_msis(9) = 0
     sw_msis(2) = 0
     call tselec(sw_msis)
  endif
  ! calc_rates.Earth.f90:83-84
  ThermalDiffCoefS(1:nSpecies) = 0.0
  ThermalDiffCoefS(iHe_) = -0.38

  ! Fields advance_vertical reads but the vertical step never uses
  KappaTemp = 0.0
  dAltDLon_CB = 0.0
  dAltDLat_CB = 0.0
  ViscCoef = 0.0
  ViscCoefS = 0.0
  cp = 0.0
  EuvHeating = 0.0
  Latitude = 0.0
  Longitude = 0.0
  MLatitude = 0.0
  SZA = 0.0
  IVelocity = 0.0
  IDensityS = 1.0e6

  status = 0

  do iCol = 1, nCols
     do s = 1, nSpecies
        NDensityS(1,1,-1:nA+2,s,1) = nds(:,s,iCol)
        VerticalVelocity(1,1,-1:nA+2,s,1) = vvel(:,s,iCol)
     enddo
     do d = 1, 3
        Velocity(1,1,-1:nA+2,d,1) = vel(:,d,iCol)
     enddo
     Temperature(1,1,-1:nA+2,1) = temp(:,iCol)
     Rho(1,1,-1:nA+2,1) = rhoH(:,iCol)
     TempUnit(1,1,-1:nA+2) = tempunit_(:,iCol)
     MeanMajorMass(1,1,-1:nA+2) = mmm(:,iCol)
     Gamma(1,1,-1:nA+2,1) = gam(:,iCol)
     cMax_GDB(1,1,0:nA+1,iUp_,1) = cmax(2:L-1,iCol)
     KappaEddyDiffusion(1,1,-1:nA+2,1) = eddy(:,iCol)
     Altitude_GB(1,1,-1:nA+2,1) = alt(:,iCol)
     Gravity_GB(1,1,-1:nA+2,1) = grav(:,iCol)
     InvRadialDistance_GB(1,1,-1:nA+2,1) = invr(:,iCol)
     dAlt_GB(1,1,-1:nA+2,1) = dalt(:,iCol)
     CosLatitude(1,1) = coslat(iCol)
     if (doIons) then
        do s = 1, nIons
           IDensityS(1,1,-1:nA+2,s,1) = ions(:,s,iCol)
        enddo
        do d = 1, 3
           IVelocity(1,1,-1:nA+2,d,1) = ivel(:,d,iCol)
        enddo
        SZA(1,1,1) = sza_(iCol)
        MLatitude(1,1,nA,1) = mlat(iCol)
     endif
     if (doMsis) then
        Latitude(1,1) = latr(iCol)
        Longitude(1,1) = lonr(iCol)
        ! level 0's msis_bcs values, as set_vertical_bcs.Earth.f90:112-137
        ! computes them (advance_vertical.f90:137-138 for Lat and Lon)
        LatD = Latitude(1,1) * 180.0/pi
        LonD = Longitude(1,1) * 180.0/pi
        call get_HPI(CurrentTime, HP, iError)
        if (iError > 0) hp = 40.0
        Ap = min(200.,max(-40.72 + 1.3 * HP, 10.))
        AltM = Altitude_GB(1,1,0,1)/1000.0 * &
             (1.0 - &
             MsisOblateFactor/2.0 + &
             MsisOblateFactor * cos(LatD*3.1415/180.0))
        Lst = mod(UTime/3600.0+LonD/15.0,24.0)
        call msis_bcs(iJulianDay,UTime,AltM,LatD,LonD,Lst, &
             F107A,F107,AP,LogNS_Species, temptemp, LogRhoM, v)
        msis_o(1:6,iCol) = LogNS_Species
        msis_o(7,iCol) = temptemp
        msis_o(8:9,iCol) = v
     endif
     if (doTides) then
        TidesEast(1,1,1:2,1) = tides(:,1,iCol)
        TidesNorth(1,1,1:2,1) = tides(:,2,iCol)
        TidesTemp(1,1,1:2,1) = tides(:,3,iCol)
     endif

     do iStep = 1, nStepsH
        call advance_vertical(1,1,1)
     enddo

     do s = 1, nSpecies
        nds(:,s,iCol) = NDensityS(1,1,-1:nA+2,s,1)
        vvel(:,s,iCol) = VerticalVelocity(1,1,-1:nA+2,s,1)
        logns_o(:,s,iCol) = LogNS(1,1,-1:nA+2,s,1)
     enddo
     do d = 1, 3
        vel(:,d,iCol) = Velocity(1,1,-1:nA+2,d,1)
     enddo
     temp(:,iCol) = Temperature(1,1,-1:nA+2,1)
     rhoH(:,iCol) = Rho(1,1,-1:nA+2,1)
     nden_o(:,iCol) = NDensity(1,1,-1:nA+2,1)
     if (doIons) then
        do s = 1, nIons
           ions(:,s,iCol) = IDensityS(1,1,-1:nA+2,s,1)
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
  if (doMsis) write(11) msis_o
  close(11)

end program gitm_vert_ref
