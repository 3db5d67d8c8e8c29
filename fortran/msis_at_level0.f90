! msis_at_level0.f90 -- the level-0 MSIS values the drop-in's UseMSIS branch
! takes from its caller (gitm_vert_columns%msis_bc).
!
! set_vertical_bcs (set_vertical_bcs.Earth.f90:112-137) calls the
! reference's msis_bcs (init_msis.Earth.f90:367) for levels -1 and 0 at every
! boundary pass.  Its inputs (time, place, solar indices) do not change
! within a step, and the library uses only level 0's values (the boundary
! conditions overwrite every level -1 value), so one call per column and
! step gives what all the passes use.  msis_at_level0 makes that call with
! the arguments set_vertical_bcs forms; fill_msis_bc does it for every
! column into MsisBC (ModVerticalGPU).

subroutine msis_at_level0(iLon, iLat, iBlock, LogNS, Temp, v)

  use ModPlanet, only: nSpecies
  use ModGITM, only: Latitude, Longitude, Altitude_GB
  use ModInputs, only: MsisOblateFactor, f107, f107a
  use ModConstants, only: pi
  use ModTime, only: UTime, iJulianDay, CurrentTime
  use ModIndicesInterfaces, only: get_HPI

  implicit none

  integer, intent(in) :: iLon, iLat, iBlock
  real, intent(out) :: LogNS(nSpecies), Temp, v(2)

  real :: Lat, Lon, Alt, Lst, HP, Ap, LogRho
  integer :: iError

  ! the column's place in degrees (advance_vertical.f90:137-138)
  Lat = Latitude(iLat, iBlock) * 180.0/pi
  Lon = Longitude(iLon, iBlock) * 180.0/pi

  ! Ap from the hemispheric power, then msis_bcs at Altitude(0)
  ! (set_vertical_bcs.Earth.f90:114-131, iAlt = 0)
  call get_HPI(CurrentTime, HP, iError)
  if (iError > 0) HP = 40.0
  Ap = min(200.,max(-40.72 + 1.3 * HP, 10.))
  Alt = Altitude_GB(iLon,iLat,0,iBlock)/1000.0 * &
       (1.0 - &
       MsisOblateFactor/2.0 + &
       MsisOblateFactor * cos(Lat*3.1415/180.0))
  Lst = mod(UTime/3600.0+Lon/15.0,24.0)

  call msis_bcs(iJulianDay,UTime,Alt,Lat,Lon,Lst, &
       F107A,F107,AP,LogNS, Temp, LogRho, v)

end subroutine msis_at_level0

! MsisBC(1:9, iLon, iLat, iBlock): LogNS(0,1:6), Temp(0), Vel(0, east),
! Vel(0, north) of every column, from msis_at_level0.
subroutine fill_msis_bc

  use ModSizeGitm, only: nLons, nLats
  use ModGITM, only: nBlocks
  use ModVerticalGPU

  implicit none

  integer :: iLon, iLat, iBlock

  if (.not. allocated(MsisBC)) allocate(MsisBC(9, nLons, nLats, nBlocks))
  do iBlock = 1, nBlocks
     do iLat = 1, nLats
        do iLon = 1, nLons
           call msis_at_level0(iLon, iLat, iBlock, MsisBC(1:6,iLon,iLat,iBlock), &
                MsisBC(7,iLon,iLat,iBlock), MsisBC(8:9,iLon,iLat,iBlock))
        enddo
     enddo
  enddo

end subroutine fill_msis_bc
