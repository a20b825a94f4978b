! ext_host_modules.F90 -- for the test driver ext_driver.F90 only: modules
! named and shaped like the host model's domain, variables_3d and
! internal_pressure (the names getm_external.F90 uses: bounds, masks, H, dx,
! dy; hn, hun, hvn, buoy, sseo, idpdx, idpdy; zz), so the external routines
! compile and run outside a GETM build.  In a GETM build the model's own
! modules take their place and this file is not compiled.
#define HALO 2
   module domain
   implicit none
   integer :: imin, imax, jmin, jmax, kmax
   real(8) :: dx, dy
   integer, allocatable :: az(:,:), au(:,:), av(:,:), ax(:,:)
   real(8), allocatable :: H(:,:)
   end module domain

   module variables_3d
   implicit none
   real(8), allocatable :: hn(:,:,:), hun(:,:,:), hvn(:,:,:), buoy(:,:,:)
   real(8), allocatable :: idpdx(:,:,:), idpdy(:,:,:), sseo(:,:)
   end module variables_3d

   module internal_pressure
   use domain, only: imin, imax, jmin, jmax, kmax, az, au, av, H, dx, dy
   use variables_3d, only: hn, hun, hvn, buoy, idpdx, idpdy
   implicit none
   real(8), allocatable :: zz(:,:,:)
   end module internal_pressure
