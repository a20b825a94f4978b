! getm_external.F90 -- the reference's external subroutines with their own
! argument lists, for a GETM build (Cartesian: no SPHERICAL/CURVILINEAR)
! that links libgetm_transport in place of
!   src/2d/adv_split_u.F90, src/2d/adv_split_v.F90, src/2d/adv_upstream_2dh.F90,
!   src/3d/adv_split_w.F90, src/3d/ip_blumberg_mellor.F90,
!   src/3d/ip_shchepetkin_mcwilliams.F90.
! The advection routines take their bounds from the domain module, as the
! explicit interfaces of advection.F90:91-194 and advection_3d.F90:55-68
! declare them; the argument-less pressure routines read and write the
! module variables of internal_pressure (internal_pressure.F90:62-92) and
! sseo of variables_3d, as the reference's do.  getm_transport_gpu must have
! been initialised (gt_init) on the same grid.  Each call copies its host
! arrays to the device, runs the routine there and copies the in/out arrays
! back.
#define HALO 2

   subroutine adv_split_u(dt,f,fi,Di,adv,U,DU,splitfac,scheme,AH,mask_flux,mask_update)
   use, intrinsic :: iso_c_binding, only: c_double
   use domain, only: imin,imax,jmin,jmax
   use getm_transport_gpu, only: adv_split_u_gpu
   implicit none
   real(c_double), intent(in) :: dt, splitfac, AH
   real(c_double), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: f, U, DU
   integer, intent(in) :: scheme
   logical, dimension(:,:), pointer, intent(in) :: mask_flux
   logical, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: mask_update
   real(c_double), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(inout) :: &
        fi, Di, adv
   call adv_split_u_gpu(dt, f, fi, Di, adv, U, DU, splitfac, scheme, AH, mask_flux, &
                        mask_update)
   end subroutine adv_split_u

   subroutine adv_split_v(dt,f,fi,Di,adv,V,DV,splitfac,scheme,AH,mask_flux,mask_update)
   use, intrinsic :: iso_c_binding, only: c_double
   use domain, only: imin,imax,jmin,jmax
   use getm_transport_gpu, only: adv_split_v_gpu
   implicit none
   real(c_double), intent(in) :: dt, splitfac, AH
   real(c_double), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: f, V, DV
   integer, intent(in) :: scheme
   logical, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO-1), intent(in) :: mask_flux
   logical, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: mask_update
   real(c_double), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(inout) :: &
        fi, Di, adv
   call adv_split_v_gpu(dt, f, fi, Di, adv, V, DV, splitfac, scheme, AH, mask_flux, &
                        mask_update)
   end subroutine adv_split_v

   subroutine adv_split_w(dt,f,fi,hi,adv,ww,splitfac,scheme,tag,az,itersmax)
   use, intrinsic :: iso_c_binding, only: c_double
   use domain, only: imin,imax,jmin,jmax,kmax
   use getm_transport_gpu, only: adv_split_w_gpu
   implicit none
   real(c_double), intent(in) :: dt, splitfac
   real(c_double), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO,0:kmax), intent(in) :: &
        f, ww
   integer, intent(in) :: scheme, tag, itersmax
   integer, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: az
   real(c_double), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO,0:kmax), &
        intent(inout) :: fi, hi, adv
   call adv_split_w_gpu(dt, f, fi, hi, adv, ww, splitfac, scheme, tag, az, itersmax)
   end subroutine adv_split_w

   subroutine adv_upstream_2dh(dt,f,fi,Di,adv,U,V,Dn,DU,DV,AH,az)
   use, intrinsic :: iso_c_binding, only: c_double
   use domain, only: imin,imax,jmin,jmax
   use getm_transport_gpu, only: adv_upstream_2dh_gpu
   implicit none
   real(c_double), intent(in) :: dt, AH
   real(c_double), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: &
        f, U, V, Dn, DU, DV
   integer, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: az
   real(c_double), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(inout) :: &
        fi, Di, adv
   call adv_upstream_2dh_gpu(dt, f, fi, Di, adv, U, V, Dn, DU, DV, AH, az)
   end subroutine adv_upstream_2dh

   subroutine ip_blumberg_mellor()
   use internal_pressure
   use getm_transport_gpu, only: ip_blumberg_mellor_gpu
   implicit none
   call ip_blumberg_mellor_gpu(H, hn, hun, hvn, buoy, zz, idpdx, idpdy)
   end subroutine ip_blumberg_mellor

   subroutine ip_shchepetkin_mcwilliams()
   use internal_pressure
   use variables_3d, only: buoy, sseo
   use getm_transport_gpu, only: ip_shchepetkin_mcwilliams_gpu
   implicit none
   call ip_shchepetkin_mcwilliams_gpu(H, hn, hun, hvn, buoy, sseo, zz, idpdx, idpdy)
   end subroutine ip_shchepetkin_mcwilliams
