! ext_driver.F90 -- a Fortran caller of the external routines of
! getm_external.F90 with the reference's argument lists, used by the
! Fortran parity test.  It reads a case written by the test (stream I/O),
! builds the advection masks as init_advection does (advection.F90:263-318:
! the H grid's, and the U grid's shifted flux mask of size NX-1), calls
!   adv_split_u (H grid and U grid), adv_split_v (H grid, the NY-1-row view
!   of the flux mask), adv_split_w, adv_upstream_2dh on slice k of the case,
!   ip_blumberg_mellor() and ip_shchepetkin_mcwilliams() on the module
!   variables,
! and writes every result back.
!   usage: ext_driver <input file> <output file>
#define HALO 2
program ext_driver
   use, intrinsic :: iso_c_binding
   use getm_transport_gpu, only: gt_init, gt_finalize
   use domain
   use variables_3d
   use internal_pressure, only: zz
   implicit none
   interface
      subroutine adv_split_u(dt,f,fi,Di,adv,U,DU,splitfac,scheme,AH,mask_flux,mask_update)
      use domain, only: imin,imax,jmin,jmax
      real(8), intent(in) :: dt, splitfac, AH
      real(8), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: f, U, DU
      integer, intent(in) :: scheme
      logical, dimension(:,:), pointer, intent(in) :: mask_flux
      logical, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: mask_update
      real(8), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(inout) :: fi, Di, adv
      end subroutine adv_split_u
      subroutine adv_split_v(dt,f,fi,Di,adv,V,DV,splitfac,scheme,AH,mask_flux,mask_update)
      use domain, only: imin,imax,jmin,jmax
      real(8), intent(in) :: dt, splitfac, AH
      real(8), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: f, V, DV
      integer, intent(in) :: scheme
      logical, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO-1), intent(in) :: mask_flux
      logical, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: mask_update
      real(8), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(inout) :: fi, Di, adv
      end subroutine adv_split_v
      subroutine adv_split_w(dt,f,fi,hi,adv,ww,splitfac,scheme,tag,az,itersmax)
      use domain, only: imin,imax,jmin,jmax,kmax
      real(8), intent(in) :: dt, splitfac
      real(8), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO,0:kmax), intent(in) :: f, ww
      integer, intent(in) :: scheme, tag, itersmax
      integer, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: az
      real(8), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO,0:kmax), intent(inout) :: &
           fi, hi, adv
      end subroutine adv_split_w
      subroutine adv_upstream_2dh(dt,f,fi,Di,adv,U,V,Dn,DU,DV,AH,az)
      use domain, only: imin,imax,jmin,jmax
      real(8), intent(in) :: dt, AH
      real(8), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: &
           f, U, V, Dn, DU, DV
      integer, dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(in) :: az
      real(8), dimension(imin-HALO:imax+HALO,jmin-HALO:jmax+HALO), intent(inout) :: fi, Di, adv
      end subroutine adv_upstream_2dh
   end interface
   integer(c_int) :: hdr(8)
   real(c_double) :: rhdr(5)
   integer :: scheme, k, itersmax
   real(8) :: dt, AH, splitfac
   real(8), allocatable, dimension(:,:,:) :: f, uu, vv, ww, ho, fi3, hi3, adv3
   real(8), allocatable, dimension(:,:) :: fi, Di, adv
   logical, allocatable, target :: mask_updateH(:,:), mask_uflux(:,:), mask_vflux(:,:)
   logical, allocatable, target :: mask_ufluxU(:,:), mask_uupdateU(:,:)
   logical, pointer :: pflux(:,:)
   character(len=1024) :: fin, fout
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(10, file=trim(fin), access='stream', form='unformatted', status='old')
   read(10) hdr
   read(10) rhdr
   imin = hdr(1); imax = hdr(2); jmin = hdr(3); jmax = hdr(4); kmax = hdr(5)
   scheme = hdr(6); k = hdr(7); itersmax = hdr(8)
   dx = rhdr(1); dy = rhdr(2); dt = rhdr(3); AH = rhdr(4); splitfac = rhdr(5)
   allocate(az(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO))
   allocate(au, av, ax, mold=az)
   allocate(H(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO))
   allocate(sseo, fi, Di, adv, mold=H)
   allocate(f(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO, 0:kmax))
   allocate(uu, vv, ww, ho, hn, hun, hvn, buoy, idpdx, idpdy, zz, fi3, hi3, adv3, mold=f)
   read(10) az, au, av, ax
   read(10) H, sseo
   read(10) f, uu, vv, ww, hun, hvn, ho, hn, buoy
   close(10)
   call gt_init(imin, imax, jmin, jmax, kmax, dx, dy, az, au, av, ax)

   ! the masks of init_advection (advection.F90:263-318)
   allocate(mask_updateH(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO))
   allocate(mask_uflux, mask_vflux, mask_uupdateU, mold=mask_updateH)
   mask_updateH = (az == 1)
   mask_uflux = (au == 1 .or. au == 2)
   mask_vflux = (av == 1 .or. av == 2)
   mask_uupdateU = (au == 1)
   mask_uflux(imax+HALO, :) = .false.
   mask_vflux(:, jmax+HALO) = .false.
   allocate(mask_ufluxU(imin-HALO:imax+HALO-1, jmin-HALO:jmax+HALO))
   mask_ufluxU = mask_updateH(imin-HALO+1:imax+HALO, jmin-HALO:jmax+HALO)

   open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
   ! u, H grid (adv_gridH%mask_uflux => mask_uflux)
   fi = f(:, :, k); Di = ho(:, :, k); adv = 0.0d0
   pflux => mask_uflux
   call adv_split_u(dt, f(:, :, k), fi, Di, adv, uu(:, :, k), hun(:, :, k), splitfac, &
                    scheme, AH, pflux, mask_updateH)
   write(11) fi, Di, adv
   ! v, H grid (adv_gridH%mask_vflux => mask_vflux(:, jmin-HALO:jmax+HALO-1))
   fi = f(:, :, k); Di = ho(:, :, k); adv = 0.0d0
   call adv_split_v(dt, f(:, :, k), fi, Di, adv, vv(:, :, k), hvn(:, :, k), splitfac, &
                    scheme, AH, mask_vflux(:, jmin-HALO:jmax+HALO-1), mask_updateH)
   write(11) fi, Di, adv
   ! u, U grid (the shifted flux mask, NX-1 columns)
   fi = f(:, :, k); Di = ho(:, :, k); adv = 0.0d0
   pflux => mask_ufluxU
   call adv_split_u(dt, f(:, :, k), fi, Di, adv, uu(:, :, k), hun(:, :, k), splitfac, &
                    scheme, AH, pflux, mask_uupdateU)
   write(11) fi, Di, adv
   ! upstream 2DH
   fi = f(:, :, k); Di = ho(:, :, k); adv = 0.0d0
   call adv_upstream_2dh(dt, f(:, :, k), fi, Di, adv, uu(:, :, k), vv(:, :, k), hn(:, :, k), &
                         hun(:, :, k), hvn(:, :, k), AH, az)
   write(11) fi, Di, adv
   ! w, full columns
   fi3 = f; hi3 = ho; adv3 = 0.0d0
   call adv_split_w(dt, f, fi3, hi3, adv3, ww, 1.0d0, scheme, 10, az, itersmax)
   write(11) fi3, hi3, adv3
   ! internal pressure on the module variables
   zz = 0.0d0; idpdx = 0.0d0; idpdy = 0.0d0
   call ip_blumberg_mellor()
   write(11) zz, idpdx, idpdy
   zz = 0.0d0; idpdx = 0.0d0; idpdy = 0.0d0
   call ip_shchepetkin_mcwilliams()
   write(11) zz, idpdx, idpdy
   close(11)
   call gt_finalize()
end program ext_driver
