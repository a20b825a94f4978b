! getm_transport_gpu.F90 -- Fortran binding of libgetm_transport.so.
!
! Keeps the Fortran call signatures of the reference routines
! (getm-model/code) so that the existing driver can call the MI355X kernels:
!   do_advection_3d    src/3d/advection_3d.F90:120-598
!   do_advection       src/2d/advection.F90:362-702
!   adv_split_u/v/w    src/2d/adv_split_u.F90, adv_split_v.F90, src/3d/adv_split_w.F90
!   adv_upstream_2dh   src/2d/adv_upstream_2dh.F90
!   update_3d_halo     src/futils/halo_zones.F90:141-187
! plus do_internal_pressure_gpu / do_eqstate_linear_gpu / do_eqstate_gpu /
! eos_internal_pressure_gpu and eqstate_internal_pressure_gpu (the linear or
! any EOS fused into the pressure methods),
! which take as arguments the module variables the argument-less reference
! routines (internal_pressure.F90, eqstate.F90) read, and the device-resident
! forms of the routines either side of the step:
!   start_macro_gpu        src/3d/start_macro.F90:54-86
!   coordinates_gpu        src/3d/coordinates.F90:66-107 (sigma, general)
!   ww_momentum_3d_gpu     src/3d/ww_momentum_3d.F90:66-119
!   salinity_diffusion_gpu, temperature_diffusion_gpu
!                          src/3d/salinity.F90:346-393, temperature.F90:463-551
!   uv_advect_3d_gpu       src/3d/uv_advect_3d.F90:8-466 (not J7)
!
! Host arrays are copied to device mirrors owned by this module, the step
! runs on the GPU, and the outputs are copied back.  Callers that keep their
! fields on the device use the *_dev entry points with type(c_ptr) handles
! from gt_malloc.  Errors stop with the reference's message.
#define HALO 2
   module getm_transport_gpu
   use, intrinsic :: iso_c_binding
   implicit none
   private

   public gt_init, gt_finalize
   public do_advection_3d, do_advection_3d_dev
   public do_internal_pressure_gpu, do_eqstate_linear_gpu, update_3d_halo
   public eos_internal_pressure_gpu, eqstate_internal_pressure_gpu
   public gt_malloc_f, gt_free_f, gt_to_device, gt_to_host, slice_to_device, slice_to_host
   public do_eqstate_gpu, start_macro_gpu, set_coordinates_gpu, coordinates_gpu
   public ww_momentum_3d_gpu, salinity_diffusion_gpu, temperature_diffusion_gpu
   public uv_advect_3d_gpu
   public adv_split_u_gpu, adv_split_v_gpu, adv_split_w_gpu, adv_upstream_2dh_gpu
   public ip_blumberg_mellor_gpu, ip_shchepetkin_mcwilliams_gpu

   integer, public, parameter :: NOSPLIT=0, FULLSPLIT=1, HALFSPLIT=2, HVSPLIT=3
   integer, public, parameter :: NOADV=0, UPSTREAM=1, UPSTREAM_2DH=2, P2=3
   integer, public, parameter :: SUPERBEE=4, MUSCL=5, P2_PDM=6
   integer, public, parameter :: H_TAG=10, D_TAG=20, U_TAG=31, V_TAG=32, W_TAG=33
   integer, public, parameter :: HU_TAG=11, HV_TAG=12, DU_TAG=21, DV_TAG=22, z_TAG=30

   type(c_ptr), save :: ctx = c_null_ptr
   integer, save :: il, ih, jl, jh, km
   integer(c_size_t), save :: n3 = 0, n2 = 0
   ! device mirrors: f uu vv ww hu hv ho hn hires advres
   type(c_ptr), save :: dbuf(10) = c_null_ptr
   ! device masks of the per-direction routines: two byte masks and one
   ! integer mask (adv_grid%az), E2DFIELD each
   type(c_ptr), save :: dmask(3) = c_null_ptr

   interface
      function c_gt_create(ctx, imin, imax, jmin, jmax, kmax, dx, dy, az, au, av, ax) &
               bind(C, name='gt_create') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), intent(out) :: ctx
         integer(c_int), value :: imin, imax, jmin, jmax, kmax
         real(c_double), value :: dx, dy
         integer(c_int), intent(in) :: az(*), au(*), av(*), ax(*)
         integer(c_int) :: rc
      end function
      function c_gt_destroy(ctx) bind(C, name='gt_destroy') result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int) :: rc
      end function
      function c_gt_synchronize(ctx) bind(C, name='gt_synchronize') result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int) :: rc
      end function
      function c_gt_do_advection_3d(ctx, dt, f, uu, vv, ww, hu, hv, ho, hn, split, &
               hscheme, vscheme, AH, tag, hires, advres) &
               bind(C, name='gt_do_advection_3d') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, f, uu, vv, ww, hu, hv, ho, hn, hires, advres
         real(c_double), value :: dt, AH
         integer(c_int), value :: split, hscheme, vscheme, tag
         integer(c_int) :: rc
      end function
      function c_gt_adv_split_u_m(ctx, dt, f, fi, Di, adv, U, DU, splitfac, scheme, AH, &
               mflux, mupd) bind(C, name='gt_adv_split_u_m') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, f, fi, Di, adv, U, DU, mflux, mupd
         real(c_double), value :: dt, splitfac, AH
         integer(c_int), value :: scheme
         integer(c_int) :: rc
      end function
      function c_gt_adv_split_v_m(ctx, dt, f, fi, Di, adv, V, DV, splitfac, scheme, AH, &
               mflux, mupd) bind(C, name='gt_adv_split_v_m') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, f, fi, Di, adv, V, DV, mflux, mupd
         real(c_double), value :: dt, splitfac, AH
         integer(c_int), value :: scheme
         integer(c_int) :: rc
      end function
      function c_gt_adv_split_w_m(ctx, dt, f, fi, hi, adv, ww, splitfac, scheme, tag, az, &
               itersmax) bind(C, name='gt_adv_split_w_m') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, f, fi, hi, adv, ww, az
         real(c_double), value :: dt, splitfac
         integer(c_int), value :: scheme, tag, itersmax
         integer(c_int) :: rc
      end function
      function c_gt_adv_upstream_2dh_m(ctx, dt, f, fi, Di, adv, U, V, Dn, DU, DV, AH, az) &
               bind(C, name='gt_adv_upstream_2dh_m') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, f, fi, Di, adv, U, V, Dn, DU, DV, az
         real(c_double), value :: dt, AH
         integer(c_int) :: rc
      end function
      function c_gt_ip_blumberg_mellor(ctx, H, hn, hun, hvn, buoy, zz, idpdx, idpdy) &
               bind(C, name='gt_ip_blumberg_mellor') result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, H, hn, hun, hvn, buoy, zz, idpdx, idpdy
         integer(c_int) :: rc
      end function
      function c_gt_ip_shchepetkin_mcwilliams(ctx, H, hn, hun, hvn, buoy, sseo, zz, idpdx, &
               idpdy) bind(C, name='gt_ip_shchepetkin_mcwilliams') result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, H, hn, hun, hvn, buoy, sseo, zz, idpdx, idpdy
         integer(c_int) :: rc
      end function
      function c_gt_do_internal_pressure(ctx, method, H, hn, hun, hvn, buoy, sseo, &
               idpdx, idpdy) bind(C, name='gt_do_internal_pressure') result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, H, hn, hun, hvn, buoy, sseo, idpdx, idpdy
         integer(c_int), value :: method
         integer(c_int) :: rc
      end function
      function c_gt_eos_linear(ctx, T, S, rho, buoy, T0, S0, dtr0, dsr0) &
               bind(C, name='gt_eos_linear') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, T, S, rho, buoy
         real(c_double), value :: T0, S0, dtr0, dsr0
         integer(c_int) :: rc
      end function
      function c_gt_eos_internal_pressure(ctx, methods, T, S, T0, S0, dtr0, dsr0, H, hn, &
               hun, hvn, sseo, zz_bm, idpdx_bm, idpdy_bm, zz_sm, idpdx_sm, idpdy_sm) &
               bind(C, name='gt_eos_internal_pressure') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, T, S, H, hn, hun, hvn, sseo
         type(c_ptr), value :: zz_bm, idpdx_bm, idpdy_bm, zz_sm, idpdx_sm, idpdy_sm
         integer(c_int), value :: methods
         real(c_double), value :: T0, S0, dtr0, dsr0
         integer(c_int) :: rc
      end function
      function c_gt_eqstate_internal_pressure(ctx, method, nonneg, methods, T, S, rho, T0, &
               S0, dtr0, dsr0, H, hn, hun, hvn, sseo, zz_bm, idpdx_bm, idpdy_bm, zz_sm, &
               idpdx_sm, idpdy_sm) bind(C, name='gt_eqstate_internal_pressure') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, T, S, rho, H, hn, hun, hvn, sseo
         type(c_ptr), value :: zz_bm, idpdx_bm, idpdy_bm, zz_sm, idpdx_sm, idpdy_sm
         integer(c_int), value :: method, nonneg, methods
         real(c_double), value :: T0, S0, dtr0, dsr0
         integer(c_int) :: rc
      end function
      function c_gt_update_3d_halo(ctx, f1, f2, tag, mirror) &
               bind(C, name='gt_update_3d_halo') result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, f1, f2
         integer(c_int), value :: tag, mirror
         integer(c_int) :: rc
      end function
      function c_gt_malloc(p, bytes) bind(C, name='gt_malloc') result(rc)
         import :: c_ptr, c_int, c_size_t
         type(c_ptr), intent(out) :: p
         integer(c_size_t), value :: bytes
         integer(c_int) :: rc
      end function
      function c_gt_free(p) bind(C, name='gt_free') result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: p
         integer(c_int) :: rc
      end function
      function c_gt_memcpy_h2d(dst, src, bytes) bind(C, name='gt_memcpy_h2d') result(rc)
         import :: c_ptr, c_int, c_size_t
         type(c_ptr), value :: dst, src
         integer(c_size_t), value :: bytes
         integer(c_int) :: rc
      end function
      function c_gt_memcpy_d2h(dst, src, bytes) bind(C, name='gt_memcpy_d2h') result(rc)
         import :: c_ptr, c_int, c_size_t
         type(c_ptr), value :: dst, src
         integer(c_size_t), value :: bytes
         integer(c_int) :: rc
      end function
      function c_gt_last_error() bind(C, name='gt_last_error') result(p)
         import :: c_ptr
         type(c_ptr) :: p
      end function
      function c_gt_do_eqstate(ctx, method, nonneg, T, S, hn, rho, buoy, alpha, beta, &
               T0, S0, dtr0, dsr0) bind(C, name='gt_do_eqstate') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, T, S, hn, rho, buoy, alpha, beta
         integer(c_int), value :: method, nonneg
         real(c_double), value :: T0, S0, dtr0, dsr0
         integer(c_int) :: rc
      end function
      function c_gt_start_macro(ctx, M, min_depth, H, HU, HV, z, sseo, ssen, ssuo, ssun, &
               ssvo, ssvn, Dn, Dun, Dvn, Uint, Vint) bind(C, name='gt_start_macro') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, H, HU, HV, z, sseo, ssen, ssuo, ssun, ssvo, ssvn
         type(c_ptr), value :: Dn, Dun, Dvn, Uint, Vint
         integer(c_int), value :: M
         real(c_double), value :: min_depth
         integer(c_int) :: rc
      end function
      function c_gt_set_coordinates(ctx, vert_cord, ddu, ddl, d_gamma, gamma_surf, min_depth) &
               bind(C, name='gt_set_coordinates') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: vert_cord, gamma_surf
         real(c_double), value :: ddu, ddl, d_gamma, min_depth
         integer(c_int) :: rc
      end function
      function c_gt_coordinates(ctx, dt, cord_relax, maxdepth, H, HU, HV, sseo, ssen, ssuo, &
               ssun, ssvo, ssvn, ho, hn, huo, hun, hvo, hvn, zc) &
               bind(C, name='gt_coordinates') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, H, HU, HV, sseo, ssen, ssuo, ssun, ssvo, ssvn
         type(c_ptr), value :: ho, hn, huo, hun, hvo, hvn, zc
         real(c_double), value :: dt, cord_relax, maxdepth
         integer(c_int) :: rc
      end function
      function c_gt_ww_momentum_3d(ctx, dt, uu, vv, ho, hn, ww) &
               bind(C, name='gt_ww_momentum_3d') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, uu, vv, ho, hn, ww
         real(c_double), value :: dt
         integer(c_int) :: rc
      end function
      function c_gt_uv_advect_3d(ctx, dt, uu, vv, ww, ho, hn, hun, hvn, split, hscheme, &
               vscheme, uuEx, vvEx) bind(C, name='gt_uv_advect_3d') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, uu, vv, ww, ho, hn, hun, hvn, uuEx, vvEx
         real(c_double), value :: dt
         integer(c_int), value :: split, hscheme, vscheme
         integer(c_int) :: rc
      end function
      function c_gt_salinity_diffusion(ctx, dt, cnpar, hn, nuh, S) &
               bind(C, name='gt_salinity_diffusion') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, hn, nuh, S
         real(c_double), value :: dt, cnpar
         integer(c_int) :: rc
      end function
      function c_gt_temperature_diffusion(ctx, dt, cnpar, hn, nuh, swr, shf, A, g1, g2, &
               dry_z, S, refl, minfrac, rad, T) &
               bind(C, name='gt_temperature_diffusion') result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, hn, nuh, swr, shf, A, g1, g2, dry_z, S, rad, T
         real(c_double), value :: dt, cnpar, refl, minfrac
         integer(c_int) :: rc
      end function
   end interface

   contains

   subroutine check(rc, where)
   integer(c_int), intent(in) :: rc
   character(len=*), intent(in) :: where
   character(kind=c_char), pointer :: msg(:)
   character(len=512) :: text
   integer :: n
   if (rc == 0) return
   call c_f_pointer(c_gt_last_error(), msg, [512])
   text = ' '
   do n = 1, 512
      if (msg(n) == c_null_char) exit
      text(n:n) = msg(n)
   end do
   write(0,*) 'FATAL ERROR: ', where, ': ', trim(text)
   stop 1
   end subroutine check

   ! masks of init_domain (src/domain/domain.F90:249-346); au/av/ax as there
   subroutine gt_init(imin, imax, jmin, jmax, kmax, dx, dy, az, au, av, ax)
   integer, intent(in) :: imin, imax, jmin, jmax, kmax
   real(c_double), intent(in) :: dx, dy
   integer(c_int), intent(in) :: az(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO)
   integer(c_int), intent(in) :: au(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO)
   integer(c_int), intent(in) :: av(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO)
   integer(c_int), intent(in) :: ax(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO)
   integer :: q
   il = imin; ih = imax; jl = jmin; jh = jmax; km = kmax
   call check(c_gt_create(ctx, imin, imax, jmin, jmax, kmax, dx, dy, az, au, av, ax), &
              'gt_init')
   n2 = int(imax - imin + 1 + 2*HALO, c_size_t) * int(jmax - jmin + 1 + 2*HALO, c_size_t)
   n3 = n2 * int(kmax + 1, c_size_t)
   do q = 1, size(dbuf)
      call check(c_gt_malloc(dbuf(q), 8_c_size_t * n3), 'gt_init')
   end do
   do q = 1, size(dmask)
      call check(c_gt_malloc(dmask(q), 4_c_size_t * n2), 'gt_init')
   end do
   end subroutine gt_init

   subroutine gt_finalize()
   integer :: q
   do q = 1, size(dbuf)
      if (c_associated(dbuf(q))) call check(c_gt_free(dbuf(q)), 'gt_finalize')
      dbuf(q) = c_null_ptr
   end do
   do q = 1, size(dmask)
      if (c_associated(dmask(q))) call check(c_gt_free(dmask(q)), 'gt_finalize')
      dmask(q) = c_null_ptr
   end do
   if (c_associated(ctx)) call check(c_gt_destroy(ctx), 'gt_finalize')
   ctx = c_null_ptr
   end subroutine gt_finalize

   subroutine gt_malloc_f(p, nbytes)
   type(c_ptr), intent(out) :: p
   integer(c_size_t), intent(in) :: nbytes
   call check(c_gt_malloc(p, nbytes), 'gt_malloc')
   end subroutine gt_malloc_f

   subroutine gt_free_f(p)
   type(c_ptr), intent(in) :: p
   call check(c_gt_free(p), 'gt_free')
   end subroutine gt_free_f

   subroutine gt_to_device(dst, src)
   type(c_ptr), intent(in) :: dst
   real(c_double), intent(in), target :: src(*)
   call check(c_gt_memcpy_h2d(dst, c_loc(src), 8_c_size_t * n3), 'gt_to_device')
   end subroutine gt_to_device

   subroutine gt_to_host(dst, src)
   real(c_double), intent(inout), target :: dst(*)
   type(c_ptr), intent(in) :: src
   call check(c_gt_memcpy_d2h(c_loc(dst), src, 8_c_size_t * n3), 'gt_to_host')
   end subroutine gt_to_host

   ! do_advection_3d with the reference's argument list on host arrays
   subroutine do_advection_3d(dt, f, uu, vv, ww, hu, hv, ho, hn, split, hscheme, vscheme, &
                              AH, tag, hires, advres)
   real(c_double), intent(in) :: dt, AH
   real(c_double), intent(in), target :: uu(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km)
   real(c_double), intent(in), target :: vv(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km)
   real(c_double), intent(in), target :: ww(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km)
   real(c_double), intent(in), target :: hu(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km)
   real(c_double), intent(in), target :: hv(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km)
   real(c_double), intent(in), target :: ho(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km)
   real(c_double), intent(in), target :: hn(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km)
   integer, intent(in) :: split, hscheme, vscheme, tag
   real(c_double), intent(inout), target :: f(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km)
   real(c_double), intent(out), target, optional :: &
        hires(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km), &
        advres(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km)
   type(c_ptr) :: ph, pa
   call gt_to_device(dbuf(1), f)
   call gt_to_device(dbuf(2), uu)
   call gt_to_device(dbuf(3), vv)
   call gt_to_device(dbuf(4), ww)
   call gt_to_device(dbuf(5), hu)
   call gt_to_device(dbuf(6), hv)
   call gt_to_device(dbuf(7), ho)
   call gt_to_device(dbuf(8), hn)
   ph = c_null_ptr
   pa = c_null_ptr
   if (present(hires)) ph = dbuf(9)
   if (present(advres)) pa = dbuf(10)
   call do_advection_3d_dev(dt, dbuf(1), dbuf(2), dbuf(3), dbuf(4), dbuf(5), dbuf(6), &
                            dbuf(7), dbuf(8), split, hscheme, vscheme, AH, tag, ph, pa)
   call gt_to_host(f, dbuf(1))
   if (present(hires)) call gt_to_host(hires, dbuf(9))
   if (present(advres)) call gt_to_host(advres, dbuf(10))
   end subroutine do_advection_3d

   ! the same on device-resident fields (c_null_ptr for absent outputs)
   subroutine do_advection_3d_dev(dt, f, uu, vv, ww, hu, hv, ho, hn, split, hscheme, &
                                  vscheme, AH, tag, hires, advres)
   real(c_double), intent(in) :: dt, AH
   type(c_ptr), intent(in) :: f, uu, vv, ww, hu, hv, ho, hn, hires, advres
   integer, intent(in) :: split, hscheme, vscheme, tag
   call check(c_gt_do_advection_3d(ctx, dt, f, uu, vv, ww, hu, hv, ho, hn, split, &
              hscheme, vscheme, AH, tag, hires, advres), 'do_advection_3d')
   call check(c_gt_synchronize(ctx), 'do_advection_3d')
   end subroutine do_advection_3d_dev

   ! do_eqstate (eqstate_method=1) on device-resident T, S -> buoy
   subroutine do_eqstate_linear_gpu(T, S, buoy, T0, S0, dtr0, dsr0)
   type(c_ptr), intent(in) :: T, S, buoy
   real(c_double), intent(in) :: T0, S0, dtr0, dsr0
   call check(c_gt_eos_linear(ctx, T, S, c_null_ptr, buoy, T0, S0, dtr0, dsr0), &
              'do_eqstate')
   end subroutine do_eqstate_linear_gpu

   ! do_internal_pressure (internal_pressure.F90:181-274), device-resident
   subroutine do_internal_pressure_gpu(ip_method, H, hn, hun, hvn, buoy, sseo, idpdx, idpdy)
   integer, intent(in) :: ip_method
   type(c_ptr), intent(in) :: H, hn, hun, hvn, buoy, sseo, idpdx, idpdy
   call check(c_gt_do_internal_pressure(ctx, ip_method, H, hn, hun, hvn, buoy, sseo, &
              idpdx, idpdy), 'do_internal_pressure()')
   call check(c_gt_synchronize(ctx), 'do_internal_pressure()')
   end subroutine do_internal_pressure_gpu

   ! the linear do_eqstate (eqstate.F90:127-149,252-261) fused into
   ! ip_blumberg_mellor (methods 1), ip_shchepetkin_mcwilliams (2) or both (3),
   ! device-resident: the gradients of do_eqstate_linear_gpu followed by the
   ! methods, bit for bit, without storing buoy; the outputs of a method not
   ! asked for (and sseo without Shchepetkin-McWilliams) may be c_null_ptr
   subroutine eos_internal_pressure_gpu(methods, T, S, T0, S0, dtr0, dsr0, H, hn, hun, &
                                        hvn, sseo, zz_bm, idpdx_bm, idpdy_bm, zz_sm, &
                                        idpdx_sm, idpdy_sm)
   integer, intent(in) :: methods
   type(c_ptr), intent(in) :: T, S, H, hn, hun, hvn, sseo
   type(c_ptr), intent(in) :: zz_bm, idpdx_bm, idpdy_bm, zz_sm, idpdx_sm, idpdy_sm
   real(c_double), intent(in) :: T0, S0, dtr0, dsr0
   call check(c_gt_eos_internal_pressure(ctx, methods, T, S, T0, S0, dtr0, dsr0, H, hn, &
              hun, hvn, sseo, zz_bm, idpdx_bm, idpdy_bm, zz_sm, idpdx_sm, idpdy_sm), &
              'eos_internal_pressure()')
   end subroutine eos_internal_pressure_gpu

   ! do_eqstate of eqstate_method 1, 2 or 3 (eqstate.F90:94-261) fused into the
   ! pressure methods as eos_internal_pressure_gpu, device-resident: rho is
   ! read on land columns (methods 2 and 3; c_null_ptr allowed for 1) and
   ! never written; the gradients are those of do_eqstate_gpu followed by
   ! the methods, bit for bit
   subroutine eqstate_internal_pressure_gpu(method, methods, T, S, rho, T0, S0, dtr0, dsr0, &
                                            H, hn, hun, hvn, sseo, zz_bm, idpdx_bm, &
                                            idpdy_bm, zz_sm, idpdx_sm, idpdy_sm)
   integer, intent(in) :: method, methods
   type(c_ptr), intent(in) :: T, S, rho, H, hn, hun, hvn, sseo
   type(c_ptr), intent(in) :: zz_bm, idpdx_bm, idpdy_bm, zz_sm, idpdx_sm, idpdy_sm
   real(c_double), intent(in) :: T0, S0, dtr0, dsr0
   integer(c_int) :: nonneg
   nonneg = 0
#ifdef NONNEGSALT
   nonneg = 1
#endif
   call check(c_gt_eqstate_internal_pressure(ctx, method, nonneg, methods, T, S, rho, T0, &
              S0, dtr0, dsr0, H, hn, hun, hvn, sseo, zz_bm, idpdx_bm, idpdy_bm, zz_sm, &
              idpdx_sm, idpdy_sm), 'eqstate_internal_pressure()')
   end subroutine eqstate_internal_pressure_gpu

   ! do_eqstate (eqstate.F90:94-271), eqstate_method 1, 2 or 3, device-resident;
   ! alpha/beta may be c_null_ptr (the _OLD_BVF_ build)
   subroutine do_eqstate_gpu(method, T, S, hn, rho, buoy, alpha, beta, T0, S0, dtr0, dsr0)
   integer, intent(in) :: method
   type(c_ptr), intent(in) :: T, S, hn, rho, buoy, alpha, beta
   real(c_double), intent(in) :: T0, S0, dtr0, dsr0
   integer(c_int) :: nonneg
   nonneg = 0
#ifdef NONNEGSALT
   nonneg = 1
#endif
   call check(c_gt_do_eqstate(ctx, method, nonneg, T, S, hn, rho, buoy, alpha, beta, &
              T0, S0, dtr0, dsr0), 'do_eqstate()')
   end subroutine do_eqstate_gpu

   ! start_macro (start_macro.F90:54-86), device-resident 2-D fields
   subroutine start_macro_gpu(M, min_depth, H, HU, HV, z, sseo, ssen, ssuo, ssun, ssvo, &
                              ssvn, Dn, Dun, Dvn, Uint, Vint)
   integer, intent(in) :: M
   real(c_double), intent(in) :: min_depth
   type(c_ptr), intent(in) :: H, HU, HV, z, sseo, ssen, ssuo, ssun, ssvo, ssvn
   type(c_ptr), intent(in) :: Dn, Dun, Dvn, Uint, Vint
   call check(c_gt_start_macro(ctx, M, min_depth, H, HU, HV, z, sseo, ssen, ssuo, ssun, &
              ssvo, ssvn, Dn, Dun, Dvn, Uint, Vint), 'start_macro()')
   end subroutine start_macro_gpu

   ! the domain-namelist values coordinates() uses (vert_cord 1, 2 or 3)
   subroutine set_coordinates_gpu(vert_cord, ddu, ddl, d_gamma, gamma_surf, min_depth)
   integer, intent(in) :: vert_cord
   real(c_double), intent(in) :: ddu, ddl, d_gamma, min_depth
   logical, intent(in) :: gamma_surf
   integer(c_int) :: gs
   gs = 0
   if (gamma_surf) gs = 1
   call check(c_gt_set_coordinates(ctx, vert_cord, ddu, ddl, d_gamma, gs, min_depth), &
              'coordinates()')
   end subroutine set_coordinates_gpu

   ! coordinates (coordinates.F90:66-107), device-resident; zc may be c_null_ptr
   subroutine coordinates_gpu(dt, cord_relax, maxdepth, H, HU, HV, sseo, ssen, ssuo, ssun, &
                              ssvo, ssvn, ho, hn, huo, hun, hvo, hvn, zc)
   real(c_double), intent(in) :: dt, cord_relax, maxdepth
   type(c_ptr), intent(in) :: H, HU, HV, sseo, ssen, ssuo, ssun, ssvo, ssvn
   type(c_ptr), intent(in) :: ho, hn, huo, hun, hvo, hvn, zc
   call check(c_gt_coordinates(ctx, dt, cord_relax, maxdepth, H, HU, HV, sseo, ssen, ssuo, &
              ssun, ssvo, ssvn, ho, hn, huo, hun, hvo, hvn, zc), 'coordinates')
   end subroutine coordinates_gpu

   ! ww_momentum_3d (ww_momentum_3d.F90:66-119), device-resident
   subroutine ww_momentum_3d_gpu(dt, uu, vv, ho, hn, ww)
   real(c_double), intent(in) :: dt
   type(c_ptr), intent(in) :: uu, vv, ho, hn, ww
   call check(c_gt_ww_momentum_3d(ctx, dt, uu, vv, ho, hn, ww), 'ww_momentum_3d()')
   end subroutine ww_momentum_3d_gpu

   ! uv_advect_3d (uv_advect_3d.F90:8-466), device-resident; the namelist
   ! values vel3d_adv_split/hor/ver of m3d are arguments
   subroutine uv_advect_3d_gpu(dt, uu, vv, ww, ho, hn, hun, hvn, split, hscheme, vscheme, &
                               uuEx, vvEx)
   real(c_double), intent(in) :: dt
   type(c_ptr), intent(in) :: uu, vv, ww, ho, hn, hun, hvn, uuEx, vvEx
   integer, intent(in) :: split, hscheme, vscheme
   call check(c_gt_uv_advect_3d(ctx, dt, uu, vv, ww, ho, hn, hun, hvn, split, hscheme, &
              vscheme, uuEx, vvEx), 'uv_advect_3d()')
   end subroutine uv_advect_3d_gpu

   ! the vertical diffusion of do_salinity (salinity.F90:346-393)
   subroutine salinity_diffusion_gpu(dt, cnpar, hn, nuh, S)
   real(c_double), intent(in) :: dt, cnpar
   type(c_ptr), intent(in) :: hn, nuh, S
   call check(c_gt_salinity_diffusion(ctx, dt, cnpar, hn, nuh, S), 'do_salinity()')
   end subroutine salinity_diffusion_gpu

   ! the short-wave and vertical diffusion of do_temperature (temperature.F90:463-551)
   subroutine temperature_diffusion_gpu(dt, cnpar, hn, nuh, swr, shf, A, g1, g2, dry_z, S, &
                                        swr_bot_refl_frac, swr_min_bot_frac, rad, T)
   real(c_double), intent(in) :: dt, cnpar, swr_bot_refl_frac, swr_min_bot_frac
   type(c_ptr), intent(in) :: hn, nuh, swr, shf, A, g1, g2, dry_z, S, rad, T
   call check(c_gt_temperature_diffusion(ctx, dt, cnpar, hn, nuh, swr, shf, A, g1, g2, &
              dry_z, S, swr_bot_refl_frac, swr_min_bot_frac, rad, T), 'do_temperature()')
   end subroutine temperature_diffusion_gpu

   ! ---- the per-direction routines with the reference's argument lists ----
   ! (advection.F90:91-194, advection_3d.F90:55-68; Cartesian build).  The
   ! slices and masks are copied to the device, the routine runs there, and
   ! the in/out arrays come back.

   ! (public: E2DFIELD host arrays to and from device buffers, e.g. H and
   ! sseo for the device-resident pressure routines)
   subroutine slice_to_device(dst, src)
   type(c_ptr), intent(in) :: dst
   real(c_double), intent(in), target :: src(*)
   call check(c_gt_memcpy_h2d(dst, c_loc(src), 8_c_size_t * n2), 'slice_to_device')
   end subroutine slice_to_device

   subroutine slice_to_host(dst, src)
   real(c_double), intent(inout), target :: dst(*)
   type(c_ptr), intent(in) :: src
   call check(c_gt_memcpy_d2h(c_loc(dst), src, 8_c_size_t * n2), 'slice_to_host')
   end subroutine slice_to_host

   ! a logical mask as E2DFIELD bytes: the entries the caller's array (or
   ! pointer view) holds, .false. elsewhere (never read there)
   subroutine mask_to_device(dst, m, i0, j0)
   type(c_ptr), intent(in) :: dst
   logical, intent(in) :: m(:,:)
   integer, intent(in) :: i0, j0   ! the caller's lower bounds of m
   integer(c_int8_t), allocatable, target :: b(:,:)
   integer :: i, j
   allocate(b(il-HALO:ih+HALO, jl-HALO:jh+HALO))
   b = 0_c_int8_t
   do j = max(j0, jl-HALO), min(j0 + size(m, 2) - 1, jh+HALO)
      do i = max(i0, il-HALO), min(i0 + size(m, 1) - 1, ih+HALO)
         if (m(i-i0+1, j-j0+1)) b(i, j) = 1_c_int8_t
      end do
   end do
   call check(c_gt_memcpy_h2d(dst, c_loc(b), n2), 'mask_to_device')
   end subroutine mask_to_device

   subroutine az_to_device(dst, az)
   type(c_ptr), intent(in) :: dst
   integer, intent(in), target :: az(il-HALO:ih+HALO, jl-HALO:jh+HALO)
   integer(c_int), allocatable, target :: b(:,:)
   allocate(b(il-HALO:ih+HALO, jl-HALO:jh+HALO))
   b = int(az, c_int)
   call check(c_gt_memcpy_h2d(dst, c_loc(b), 4_c_size_t * n2), 'az_to_device')
   end subroutine az_to_device

   ! adv_split_u (adv_split_u.F90:8-295); mask_flux as the reference's
   ! pointer dummy (its lower bounds kept), mask_update E2DFIELD
   subroutine adv_split_u_gpu(dt, f, fi, Di, adv, U, DU, splitfac, scheme, AH, mask_flux, &
                              mask_update)
   real(c_double), intent(in) :: dt, splitfac, AH
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(in) :: f, U, DU
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(inout) :: fi, Di, adv
   integer, intent(in) :: scheme
   logical, dimension(:,:), pointer, intent(in) :: mask_flux
   logical, dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(in) :: mask_update
   call slice_to_device(dbuf(1), f)
   call slice_to_device(dbuf(2), fi)
   call slice_to_device(dbuf(3), Di)
   call slice_to_device(dbuf(4), adv)
   call slice_to_device(dbuf(5), U)
   call slice_to_device(dbuf(6), DU)
   call mask_to_device(dmask(1), mask_flux, lbound(mask_flux, 1), lbound(mask_flux, 2))
   call mask_to_device(dmask(2), mask_update, il-HALO, jl-HALO)
   call check(c_gt_adv_split_u_m(ctx, dt, dbuf(1), dbuf(2), dbuf(3), dbuf(4), dbuf(5), &
              dbuf(6), splitfac, int(scheme, c_int), AH, dmask(1), dmask(2)), 'adv_split_u')
   call slice_to_host(fi, dbuf(2))
   call slice_to_host(Di, dbuf(3))
   call slice_to_host(adv, dbuf(4))
   end subroutine adv_split_u_gpu

   ! adv_split_v (adv_split_v.F90:8-145); mask_flux has NY-1 rows
   subroutine adv_split_v_gpu(dt, f, fi, Di, adv, V, DV, splitfac, scheme, AH, mask_flux, &
                              mask_update)
   real(c_double), intent(in) :: dt, splitfac, AH
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(in) :: f, V, DV
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(inout) :: fi, Di, adv
   integer, intent(in) :: scheme
   logical, dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO-1), intent(in) :: mask_flux
   logical, dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(in) :: mask_update
   call slice_to_device(dbuf(1), f)
   call slice_to_device(dbuf(2), fi)
   call slice_to_device(dbuf(3), Di)
   call slice_to_device(dbuf(4), adv)
   call slice_to_device(dbuf(5), V)
   call slice_to_device(dbuf(6), DV)
   call mask_to_device(dmask(1), mask_flux, il-HALO, jl-HALO)
   call mask_to_device(dmask(2), mask_update, il-HALO, jl-HALO)
   call check(c_gt_adv_split_v_m(ctx, dt, dbuf(1), dbuf(2), dbuf(3), dbuf(4), dbuf(5), &
              dbuf(6), splitfac, int(scheme, c_int), AH, dmask(1), dmask(2)), 'adv_split_v')
   call slice_to_host(fi, dbuf(2))
   call slice_to_host(Di, dbuf(3))
   call slice_to_host(adv, dbuf(4))
   end subroutine adv_split_v_gpu

   ! adv_split_w (adv_split_w.F90:7-360), I3DFIELD arrays
   subroutine adv_split_w_gpu(dt, f, fi, hi, adv, ww, splitfac, scheme, tag, az, itersmax)
   real(c_double), intent(in) :: dt, splitfac
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km), intent(in) :: f, ww
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km), intent(inout) :: &
        fi, hi, adv
   integer, intent(in) :: scheme, tag, itersmax
   integer, dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(in) :: az
   call gt_to_device(dbuf(1), f)
   call gt_to_device(dbuf(2), fi)
   call gt_to_device(dbuf(3), hi)
   call gt_to_device(dbuf(4), adv)
   call gt_to_device(dbuf(5), ww)
   call az_to_device(dmask(3), az)
   call check(c_gt_adv_split_w_m(ctx, dt, dbuf(1), dbuf(2), dbuf(3), dbuf(4), dbuf(5), &
              splitfac, int(scheme, c_int), int(tag, c_int), dmask(3), &
              int(itersmax, c_int)), 'adv_split_w')
   call gt_to_host(fi, dbuf(2))
   call gt_to_host(hi, dbuf(3))
   call gt_to_host(adv, dbuf(4))
   end subroutine adv_split_w_gpu

   ! adv_upstream_2dh (adv_upstream_2dh.F90:8-169)
   subroutine adv_upstream_2dh_gpu(dt, f, fi, Di, adv, U, V, Dn, DU, DV, AH, az)
   real(c_double), intent(in) :: dt, AH
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(in) :: &
        f, U, V, Dn, DU, DV
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(inout) :: fi, Di, adv
   integer, dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(in) :: az
   call slice_to_device(dbuf(1), f)
   call slice_to_device(dbuf(2), fi)
   call slice_to_device(dbuf(3), Di)
   call slice_to_device(dbuf(4), adv)
   call slice_to_device(dbuf(5), U)
   call slice_to_device(dbuf(6), V)
   call slice_to_device(dbuf(7), Dn)
   call slice_to_device(dbuf(8), DU)
   call slice_to_device(dbuf(9), DV)
   call az_to_device(dmask(3), az)
   call check(c_gt_adv_upstream_2dh_m(ctx, dt, dbuf(1), dbuf(2), dbuf(3), dbuf(4), dbuf(5), &
              dbuf(6), dbuf(7), dbuf(8), dbuf(9), AH, dmask(3)), 'adv_upstream_2dh')
   call slice_to_host(fi, dbuf(2))
   call slice_to_host(Di, dbuf(3))
   call slice_to_host(adv, dbuf(4))
   end subroutine adv_upstream_2dh_gpu

   ! ip_blumberg_mellor (ip_blumberg_mellor.F90:8-152) on the module
   ! variables it reads, passed as host arrays: zz, idpdx, idpdy come back
   subroutine ip_blumberg_mellor_gpu(H, hn, hun, hvn, buoy, zz, idpdx, idpdy)
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(in) :: H
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km), intent(in) :: &
        hn, hun, hvn, buoy
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km), intent(inout) :: &
        zz, idpdx, idpdy
   call slice_to_device(dbuf(1), H)
   call gt_to_device(dbuf(2), hn)
   call gt_to_device(dbuf(3), hun)
   call gt_to_device(dbuf(4), hvn)
   call gt_to_device(dbuf(5), buoy)
   call gt_to_device(dbuf(6), zz)
   call gt_to_device(dbuf(7), idpdx)
   call gt_to_device(dbuf(8), idpdy)
   call check(c_gt_ip_blumberg_mellor(ctx, dbuf(1), dbuf(2), dbuf(3), dbuf(4), dbuf(5), &
              dbuf(6), dbuf(7), dbuf(8)), 'ip_blumberg_mellor()')
   call gt_to_host(zz, dbuf(6))
   call gt_to_host(idpdx, dbuf(7))
   call gt_to_host(idpdy, dbuf(8))
   end subroutine ip_blumberg_mellor_gpu

   ! ip_shchepetkin_mcwilliams (ip_shchepetkin_mcwilliams.F90:8-291)
   subroutine ip_shchepetkin_mcwilliams_gpu(H, hn, hun, hvn, buoy, sseo, zz, idpdx, idpdy)
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO), intent(in) :: H, sseo
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km), intent(in) :: &
        hn, hun, hvn, buoy
   real(c_double), dimension(il-HALO:ih+HALO, jl-HALO:jh+HALO, 0:km), intent(inout) :: &
        zz, idpdx, idpdy
   call slice_to_device(dbuf(1), H)
   call gt_to_device(dbuf(2), hn)
   call gt_to_device(dbuf(3), hun)
   call gt_to_device(dbuf(4), hvn)
   call gt_to_device(dbuf(5), buoy)
   call slice_to_device(dbuf(9), sseo)
   call gt_to_device(dbuf(6), zz)
   call gt_to_device(dbuf(7), idpdx)
   call gt_to_device(dbuf(8), idpdy)
   call check(c_gt_ip_shchepetkin_mcwilliams(ctx, dbuf(1), dbuf(2), dbuf(3), dbuf(4), &
              dbuf(5), dbuf(9), dbuf(6), dbuf(7), dbuf(8)), 'ip_shchepetkin_mcwilliams()')
   call gt_to_host(zz, dbuf(6))
   call gt_to_host(idpdx, dbuf(7))
   call gt_to_host(idpdy, dbuf(8))
   end subroutine ip_shchepetkin_mcwilliams_gpu

   ! update_3d_halo (halo_zones.F90:141-187) on host arrays, serial semantics
   subroutine update_3d_halo(f1, f2, mask, imin, jmin, imax, jmax, kmax, tag, mirror)
   integer, intent(in) :: imin, jmin, imax, jmax, kmax, tag
   integer, intent(in) :: mask(-HALO+1:, -HALO+1:)
   logical, optional, intent(in) :: mirror
   real(c_double), intent(inout), target :: f1(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO, 0:kmax)
   real(c_double), intent(inout), target :: f2(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO, 0:kmax)
   integer(c_int) :: m
   if (.false.) print *, mask(-HALO+1, -HALO+1)
   m = -1
   if (present(mirror)) then
      m = 0
      if (mirror) m = 1
   end if
   if (m /= 1) return
   call gt_to_device(dbuf(1), f1)
   if (c_associated(c_loc(f1), c_loc(f2))) then
      call check(c_gt_update_3d_halo(ctx, dbuf(1), dbuf(1), tag, m), 'update_3d_halo')
   else
      call gt_to_device(dbuf(2), f2)
      call check(c_gt_update_3d_halo(ctx, dbuf(1), dbuf(2), tag, m), 'update_3d_halo')
   end if
   call gt_to_host(f1, dbuf(1))
   end subroutine update_3d_halo

   end module getm_transport_gpu
