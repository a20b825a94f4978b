! rows_driver.F90 -- a Fortran caller of the device-resident routines either
! side of the transport step (getm_transport_gpu), used by the Fortran parity
! test: it reads T, S, hn, nuh written by the test (stream I/O), moves them to
! the device, runs do_eqstate (eqstate_method 3, eqstate.F90:94-271) and the
! salinity diffusion of do_salinity (salinity.F90:346-393), and writes rho,
! buoy, alpha, beta and the diffused S back; between the two it runs the
! fused equation of state + pressure routines (Jackett et al. and linear,
! both methods) on T, S and the rho do_eqstate left, and writes their zz and
! gradients.
!   usage: rows_driver <input file> <output file>
#define HALO 2
program rows_driver
   use, intrinsic :: iso_c_binding
   use getm_transport_gpu
   implicit none
   integer(c_int) :: hdr(5)
   real(c_double) :: rhdr(4)
   integer :: imin, imax, jmin, jmax, kmax, q
   integer(c_int), allocatable :: az(:,:), au(:,:), av(:,:), ax(:,:)
   real(c_double), allocatable, dimension(:,:,:) :: T, S, hn, nuh, rho, buoy, alpha, beta
   real(c_double), allocatable, dimension(:,:,:) :: hun, hvn, zero3, o3
   real(c_double), allocatable, dimension(:,:) :: H, sseo
   type(c_ptr) :: d(24)
   integer(c_size_t) :: nbytes, nbytes2
   real(c_double) :: T0, S0, dtr0, dsr0
   character(len=1024) :: fin, fout
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(10, file=trim(fin), access='stream', form='unformatted', status='old')
   read(10) hdr
   read(10) rhdr
   imin = hdr(1); imax = hdr(2); jmin = hdr(3); jmax = hdr(4); kmax = hdr(5)
   allocate(az(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO))
   allocate(au, av, ax, mold=az)
   allocate(T(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO, 0:kmax))
   allocate(S, hn, nuh, rho, buoy, alpha, beta, hun, hvn, zero3, o3, mold=T)
   allocate(H(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO))
   allocate(sseo, mold=H)
   read(10) az, au, av, ax
   read(10) T, S, hn, nuh
   read(10) hun, hvn, H, sseo
   close(10)
   zero3 = 0.0d0
   call gt_init(imin, imax, jmin, jmax, kmax, rhdr(1), rhdr(2), az, au, av, ax)
   nbytes = 8_c_size_t * size(T, kind=c_size_t)
   nbytes2 = 8_c_size_t * size(H, kind=c_size_t)
   do q = 1, 24
      if (q == 9 .or. q == 10) then
         call gt_malloc_f(d(q), nbytes2)
      else
         call gt_malloc_f(d(q), nbytes)
      end if
   end do
   call slice_to_device(d(9), H)
   call slice_to_device(d(10), sseo)
   call gt_to_device(d(11), hun)
   call gt_to_device(d(12), hvn)
   do q = 13, 24
      call gt_to_device(d(q), zero3)
   end do
   rho = 0.0d0
   alpha = 0.0d0
   call gt_to_device(d(1), T)
   call gt_to_device(d(2), S)
   call gt_to_device(d(3), hn)
   call gt_to_device(d(4), nuh)
   call gt_to_device(d(5), rho)
   call gt_to_device(d(6), rho)
   call gt_to_device(d(7), alpha)
   call gt_to_device(d(8), alpha)
   ! the namelist defaults of eqstate.F90:26-27 (default REALs)
   T0 = real(10., c_double)
   S0 = real(33.75, c_double)
   dtr0 = real(-0.17, c_double)
   dsr0 = real(0.78, c_double)
   call do_eqstate_gpu(3, d(1), d(2), d(3), d(5), d(6), d(7), d(8), T0, S0, dtr0, dsr0)
   ! Jackett et al. and the linear EOS fused into both pressure methods
   call eqstate_internal_pressure_gpu(3, 3, d(1), d(2), d(5), T0, S0, dtr0, dsr0, d(9), &
                                      d(3), d(11), d(12), d(10), d(13), d(14), d(15), &
                                      d(16), d(17), d(18))
   call eos_internal_pressure_gpu(3, d(1), d(2), T0, S0, dtr0, dsr0, d(9), d(3), d(11), &
                                  d(12), d(10), d(19), d(20), d(21), d(22), d(23), d(24))
   call salinity_diffusion_gpu(rhdr(3), rhdr(4), d(3), d(4), d(2))
   call gt_to_host(rho, d(5))
   call gt_to_host(buoy, d(6))
   call gt_to_host(alpha, d(7))
   call gt_to_host(beta, d(8))
   call gt_to_host(S, d(2))
   open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(11) rho, buoy, alpha, beta, S
   do q = 13, 24
      call gt_to_host(o3, d(q))
      write(11) o3
   end do
   close(11)
   do q = 1, 24
      call gt_free_f(d(q))
   end do
   call gt_finalize()
end program rows_driver
