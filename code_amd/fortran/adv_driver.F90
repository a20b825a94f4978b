! adv_driver.F90 -- a minimal Fortran caller of getm_transport_gpu, used by the
! Fortran parity test: it reads a case written by the test (stream I/O),
! advances it with do_advection_3d exactly as temperature.F90:398-405 calls
! it, and writes f and hires back.
!   usage: adv_driver <input file> <output file>
#define HALO 2
program adv_driver
   use, intrinsic :: iso_c_binding
   use getm_transport_gpu
   implicit none
   integer(c_int) :: hdr(10)
   real(c_double) :: rhdr(4)
   integer :: imin, imax, jmin, jmax, kmax, nsteps, split, hscheme, vscheme, tag, n
   integer(c_int), allocatable :: az(:,:), au(:,:), av(:,:), ax(:,:)
   real(c_double), allocatable, dimension(:,:,:) :: f, uu, vv, ww, hu, hv, ho, hn, hires
   character(len=1024) :: fin, fout
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(10, file=trim(fin), access='stream', form='unformatted', status='old')
   read(10) hdr
   read(10) rhdr
   imin = hdr(1); imax = hdr(2); jmin = hdr(3); jmax = hdr(4); kmax = hdr(5)
   nsteps = hdr(6); split = hdr(7); hscheme = hdr(8); vscheme = hdr(9); tag = hdr(10)
   allocate(az(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO))
   allocate(au, av, ax, mold=az)
   allocate(f(imin-HALO:imax+HALO, jmin-HALO:jmax+HALO, 0:kmax))
   allocate(uu, vv, ww, hu, hv, ho, hn, hires, mold=f)
   read(10) az, au, av, ax
   read(10) f, uu, vv, ww, hu, hv, ho, hn
   close(10)
   call gt_init(imin, imax, jmin, jmax, kmax, rhdr(1), rhdr(2), az, au, av, ax)
   do n = 1, nsteps
      call do_advection_3d(rhdr(3), f, uu, vv, ww, hu, hv, ho, hn, split, hscheme, &
                           vscheme, rhdr(4), tag, hires=hires)
   end do
   call gt_finalize()
   open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(11) f, hires
   close(11)
end program adv_driver
