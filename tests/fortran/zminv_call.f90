! Calls the library's zminv and real-axis W build through the ISO_C_BINDING
! module, as a Fortran driver would (m_llw's zminv alias, WVRllwR), on inputs
! read from the file named on the command line; writes the results back.
program zminv_call
  use, intrinsic :: iso_c_binding
  use ecalj_gpu
  implicit none
  type(c_ptr) :: h
  integer(c_int) :: rc, n, ngb, nw
  complex(c_double_complex), allocatable :: a(:, :), chi(:, :, :), wc(:, :, :), zc(:, :)
  real(c_double), allocatable :: vcoud(:)
  character(len=512) :: fin, fout
  integer :: u
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, ngb, nw
  allocate(a(n, n), chi(ngb, ngb, nw), wc(ngb, ngb, nw), vcoud(ngb), zc(ngb, ngb))
  read(u) a, chi, vcoud, zc
  close(u)
  rc = ecalj_create(h, 0)
  if (rc /= 0) stop 2
  rc = ecalj_zminv(h, a, n, n)
  if (rc /= 0) stop 3
  rc = ecalj_set_coulomb(h, ngb, vcoud, zc)
  if (rc /= 0) stop 4
  rc = ecalj_wc_build_real(h, chi, ngb, nw, 1, 1, 1, wc, ngb)
  if (rc /= 0) stop 5
  rc = ecalj_destroy(h)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) a, wc
  close(u)
end program zminv_call
