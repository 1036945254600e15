! Calls the library's tetrahedron weights (gettetwt's replacement) through
! the ISO_C_BINDING module as a Fortran driver would: once for the sizes
! (capacities 0), then with arrays of those sizes; the options npm, the core
! states and tetrahedron_matrix_linear come from the input file.  Writes the
! lists and weights back.
program tetwt_call
  use, intrinsic :: iso_c_binding
  use ecalj_gpu
  implicit none
  type(c_ptr) :: h
  integer(c_int) :: rc, nband, nqbz, ntetf, nqbzw, nwhis, npm, nctot, ml, nbnbx
  integer(c_int64_t) :: nhwtot
  integer(c_int), allocatable :: idtetf(:, :), ib1bz(:), nbnb(:, :)
  integer(c_int), allocatable :: n1b(:, :, :), n2b(:, :, :), ihw(:, :, :), nhw(:, :, :), jhw(:, :, :)
  real(c_double), allocatable :: qbzw(:, :), ek1(:, :), ek2(:, :), frhis(:), ecore(:), whw(:)
  real(c_double) :: ef, ebmx, qlat(3, 3)
  integer(c_int) :: idum(1)
  real(c_double) :: ddum(1)
  character(len=512) :: fin, fout
  integer :: u
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) nband, nqbz, ntetf, nqbzw, nwhis, npm, nctot, ml
  read(u) ef, ebmx
  allocate(idtetf(4, ntetf), ib1bz(nqbzw), qbzw(3, nqbzw), ek1(nband, nqbz), ek2(nband, nqbz), frhis(nwhis + 1))
  allocate(ecore(max(nctot, 1)))
  read(u) idtetf, qbzw, ib1bz, ek1, ek2, qlat, frhis
  if (nctot > 0) read(u) ecore(1:nctot)
  close(u)
  rc = ecalj_create(h, 0)
  if (rc /= 0) stop 2
  nbnbx = 0
  nhwtot = 0
  rc = ecalj_tetwt_full(h, nband, nqbz, ntetf, idtetf, nqbzw, qbzw, ib1bz, ek1, ek2, ef, ebmx, qlat, frhis, nwhis, &
                        npm, nctot, ecore, ml, 0_c_int, 0_c_int64_t, idum, idum, idum, idum, idum, idum, ddum, &
                        nbnbx, nhwtot)
  if (nbnbx < 1 .or. nhwtot < 1) stop 3
  allocate(nbnb(nqbz, npm), n1b(nbnbx, nqbz, npm), n2b(nbnbx, nqbz, npm), ihw(nbnbx, nqbz, npm))
  allocate(nhw(nbnbx, nqbz, npm), jhw(nbnbx, nqbz, npm), whw(nhwtot))
  rc = ecalj_tetwt_full(h, nband, nqbz, ntetf, idtetf, nqbzw, qbzw, ib1bz, ek1, ek2, ef, ebmx, qlat, frhis, nwhis, &
                        npm, nctot, ecore, ml, nbnbx, nhwtot, nbnb, n1b, n2b, ihw, nhw, jhw, whw, nbnbx, nhwtot)
  if (rc /= 0) stop 4
  rc = ecalj_destroy(h)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) nbnbx, nhwtot
  write(u) nbnb, n1b, n2b, ihw, nhw, jhw, whw
  close(u)
end program tetwt_call
