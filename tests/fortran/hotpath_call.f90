! The hot path of one q-point driven from Fortran through the ISO_C_BINDING
! module, with the calls written as a driver makes them: per k zmel (conj)
! and x0gemm as in x0kf_zxq (S/x0kf_v4h.f90:272-282), dpsion_chiq on the
! imaginary axis (:317-319), the W build of WVIllwI (S/m_llw.f90:309-373),
! then per internal k zmel (no conj), the imaginary-axis weights and the
! Sigma_c contraction of get_correlation_block (S/m_sxcf_gemm.f90:423-608).
! Inputs from the file named first, results (chi, W_c, Sigma_c) to the second.
program hotpath_call
  use, intrinsic :: iso_c_binding
  use ecalj_gpu
  implicit none
  type(c_ptr) :: h
  integer(c_int) :: rc, prec, natom, ndima, ngb, nk, no, nu, ntq, nbt, nwhis, nw1, ncoun, ncount, ppbsize, nf, k
  integer(c_int), allocatable :: nv(:), mdim(:), icounkmin(:), icounkmax(:)
  integer(c_int), allocatable :: iwini(:), iwend(:), itc(:), itpc(:), jpmc(:), icouini(:)
  real(c_double), allocatable :: ppb(:), vcoud(:), whwc(:), frhis(:), freqi(:), omega(:), ekc(:, :), wk(:)
  real(c_double), allocatable :: freqx(:), wx(:), expa(:), wgtim(:, :, :)
  real(c_double) :: esmr, ua
  complex(c_double_complex), allocatable :: zcousq(:, :), phase(:), c_occ(:, :, :), c_unocc(:, :, :)
  complex(c_double_complex), allocatable :: c_qmk(:, :, :), c_q(:, :), zmel(:, :, :), zmels(:, :, :)
  complex(c_double_complex), allocatable :: rcxq(:, :, :), zxqi(:, :, :), wc(:, :, :), zsec(:, :)
  character(len=512) :: fin, fout
  integer :: u
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) prec, natom, ndima, ngb, nk, no, nu, ntq, nbt, nwhis, nw1, ncoun, ncount, ppbsize
  nf = nw1 - 1
  allocate(nv(natom), mdim(natom), ppb(ppbsize), vcoud(ngb), zcousq(ngb, ngb), phase(natom))
  allocate(c_occ(ndima, no, nk), c_unocc(ndima, nu, nk), icounkmin(nk), icounkmax(nk))
  allocate(iwini(ncoun), iwend(ncoun), itc(ncoun), itpc(ncoun), jpmc(ncoun), icouini(ncoun), whwc(ncount))
  allocate(frhis(nwhis + 1), freqi(nw1), c_qmk(ndima, nbt, nk), c_q(ndima, ntq), omega(ntq), ekc(nbt, nk), wk(nk))
  allocate(freqx(nf), wx(nf), expa(nf))
  read(u) nv, mdim, ppb, vcoud, zcousq, phase, c_occ, c_unocc, icounkmin, icounkmax
  read(u) iwini, iwend, itc, itpc, jpmc, icouini, whwc, frhis, freqi, c_qmk, c_q, omega, ekc, wk
  read(u) freqx, wx, expa, esmr, ua
  close(u)
  allocate(zmel(ngb, no, nu), rcxq(ngb, ngb, 0:nwhis), zxqi(ngb, ngb, nw1), wc(ngb, ngb, nw1))
  allocate(zmels(ngb, nbt, ntq), wgtim(0:nf, nbt, ntq), zsec(ntq, ntq))
  rc = ecalj_create(h, 0)
  if (rc /= 0) stop 2
  rc = ecalj_set_precision(h, prec)
  if (rc /= 0) stop 3
  rc = ecalj_set_ppb(h, natom, nv, mdim, ppb)
  if (rc /= 0) stop 4
  rc = ecalj_set_coulomb(h, ngb, vcoud, zcousq)
  if (rc /= 0) stop 5
  rcxq = (0d0, 0d0)
  do k = 1, nk   ! x0kf_zxq: zmel of occupied (k) x unoccupied (k+q), conjugated, then x0gemm
    rc = ecalj_zmel(h, phase, ndima, no, c_occ(:, :, k), nu, c_unocc(:, :, k), 1_c_int, 1_c_int, 1_c_int, zmel)
    if (rc /= 0) stop 6
    rc = ecalj_x0gemm(h, rcxq, ngb, 1_c_int, ngb, nwhis, 1_c_int, 1_c_int, no, zmel, ngb, nu, &
                      icounkmin(k), icounkmax(k), iwini, iwend, itc, itpc, jpmc, icouini, whwc)
    if (rc /= 0) stop 7
  enddo
  rc = ecalj_dpsion_imag(h, rcxq, ngb, ngb, nwhis, nw1, frhis, freqi, 1d10, 1_c_int, zxqi)
  if (rc /= 0) stop 8
  rc = ecalj_wc_build(h, zxqi, ngb, nw1, 1_c_int, 1_c_int, 1_c_int, wc, ngb)
  if (rc /= 0) stop 9
  zsec = (0d0, 0d0)
  do k = 1, nk   ! get_correlation_block for internal k: zmel (M basis), weights, contraction
    rc = ecalj_zmel(h, phase, ndima, nbt, c_qmk(:, :, k), ntq, c_q, 1_c_int, 0_c_int, 1_c_int, zmels)
    if (rc /= 0) stop 10
    rc = ecalj_sigc_weights(h, nf, 1_c_int, nbt, ntq, 0_c_int, omega, ekc(:, k), esmr, ua, wk(k), freqx, wx, &
                            expa, wgtim)
    if (rc /= 0) stop 11
    rc = ecalj_sigc(h, zmels, ngb, nbt, ntq, nf, wc, ngb, wgtim, zsec, ntq)
    if (rc /= 0) stop 12
  enddo
  rc = ecalj_destroy(h)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) zxqi, wc, zsec
  close(u)
end program hotpath_call
