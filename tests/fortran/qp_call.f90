! Calls the host-only QP update of libecalj_gpu.so through the ISO_C_BINDING
! module, as hqpe_sc would, and prints the results for the Python test.
program qp_call
  use, intrinsic :: iso_c_binding
  use ecalj_gpu
  implicit none
  integer, parameter :: n = 3
  real(c_double) :: elda(n) = [-1.5d0, 0.2d0, 3.1d0], sex(n) = [-0.5d0, -0.3d0, -0.1d0]
  real(c_double) :: sexcore(n) = [-0.01d0, -0.02d0, 0.0d0], vxc(n) = [-12.d0, -10.d0, -9.d0], eqp(n)
  complex(c_double_complex) :: sec(n, n)
  integer(c_int) :: rc, i
  sec = (0d0, 0d0)
  do i = 1, n
    sec(i, i) = cmplx(-0.1d0 * i, -0.01d0 * i, kind=c_double_complex)
  enddo
  rc = ecalj_qp_energy(n, elda, sex, sexcore, sec, n, vxc, eqp)
  print '(i2,3es24.16)', rc, eqp
end program qp_call
