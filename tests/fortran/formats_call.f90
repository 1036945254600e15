! Writes and reads the files the Fortran steps exchange, with the record
! statements of those steps (hvccfp0's __Vcoud, hsfp0's SEC2U, WVIllwI's
! fixed-length W records), for the library's readers and writers to be
! checked against.  Usage: formats_call <mode> <file> <stream-io-file>
!   vc_w / vc_r   __Vcoud: ngb; q(3); vcoud(ngb); zcousq(ngb,ngb)
!   s2_w / s2_r   SEC2U: nspin nq ntq nqbz nqibz n1 n2 n3; per q: is, q(3), zsec(ntq,ntq)
!   wv_w / wv_r   direct access, recl = 16*ld*ld bytes, records 1..nrec
program formats_call
  implicit none
  character(len=8) :: mode
  character(len=512) :: fname, sname
  integer :: u, s, ngb, h8(8), nq, ntq, ip, ld, nrec, r
  real(8) :: q(3)
  real(8), allocatable :: v(:), qq(:,:)
  integer, allocatable :: isp(:)
  complex(8), allocatable :: z(:,:), zs(:,:,:), w(:,:,:)
  call get_command_argument(1, mode)
  call get_command_argument(2, fname)
  call get_command_argument(3, sname)
  select case (trim(mode))
  case ('vc_w')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) ngb
    allocate(v(ngb), z(ngb, ngb))
    read(s) q, v, z
    close(s)
    open(newunit=u, file=trim(fname), form='unformatted')
    write(u) ngb
    write(u) q
    write(u) v
    write(u) z
    close(u)
  case ('vc_r')
    open(newunit=u, file=trim(fname), form='unformatted', status='old')
    read(u) ngb
    allocate(v(ngb), z(ngb, ngb))
    read(u) q
    read(u) v
    read(u) z
    close(u)
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='replace')
    write(s) ngb, q, v, z
    close(s)
  case ('s2_w')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) h8
    nq = h8(2); ntq = h8(3)
    allocate(isp(nq), qq(3, nq), zs(ntq, ntq, nq))
    read(s) isp, qq, zs
    close(s)
    open(newunit=u, file=trim(fname), form='unformatted')
    write(u) h8
    do ip = 1, nq
      write(u) isp(ip), qq(1:3, ip), zs(1:ntq, 1:ntq, ip)
    enddo
    close(u)
  case ('s2_r')
    open(newunit=u, file=trim(fname), form='unformatted', status='old')
    read(u) h8
    nq = h8(2); ntq = h8(3)
    allocate(isp(nq), qq(3, nq), zs(ntq, ntq, nq))
    do ip = 1, nq
      read(u) isp(ip), qq(1:3, ip), zs(1:ntq, 1:ntq, ip)
    enddo
    close(u)
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='replace')
    write(s) h8, isp, qq, zs
    close(s)
  case ('wv_w')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) ld, nrec
    allocate(w(ld, ld, nrec))
    read(s) w
    close(s)
    open(newunit=u, file=trim(fname), form='unformatted', access='direct', recl=16*ld*ld)
    do r = 1, nrec
      write(u, rec=r) w(:, :, r)
    enddo
    close(u)
  case ('wv_r')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) ld, nrec
    close(s)
    allocate(w(ld, ld, nrec))
    open(newunit=u, file=trim(fname), form='unformatted', access='direct', recl=16*ld*ld, status='old')
    do r = 1, nrec
      read(u, rec=r) w(:, :, r)
    enddo
    close(u)
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='replace')
    write(s) w
    close(s)
  end select
end program formats_call
