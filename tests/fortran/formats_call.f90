! Writes and reads the files the Fortran steps exchange, with the record
! statements of those steps (hvccfp0's __Vcoud, hsfp0's SEC2U, WVIllwI's
! fixed-length W records), for the library's readers and writers to be
! checked against.  Usage: formats_call <mode> <file> <stream-io-file>
!   vc_w / vc_r   __Vcoud: ngb; q(3); vcoud(ngb); zcousq(ngb,ngb)
!   s2_w / s2_r   SEC2U: nspin nq ntq nqbz nqibz n1 n2 n3; per q: is, q(3), zsec(ntq,ntq)
!   wv_w / wv_r   direct access, recl = 16*ld*ld bytes, records 1..nrec
!   w4_w / w4_r   the same in complex(4) (the __MP build), recl = 8*ld*ld bytes
!   qg_w          __QGpsi (mkqg): header; per q (q, ngp, irr), (ngvecp, ngvecprev)
!   ev_w          __EValue (sugw): (nband, nqi, nsp, nspc), qirr(3,nqi), evl(nband,nqi,nspx)
!   cp_w          __CPHI-like direct access: recl bytes, nrec records of n complex(8)
!   mn_w          @MNLA_CPHI (sugw): header line, then m n l ibas iord iorb per row
program formats_call
  implicit none
  character(len=8) :: mode
  character(len=512) :: fname, sname
  integer :: u, s, ngb, h8(8), nq, ntq, ip, ld, nrec, r
  integer :: h7(6), ngp, irr, imx, nqi, nband, nsp, nspc, nspx, recl, n, nrow, mx, nx, lx, ibas, iord, iorb
  real(8) :: cut
  integer, allocatable :: gv(:,:), gr(:,:,:)
  real(8), allocatable :: ev(:,:,:)
  complex(8), allocatable :: c1(:)
  real(8) :: q(3)
  real(8), allocatable :: v(:), qq(:,:)
  integer, allocatable :: isp(:)
  complex(8), allocatable :: z(:,:), zs(:,:,:), w(:,:,:)
  complex(4), allocatable :: w4(:,:,:)
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
  case ('w4_w')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) ld, nrec
    allocate(w4(ld, ld, nrec))
    read(s) w4
    close(s)
    open(newunit=u, file=trim(fname), form='unformatted', access='direct', recl=8*ld*ld)
    do r = 1, nrec
      write(u, rec=r) w4(:, :, r)
    enddo
    close(u)
  case ('w4_r')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) ld, nrec
    close(s)
    allocate(w4(ld, ld, nrec))
    open(newunit=u, file=trim(fname), form='unformatted', access='direct', recl=8*ld*ld, status='old')
    do r = 1, nrec
      read(u, rec=r) w4(:, :, r)
    enddo
    close(u)
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='replace')
    write(s) w4
    close(s)
  case ('qg_w')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) h7(1:2), cut, h7(3:6)
    nq = h7(1); imx = h7(5)
    allocate(gr(-imx:imx, -imx:imx, -imx:imx))
    open(newunit=u, file=trim(fname), form='unformatted')
    write(u) h7(1), h7(2), cut, h7(3), h7(4), h7(5), h7(6)
    do ip = 1, nq
      read(s) q, ngp, irr
      allocate(gv(3, max(ngp, 1)))
      read(s) gv, gr
      write(u) q, ngp, irr
      write(u) gv, gr
      deallocate(gv)
    enddo
    close(u)
    close(s)
  case ('ev_w')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) nband, nqi, nsp, nspc, nspx
    allocate(qq(3, nqi), ev(nband, nqi, nspx))
    read(s) qq, ev
    close(s)
    open(newunit=u, file=trim(fname), form='unformatted')
    write(u) nband, nqi, nsp, nspc
    write(u) qq(1:3, 1:nqi)
    write(u) ev(1:nband, 1:nqi, 1:nspx)
    close(u)
  case ('cp_w')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) recl, nrec, n
    allocate(c1(n))
    open(newunit=u, file=trim(fname), form='unformatted', access='direct', recl=recl)
    do r = 1, nrec
      read(s) c1
      write(u, rec=r) c1
    enddo
    close(u)
    close(s)
  case ('mn_w')
    open(newunit=s, file=trim(sname), access='stream', form='unformatted', status='old')
    read(s) nrow
    open(newunit=u, file=trim(fname))
    write(u, "('    m    n    l ibas')")
    do r = 1, nrow
      read(s) mx, nx, lx, ibas, iord, iorb
      write(u, "(10i6)") mx, nx, lx, ibas, iord, iorb
    enddo
    close(u)
    close(s)
  end select
end program formats_call
