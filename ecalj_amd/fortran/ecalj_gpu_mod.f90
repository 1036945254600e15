!> ISO_C_BINDING interfaces to libecalj_gpu.so (include/ecalj_gpu.h).
!! The argument layout follows the reference routines each entry point
!! replaces; Fortran arrays are passed by address (column-major, complex(8)
!! interleaved), scalars by value.  Every function returns an integer
!! status; 0 is success (see ecalj_status_string).
module ecalj_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: ecalj_create, ecalj_destroy, ecalj_set_precision, ecalj_sync, ecalj_stream_wait
  public :: ecalj_set_ppb, ecalj_set_coulomb, ecalj_zmel, ecalj_x0gemm, ecalj_dpsion_imag
  public :: ecalj_wc_build, ecalj_sigc_weights, ecalj_sigc, ecalj_qp_energy, ecalj_sigma_herm
  public :: ecalj_zmel_dev, ecalj_x0gemm_batched_dev, ecalj_dpsion_imag_dev, ecalj_wc_build_dev
  public :: ecalj_sigc_weights_dev, ecalj_sigc_dev, ecalj_sigc_grouped_dev
  public :: ecalj_set_tuning, ecalj_graph_begin, ecalj_graph_end, ecalj_graph_launch, ecalj_graph_destroy
  public :: ecalj_dpsion_real, ecalj_dpsion_real_dev, ecalj_wc_build_real, ecalj_wc_build_real_dev
  public :: ecalj_zhgv, ecalj_zhgv_dev
  public :: ecalj_set_ppb_full, ecalj_zmel_init, ecalj_zmel_init_dev
  public :: ecalj_zminv, ecalj_zminv_dev, ecalj_sigc_real, ecalj_sigc_real_dev, ecalj_sigc_full_dev
  public :: ecalj_tetwt, ecalj_tetwt_full, ecalj_set_ppovl, ecalj_zmel_ipw, ecalj_zmel_ipw_dev, ecalj_zmel_assemble_dev
  public :: ecalj_wv_write_dev, ecalj_wv_read_dev, ecalj_wv4_write_dev, ecalj_wv4_read_dev
  public :: ecalj_wc_build_gamma, ecalj_wc_build_gamma_dev, ecalj_eps_head, ecalj_eps_head_dev, ecalj_getw0
  public :: ecalj_wv0_patch, ecalj_wv0_patch_dev
  public :: ecalj_set_hamindex, ecalj_set_mnla, ecalj_readcphi_dev, ecalj_readgeig_dev, ecalj_mnla_read
  public :: ecalj_qgpsi_read, ecalj_evalue_read, ecalj_rec_read_dev
  integer(c_int), parameter, public :: ECALJ_PREC_F64 = 0, ECALJ_PREC_F32 = 1, ECALJ_PREC_BF16X3 = 2, &
                                        ECALJ_PREC_BF16X3_F8 = 3

  interface
    ! gpu_init (S/m_gpu.f90:14-68)
    integer(c_int) function ecalj_create(h, device) bind(C, name="ecalj_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: device
    end function
    integer(c_int) function ecalj_destroy(h) bind(C, name="ecalj_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function ecalj_set_precision(h, prec) bind(C, name="ecalj_set_precision")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: prec
    end function
    integer(c_int) function ecalj_sync(h) bind(C, name="ecalj_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! no counterpart in the reference: h's stream waits for the work issued on other's
    integer(c_int) function ecalj_stream_wait(h, other) bind(C, name="ecalj_stream_wait")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, other
    end function
    ! set_ppb (S/m_zmel.f90:110-145)
    integer(c_int) function ecalj_set_ppb(h, natom, nv, mdim, ppb) bind(C, name="ecalj_set_ppb")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: natom
      integer(c_int), intent(in) :: nv(*), mdim(*)
      real(c_double), intent(in) :: ppb(*)
    end function
    ! Readvcoud + set_m2e_prod_basis (S/m_readVcoud.f90:11-43)
    integer(c_int) function ecalj_set_coulomb(h, ngb, vcoud, zcousq) bind(C, name="ecalj_set_coulomb")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      integer(c_int), value :: ngb
      real(c_double), intent(in) :: vcoud(*)
      complex(c_double_complex), intent(in) :: zcousq(*)
    end function
    ! get_zmel_init_gemm (S/m_zmel.f90:146-597)
    integer(c_int) function ecalj_zmel(h, phase, ndima, nm, cphim, nq, cphiq, is_m_basis, zmelconjg, nbatch, zmel) &
        bind(C, name="ecalj_zmel")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: phase(*)
      integer(c_int), value :: ndima, nm, nq, is_m_basis, zmelconjg, nbatch
      complex(c_double_complex), intent(in) :: cphim(*), cphiq(*)
      complex(c_double_complex), intent(out) :: zmel(*)
    end function
    integer(c_int) function ecalj_zmel_dev(h, phase, ndima, nm, cphim, nq, cphiq, is_m_basis, zmelconjg, nbatch, &
        zmel) bind(C, name="ecalj_zmel_dev")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: phase(*)
      integer(c_int), value :: ndima, nm, nq, is_m_basis, zmelconjg, nbatch
      type(c_ptr), value :: cphim, cphiq, zmel
    end function
    ! x0gemm (S/x0gemm.f90:2-97); module inputs of m_x0kf/m_zmel made explicit
    integer(c_int) function ecalj_x0gemm(h, rcxq, npr, ipr_col, npr_col, nwhis, npm, ns1, ns2, zmel, nbb, nqtot, &
        icoun_lo, icoun_hi, iwini, iwend, itc, itpc, jpmc, icouini, whwc) bind(C, name="ecalj_x0gemm")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(inout) :: rcxq(*)
      integer(c_int), value :: npr, ipr_col, npr_col, nwhis, npm, ns1, ns2, nbb, nqtot, icoun_lo, icoun_hi
      complex(c_double_complex), intent(in) :: zmel(*)
      integer(c_int), intent(in) :: iwini(*), iwend(*), itc(*), itpc(*), jpmc(*), icouini(*)
      real(c_double), intent(in) :: whwc(*)
    end function
    integer(c_int) function ecalj_x0gemm_batched_dev(h, rcxq, npr, ipr_col, npr_col, nwhis, npm, nk, ns1, ns2, &
        zmel, zmel_stride, nbb, nqtot, icounkmin, icounkmax, ncoun, iwini, iwend, itc, itpc, jpmc, icouini, ncount, &
        whwc) bind(C, name="ecalj_x0gemm_batched_dev")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h, rcxq, zmel
      integer(c_int), value :: npr, ipr_col, npr_col, nwhis, npm, nk, nbb, nqtot, ncoun, ncount
      integer(c_int64_t), value :: zmel_stride
      integer(c_int), intent(in) :: ns1(*), ns2(*), icounkmin(*), icounkmax(*)
      integer(c_int), intent(in) :: iwini(*), iwend(*), itc(*), itpc(*), jpmc(*), icouini(*)
      real(c_double), intent(in) :: whwc(*)
    end function
    ! dpsion_chiq imaginary axis (S/dpsion5.f90:211-232)
    integer(c_int) function ecalj_dpsion_imag(h, rcxq, npr, npr_col, nwhis, niw, frhis, freqi, ecut, scale_rcxq, &
        zxqi) bind(C, name="ecalj_dpsion_imag")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(inout) :: rcxq(*)
      integer(c_int), value :: npr, npr_col, nwhis, niw, scale_rcxq
      real(c_double), intent(in) :: frhis(*), freqi(*)
      real(c_double), value :: ecut
      complex(c_double_complex), intent(out) :: zxqi(*)
    end function
    integer(c_int) function ecalj_dpsion_imag_dev(h, rcxq, npr, npr_col, nwhis, niw, frhis, freqi, ecut, &
        scale_rcxq, zxqi) bind(C, name="ecalj_dpsion_imag_dev")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h, rcxq, zxqi
      integer(c_int), value :: npr, npr_col, nwhis, niw, scale_rcxq
      real(c_double), intent(in) :: frhis(*), freqi(*)
      real(c_double), value :: ecut
    end function
    ! dpsion_chiq real branch, npm=1 (S/dpsion5.f90:193-200,244-256); rcxq read only, zxq(:,:,0:nw) returned
    integer(c_int) function ecalj_dpsion_real(h, rcxq, npr, npr_col, nwhis, nw, frhis, freqr, ecut, rcxq_scaled, &
        zxq) bind(C, name="ecalj_dpsion_real")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: rcxq(*)
      integer(c_int), value :: npr, npr_col, nwhis, nw, rcxq_scaled
      real(c_double), intent(in) :: frhis(*), freqr(*)
      real(c_double), value :: ecut
      complex(c_double_complex), intent(out) :: zxq(*)
    end function
    integer(c_int) function ecalj_dpsion_real_dev(h, rcxq, npr, npr_col, nwhis, nw, frhis, freqr, ecut, &
        rcxq_scaled, zxq) bind(C, name="ecalj_dpsion_real_dev")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h, rcxq, zxq
      integer(c_int), value :: npr, npr_col, nwhis, nw, rcxq_scaled
      real(c_double), intent(in) :: frhis(*), freqr(*)
      real(c_double), value :: ecut
    end function
    ! WVRllwR body, q /= 0 (S/m_llw.f90:88-180): pivoted inversion
    integer(c_int) function ecalj_wc_build_real(h, chi, ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, wc, ldw) &
        bind(C, name="ecalj_wc_build_real")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: chi(*)
      integer(c_int), value :: ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, ldw
      complex(c_double_complex), intent(inout) :: wc(*)
    end function
    integer(c_int) function ecalj_wc_build_real_dev(h, chi, ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, wc, ldw) &
        bind(C, name="ecalj_wc_build_real_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, chi, wc
      integer(c_int), value :: ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, ldw
    end function
    ! zminv_h / zminv_d (S/m_lapack.f90:21-47,108-176)
    integer(c_int) function ecalj_zminv(h, a, n, lda) bind(C, name="ecalj_zminv")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(inout) :: a(*)
      integer(c_int), value :: n, lda
    end function
    integer(c_int) function ecalj_zminv_dev(h, a, n, lda, nbatch, stride) bind(C, name="ecalj_zminv_dev")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h, a
      integer(c_int), value :: n, lda, nbatch
      integer(c_int64_t), value :: stride
    end function
    ! set_ppb from the reference's ppb(nlnmx,nlnmx,mdimx,natom) (S/m_zmel.f90:38,110-145)
    integer(c_int) function ecalj_set_ppb_full(h, natom, nlnmx, mdimx, nlnmc, nlnmv, nblocha, ppb) &
        bind(C, name="ecalj_set_ppb_full")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: natom, nlnmx, mdimx
      integer(c_int), intent(in) :: nlnmc(*), nlnmv(*), nblocha(*)
      real(c_double), intent(in) :: ppb(*)
    end function
    ! get_zmel_init_gemm, MT part with core states and iatomp (S/m_zmel.f90:146-147,323-391)
    integer(c_int) function ecalj_zmel_init(h, phase, iatomp, ndima, ns1, ns2, nctot, cphim, nq, cphiq, &
        is_m_basis, zmelconjg, nbatch, zmel) bind(C, name="ecalj_zmel_init")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h, phase, iatomp
      integer(c_int), value :: ndima, ns1, ns2, nctot, nq, is_m_basis, zmelconjg, nbatch
      complex(c_double_complex), intent(in) :: cphim(*), cphiq(*)
      complex(c_double_complex), intent(out) :: zmel(*)
    end function
    integer(c_int) function ecalj_zmel_init_dev(h, phase, iatomp, ndima, ns1, ns2, nctot, cphim, nq, cphiq, &
        is_m_basis, zmelconjg, nbatch, zmel) bind(C, name="ecalj_zmel_init_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, phase, iatomp, cphim, cphiq, zmel
      integer(c_int), value :: ndima, ns1, ns2, nctot, nq, is_m_basis, zmelconjg, nbatch
    end function
    ! zhgv_h / zhgv_d (S/m_lapack.f90:64-104,203-229): a z = lambda b z, eigenpairs il..iu
    integer(c_int) function ecalj_zhgv(h, a, b, n, evl, il, iu, lda, ldb) bind(C, name="ecalj_zhgv")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(inout) :: a(*)
      complex(c_double_complex), intent(in) :: b(*)
      real(c_double), intent(out) :: evl(*)
      integer(c_int), value :: n, il, iu, lda, ldb
    end function
    integer(c_int) function ecalj_zhgv_dev(h, a, b, n, evl, il, iu, lda, ldb) bind(C, name="ecalj_zhgv_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, a, b, evl
      integer(c_int), value :: n, il, iu, lda, ldb
    end function
    ! gettetwt + tetwt5x_dtet4, npm=1 (S/m_tetwt.f90:36-256, S/tetwt5.f90:9-486)
    integer(c_int) function ecalj_tetwt(h, nband, nqbz, ntetf, idtetf, nqbzw, qbzw, ib1bz, ekxx1, ekxx2, ef, ebmx, &
        qlat, frhis, nwhis, nbnbx_cap, nhw_cap, nbnb, n1b, n2b, ihw, nhw, jhw, whw, nbnbx_out, nhwtot_out) &
        bind(C, name="ecalj_tetwt")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: nband, nqbz, ntetf, nqbzw, nwhis, nbnbx_cap
      integer(c_int64_t), value :: nhw_cap
      integer(c_int), intent(in) :: idtetf(*), ib1bz(*)
      real(c_double), intent(in) :: qbzw(*), ekxx1(*), ekxx2(*), qlat(*), frhis(*)
      real(c_double), value :: ef, ebmx
      integer(c_int), intent(out) :: nbnb(*), n1b(*), n2b(*), ihw(*), nhw(*), jhw(*), nbnbx_out
      real(c_double), intent(out) :: whw(*)
      integer(c_int64_t), intent(out) :: nhwtot_out
    end function
    ! gettetwt with npm = 1 or 2 and nctot core states ecore(1:nctot, is) (S/m_tetwt.f90:36-256,
    ! S/tetwt5.f90:179-185, 252-264, 349-427); lists (nbnbx_cap, nqbz, npm), nbnb(nqbz, npm)
    integer(c_int) function ecalj_tetwt_full(h, nband, nqbz, ntetf, idtetf, nqbzw, qbzw, ib1bz, ekxx1, ekxx2, ef, &
        ebmx, qlat, frhis, nwhis, npm, nctot, ecore, matrix_linear, nbnbx_cap, nhw_cap, nbnb, n1b, n2b, ihw, nhw, &
        jhw, whw, nbnbx_out, nhwtot_out) bind(C, name="ecalj_tetwt_full")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: nband, nqbz, ntetf, nqbzw, nwhis, npm, nctot, matrix_linear, nbnbx_cap
      integer(c_int64_t), value :: nhw_cap
      integer(c_int), intent(in) :: idtetf(*), ib1bz(*)
      real(c_double), intent(in) :: qbzw(*), ekxx1(*), ekxx2(*), qlat(*), frhis(*), ecore(*)
      real(c_double), value :: ef, ebmx
      integer(c_int), intent(out) :: nbnb(*), n1b(*), n2b(*), ihw(*), nhw(*), jhw(*), nbnbx_out
      real(c_double), intent(out) :: whw(*)
      integer(c_int64_t), intent(out) :: nhwtot_out
    end function
    ! m_read_ppovl tables of q, and the IPW rows of zmel (S/m_zmel.f90:418-538)
    integer(c_int) function ecalj_set_ppovl(h, nggg, ggg, ggg_lo, ggg_n, igggi, ngcgp, nvgcgp2, gcp_lo, gcp_n, &
        igcgp2i) bind(C, name="ecalj_set_ppovl")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      integer(c_int), value :: nggg, ngcgp
      complex(c_double_complex), intent(in) :: ggg(*)
      integer(c_int), intent(in) :: ggg_lo(3), ggg_n(3), igggi(*), nvgcgp2(*), gcp_lo(3), gcp_n(3), igcgp2i(*)
    end function
    integer(c_int) function ecalj_zmel_ipw(h, ngc, ngveccR, nadd, phase, ngpmx, ngp1, ngvecpB1, geigq, ntp0, ngp2, &
        ngvecpB2, dgeigqk, nm, zmel, ldz, row0) bind(C, name="ecalj_zmel_ipw")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      integer(c_int), value :: ngc, ngpmx, ngp1, ntp0, ngp2, nm, ldz, row0
      integer(c_int), intent(in) :: ngveccR(*), nadd(3), ngvecpB1(*), ngvecpB2(*)
      complex(c_double_complex), intent(in) :: phase(*), geigq(*), dgeigqk(*)
      complex(c_double_complex), intent(inout) :: zmel(*)
    end function
    integer(c_int) function ecalj_zmel_ipw_dev(h, ngc, ngveccR, nadd, phase, ngpmx, ngp1, ngvecpB1, geigq, ntp0, &
        ngp2, ngvecpB2, dgeigqk, nm, zmel, ldz, row0) bind(C, name="ecalj_zmel_ipw_dev")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h, geigq, dgeigqk, zmel
      integer(c_int), value :: ngc, ngpmx, ngp1, ntp0, ngp2, nm, ldz, row0
      integer(c_int), intent(in) :: ngveccR(*), nadd(3), ngvecpB1(*), ngvecpB2(*)
      complex(c_double_complex), intent(in) :: phase(*)
    end function
    integer(c_int) function ecalj_zmel_assemble_dev(h, mt, nbloch, ipw, ngc, ncol, is_m_basis, zmelconjg, zmel) &
        bind(C, name="ecalj_zmel_assemble_dev")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h, mt, ipw, zmel
      integer(c_int), value :: nbloch, ngc, is_m_basis, zmelconjg
      integer(c_int64_t), value :: ncol
    end function
    ! W records of __WVI/__WVR (S/m_llw.f90:102,177,308,373) to / from device memory
    integer(c_int) function ecalj_wv_write_dev(h, path, rec0, nrec, ld, w) bind(C, name="ecalj_wv_write_dev")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h, w
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: rec0, nrec, ld
    end function
    integer(c_int) function ecalj_wv_read_dev(h, path, rec0, nrec, ld, w) bind(C, name="ecalj_wv_read_dev")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h, w
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: rec0, nrec, ld
    end function
    ! the complex(4) records of the mixed-precision build (__MP, S/m_kind.f90:3-6)
    integer(c_int) function ecalj_wv4_write_dev(h, path, rec0, nrec, ld, w) bind(C, name="ecalj_wv4_write_dev")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h, w
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: rec0, nrec, ld
    end function
    integer(c_int) function ecalj_wv4_read_dev(h, path, rec0, nrec, ld, w) bind(C, name="ecalj_wv4_read_dev")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h, w
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: rec0, nrec, ld
    end function
    ! WVIllwI body, q /= 0 (S/m_llw.f90:299-373)
    integer(c_int) function ecalj_wc_build(h, chi, ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, wc, ldw) &
        bind(C, name="ecalj_wc_build")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: chi(*)
      integer(c_int), value :: ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, ldw
      complex(c_double_complex), intent(inout) :: wc(*)
    end function
    integer(c_int) function ecalj_wc_build_dev(h, chi, ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, wc, ldw) &
        bind(C, name="ecalj_wc_build_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, chi, wc
      integer(c_int), value :: ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, ldw
    end function
    ! imaginary-axis weights (S/m_sxcf_gemm.f90:432-458)
    integer(c_int) function ecalj_sigc_weights(h, niw, ns1, ns2, ntq, nctot, omega, ekc, esmr, ua, wkkr, freqx, &
        wwx, expa, wgtim) bind(C, name="ecalj_sigc_weights")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: niw, ns1, ns2, ntq, nctot
      real(c_double), intent(in) :: omega(*), ekc(*), freqx(*), wwx(*), expa(*)
      real(c_double), value :: esmr, ua, wkkr
      real(c_double), intent(out) :: wgtim(*)
    end function
    integer(c_int) function ecalj_sigc_weights_dev(h, nbatch, niw, ns1, ns2, ntq, nctot, omega, ekc, esmr, ua, &
        wkkr, freqx, wwx, expa, wgtim) bind(C, name="ecalj_sigc_weights_dev")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h, omega, ekc, wkkr, freqx, wwx, expa, wgtim
      integer(c_int), value :: nbatch, niw, ns1, ns2, ntq, nctot
      real(c_double), value :: esmr, ua
    end function
    ! Sigma_c contraction (S/m_sxcf_gemm.f90:463-496, 599-608)
    integer(c_int) function ecalj_sigc(h, zmel, ngb, nm, ntq, niw, wc, ldw, wgtim, zsec, ldz) &
        bind(C, name="ecalj_sigc")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: zmel(*), wc(*)
      integer(c_int), value :: ngb, nm, ntq, niw, ldw, ldz
      real(c_double), intent(in) :: wgtim(*)
      complex(c_double_complex), intent(inout) :: zsec(*)
    end function
    integer(c_int) function ecalj_sigc_dev(h, nbatch, zmel, ngb, nm, ntq, niw, wc, ldw, wgtim, zsec, ldz) &
        bind(C, name="ecalj_sigc_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, zmel, wc, wgtim, zsec
      integer(c_int), value :: nbatch, ngb, nm, ntq, niw, ldw, ldz
    end function
    ! real-axis pole term (S/m_sxcf_gemm.f90:501-608)
    integer(c_int) function ecalj_sigc_real(h, zmel, ngb, nm, ntq, wr, ldw, nw, freq_r, omega, ekc, ef, esmr, &
        nctot, ns2r, wkkr, zsec, ldz) bind(C, name="ecalj_sigc_real")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: zmel(*), wr(*)
      integer(c_int), value :: ngb, nm, ntq, ldw, nw, nctot, ns2r, ldz
      real(c_double), intent(in) :: freq_r(*), omega(*), ekc(*)
      real(c_double), value :: ef, esmr, wkkr
      complex(c_double_complex), intent(inout) :: zsec(*)
    end function
    integer(c_int) function ecalj_sigc_real_dev(h, nbatch, zmel, ngb, nm, ntq, wr, ldw, wr_stride, nw, freq_r, &
        omega, ekc, ef, esmr, nctot, ns2r, wkkr, zsec, ldz) bind(C, name="ecalj_sigc_real_dev")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h, zmel, wr, zsec
      integer(c_int), value :: nbatch, ngb, nm, ntq, ldw, nw, nctot, ns2r, ldz
      integer(c_int64_t), value :: wr_stride
      real(c_double), intent(in) :: freq_r(*), omega(*), ekc(*), wkkr(*)
      real(c_double), value :: ef, esmr
    end function
    ! get_correlation_block in full (S/m_sxcf_gemm.f90:423-608)
    integer(c_int) function ecalj_sigc_full_dev(h, nbatch, zmel, ngb, nm, ntq, niw, wc, ldw, wgtim, wr, ldwr, &
        wr_stride, nw, freq_r, omega, ekc, ef, esmr, nctot, ns2r, wkkr, zsec, ldz) bind(C, name="ecalj_sigc_full_dev")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h, zmel, wc, wgtim, wr, zsec
      integer(c_int), value :: nbatch, ngb, nm, ntq, niw, ldw, ldwr, nw, nctot, ns2r, ldz
      integer(c_int64_t), value :: wr_stride
      real(c_double), intent(in) :: freq_r(*), omega(*), ekc(*), wkkr(*)
      real(c_double), value :: ef, esmr
    end function
    ! Sigma_c with per-k middle-state windows (S/sxcf_fal2_count.sc.f90:118-236)
    integer(c_int) function ecalj_sigc_grouped_dev(h, nk, nm, zmel, ngb, ntq, niw, wc, ldw, wgtim, zsec, ldz) &
        bind(C, name="ecalj_sigc_grouped_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, zmel, wc, wgtim, zsec
      integer(c_int), value :: nk, ngb, ntq, niw, ldw, ldz
      integer(c_int), intent(in) :: nm(*)
    end function
    ! kernel-variant tuner and HIP graphs (no reference counterpart)
    integer(c_int) function ecalj_set_tuning(h, mode) bind(C, name="ecalj_set_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mode
    end function
    integer(c_int) function ecalj_graph_begin(h) bind(C, name="ecalj_graph_begin")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function ecalj_graph_end(h, g) bind(C, name="ecalj_graph_end")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr), intent(out) :: g
    end function
    integer(c_int) function ecalj_graph_launch(h, g) bind(C, name="ecalj_graph_launch")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, g
    end function
    integer(c_int) function ecalj_graph_destroy(g) bind(C, name="ecalj_graph_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: g
    end function
    ! QP energy (S/main_hqpe.sc.f90:152-161) and hermitised Sigma (:175-181)
    integer(c_int) function ecalj_qp_energy(ntq, elda, sex, sexcore, sec, ldsec, vxc, eqp) &
        bind(C, name="ecalj_qp_energy")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: ntq, ldsec
      real(c_double), intent(in) :: elda(*), sex(*), sexcore(*), vxc(*)
      complex(c_double_complex), intent(in) :: sec(*)
      real(c_double), intent(out) :: eqp(*)
    end function
    integer(c_int) function ecalj_sigma_herm(ntq, sex2, sexcore2, sec, vxcmat, se) bind(C, name="ecalj_sigma_herm")
      import :: c_int, c_double_complex
      integer(c_int), value :: ntq
      complex(c_double_complex), intent(in) :: sex2(*), sexcore2(*), sec(*), vxcmat(*)
      complex(c_double_complex), intent(out) :: se(*)
    end function
    ! q = 0 body of WVIllwI / WVRllwR with ix = 1 (S/m_llw.f90:98-178, 299-373); W in the E basis
    integer(c_int) function ecalj_wc_build_gamma(h, chi, ngb, nw, nspin, is_x0_m_basis, real_axis, wc, ldw) &
        bind(C, name="ecalj_wc_build_gamma")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: chi(*)
      integer(c_int), value :: ngb, nw, nspin, is_x0_m_basis, real_axis, ldw
      complex(c_double_complex), intent(inout) :: wc(*)
    end function
    integer(c_int) function ecalj_wc_build_gamma_dev(h, chi, ngb, nw, nspin, is_x0_m_basis, real_axis, wc, ldw) &
        bind(C, name="ecalj_wc_build_gamma_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, chi, wc
      integer(c_int), value :: ngb, nw, nspin, is_x0_m_basis, real_axis, ldw
    end function
    ! offset points q0i: llw = 1/epsinv(1,1) and the wing (S/m_llw.f90:183-256, 378-440)
    integer(c_int) function ecalj_eps_head(h, chi, ngb, nw, nspin, is_x0_m_basis, vcou1, real_axis, llw, wing_iw, &
        wing) bind(C, name="ecalj_eps_head")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: chi(*)
      integer(c_int), value :: ngb, nw, nspin, is_x0_m_basis, real_axis, wing_iw
      real(c_double), value :: vcou1
      complex(c_double_complex), intent(out) :: llw(*), wing(*)
    end function
    integer(c_int) function ecalj_eps_head_dev(h, chi, ngb, nw, nspin, is_x0_m_basis, vcou1, real_axis, llw, &
        wing_iw, wing) bind(C, name="ecalj_eps_head_dev")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h, chi, llw, wing
      integer(c_int), value :: ngb, nw, nspin, is_x0_m_basis, real_axis, wing_iw
      real(c_double), value :: vcou1
    end function
    ! getw0 (S/m_w0w0i.f90:146-311), host, no handle
    integer(c_int) function ecalj_getw0(ii, ie, nq0i, lmxax, llw, dmlx, epinvq0i, wklm, wbz, epinv, w0, llmat) &
        bind(C, name="ecalj_getw0")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: ii, ie, nq0i, lmxax
      complex(c_double_complex), intent(in) :: llw(*)
      real(c_double), intent(in) :: dmlx(*), epinvq0i(*), wklm(*), epinv(*)
      real(c_double), value :: wbz
      complex(c_double_complex), intent(out) :: w0(*), llmat(*)
    end function
    ! modifyWV0 on records in memory (S/m_w0w0i.f90:89-140)
    integer(c_int) function ecalj_wv0_patch(h, w, ngb, ldw, nw, w0, is_wc_m_basis) bind(C, name="ecalj_wv0_patch")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(inout) :: w(*)
      integer(c_int), value :: ngb, ldw, nw, is_wc_m_basis
      complex(c_double_complex), intent(in) :: w0(*)
    end function
    integer(c_int) function ecalj_wv0_patch_dev(h, w, ngb, ldw, nw, w0, is_wc_m_basis) &
        bind(C, name="ecalj_wv0_patch_dev")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h, w
      integer(c_int), value :: ngb, ldw, nw, is_wc_m_basis
      complex(c_double_complex), intent(in) :: w0(*)
    end function
    ! m_hamindex state of readcphi / readgeig (S/readeigen.f90:11-12)
    integer(c_int) function ecalj_set_hamindex(h, nqtt, qtt, iqmap, igmap, iqimap, ngp, ngpmx, imx, ngvecp, ngvecprev, &
        ngrp, symops, shtvg, nbas, miat, tiat, lmxax, dlmm, plat, qlat) bind(C, name="ecalj_set_hamindex")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nqtt, ngpmx, imx, ngrp, nbas, lmxax
      real(c_double), intent(in) :: qtt(*), symops(*), shtvg(*), tiat(*), dlmm(*), plat(*), qlat(*)
      integer(c_int), intent(in) :: iqmap(*), igmap(*), iqimap(*), ngp(*), ngvecp(*), ngvecprev(*), miat(*)
    end function
    ! readmnla_cphi tables (S/readeigen.f90:487-521)
    integer(c_int) function ecalj_set_mnla(h, norbt, ibas_tbl, l_tbl, k_tbl, offset_tbl) bind(C, name="ecalj_set_mnla")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: norbt
      integer(c_int), intent(in) :: ibas_tbl(*), l_tbl(*), k_tbl(*), offset_tbl(*)
    end function
    ! readcphi / readgeig for a batch of target points (S/readeigen.f90:71-181)
    integer(c_int) function ecalj_readcphi_dev(h, nk, iq, nspc, nband, cphi_irr, ldc, cphi_out) &
        bind(C, name="ecalj_readcphi_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, cphi_irr, cphi_out
      integer(c_int), value :: nk, nspc, nband, ldc
      integer(c_int), intent(in) :: iq(*)
    end function
    integer(c_int) function ecalj_readgeig_dev(h, nk, iq, nspc, nband, geig_irr, ldg, geig_out) &
        bind(C, name="ecalj_readgeig_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, geig_irr, geig_out
      integer(c_int), value :: nk, nspc, nband, ldg
      integer(c_int), intent(in) :: iq(*)
    end function
    ! readers of @MNLA_CPHI, __QGpsi, __EValue and the __CPHI / __GEIG records
    integer(c_int) function ecalj_mnla_read(path, cap, norbt, ibas_tbl, l_tbl, k_tbl, offset_tbl) &
        bind(C, name="ecalj_mnla_read")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: cap
      integer(c_int), intent(out) :: norbt, ibas_tbl(*), l_tbl(*), k_tbl(*), offset_tbl(*)
    end function
    integer(c_int) function ecalj_qgpsi_read(path, hdr6, qpgcut, nq_cap, ngpmx_cap, imx_cap, qtt, ngp, irr, ngvecp, &
        ngvecprev) bind(C, name="ecalj_qgpsi_read")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: hdr6(6), ngp(*), irr(*), ngvecp(*), ngvecprev(*)
      real(c_double), intent(out) :: qpgcut, qtt(*)
      integer(c_int), value :: nq_cap, ngpmx_cap, imx_cap
    end function
    integer(c_int) function ecalj_evalue_read(path, hdr5, cap, qtti, evud) bind(C, name="ecalj_evalue_read")
      import :: c_int, c_char, c_double, c_int64_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: hdr5(5)
      integer(c_int64_t), value :: cap
      real(c_double), intent(out) :: qtti(*), evud(*)
    end function
    integer(c_int) function ecalj_rec_read_dev(h, path, recl, rec0, step, nrec, nbytes, dst, dst_stride) &
        bind(C, name="ecalj_rec_read_dev")
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: h, dst
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: recl, nbytes, dst_stride
      integer(c_int), value :: rec0, step, nrec
    end function
  end interface
end module ecalj_gpu
