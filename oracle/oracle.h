/*
 * oracle.h -- FP64 CPU restatement of ecalj's QSGW chi0 -> W -> Sigma_c path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in ecalj_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  The product library never links or
 * calls it.
 *
 * Every function restates one reference routine (file:line under the ecalj
 * source tree, SRC/subroutines = S/).  Arrays are Fortran column-major;
 * complex arrays are interleaved (re, im) doubles (C99 double complex).
 * Band and bin indices stored inside index arrays are 1-based, as in Fortran.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   orc_gauss          numpy.polynomial.legendre.leggauss, exact polynomial
 *                      integration; the reference's 15-digit table values
 *   orc_frhis          closed form bb*(exp(aa*(i-1))-1)
 *   orc_hilbertmat     oracle/_ref (S/hilbertmat.f90 compiled with flang),
 *                      analytic Hilbert transform of piecewise-linear f
 *   orc_wfacx2/weavx2  oracle/_ref (S/wfacx.f90 compiled with flang)
 *   orc_zmel           numpy einsum on tiny inputs; unitary M->E invariance
 *   orc_zmel_ipw       numpy triple sum over (G1, G2); both branches agree
 *   orc_x0gemm         brute-force numpy sum over pairs; Hermiticity
 *   orc_dpsion_imag    numpy; delta-pair limit -2D/(D^2+nu^2)
 *   orc_rmatt          oracle/_ref hilbertmat at real zz; principal-value
 *                      quadrature of the piecewise-linear interpolant
 *   orc_dpsion_real    Kramers-Kronig pair of a two-pole response (closed
 *                      form of Re chi from binned Im chi); slot-0 term
 *   orc_chi0_freq      equals x0gemm+dpsion with composed weights
 *   orc_wc_build       numpy.linalg.inv textbook W; Dyson residual
 *   orc_wgtim          scipy.integrate.quad of the smeared-Lorentzian integral
 *   orc_sigc_imag      numpy GEMM chain; Im(Sigma_nn) clip
 *   orc_qp_energy      closed form of S/main_hqpe.sc.f90:160-161
 *   orc_lindtet6       Monte Carlo of f(ea)(1-f(eb)) delta(omega+x) over the
 *                      tetrahedron, every occupied/unoccupied cut figure
 *   orc_lindtet6_ml    corner sum = 4 x orc_lindtet6; on an uncut tetrahedron
 *                      the vertex average of the cross-section polygon at
 *                      the bin centre (brute force over the six edges);
 *                      sum_c w_c wt_c within the half bin width of the bin
 *                      centre; the point inside the occupied/unoccupied
 *                      region; invariance under an affine map of k and
 *                      covariance under relabelling the corners
 *   orc_tetwt          brute-force pair sets; degenerate-band equality;
 *                      Monte Carlo of the whole-mesh histogram
 *   orc_sylm/sylmnc    oracle/_ref (S/sylm.f90 compiled with flang)
 *   orc_gaunt          oracle/_ref scg (S/scg.f90 compiled with flang) via
 *                      the table layout of rotcg; sympy real_gaunt
 *   orc_getw0          isotropic closed form wklm(1)(4pi)^1.5/wbz (1/eps-1);
 *                      Galerkin residual of 1/L(q) by dense quadrature of the
 *                      quadratic form q.epinv.q itself; llmat sum
 *   orc_eps_head       diagonal chi (llw = the no-LFC head); Schur complement
 *                      eps11 - eps1r eps_rr^-1 eps_r1 by numpy.linalg.solve
 *   orc_wc_build_gamma numpy.linalg.inv of the mu,nu >= 2 block
 *   orc_wv0_patch      W(1,1) set; unitary E->M by numpy
 *   orc_rotmto/rotipw  the rotated Bloch function psi(g^-1 (r - s)) of a
 *                      synthetic two-atom crystal evaluated in real space
 *                      (muffin-tin Bloch sums, plane waves) against the one
 *                      built from the rotated coefficients at g k + G0
 */
#ifndef ECALJ_ORACLE_H
#define ECALJ_ORACLE_H
#include <complex.h>
#ifdef __cplusplus
extern "C" {
#endif

/* BLAS-semantics complex GEMM, column-major, op in {'N','T','C'};
 * stands for zmm_h (S/m_blas.f90:205-232). */
void orc_zgemm(char opa, char opb, int m, int n, int k, double complex alpha,
               const double complex* a, long lda, const double complex* b, long ldb,
               double complex beta, double complex* c, long ldc);
/* x0kf_v4hz_init (S/x0kf_v4h.f90:48-142), jobs 0 and 1: the flattened pair
 * and bin lists of x0gemm from tetwt's per-k pair tables.  Pair tables are
 * Fortran arrays (nbnbx, nqbz, npm); nbnb is (nqbz, npm); indices 1-based. */
int orc_x0kf_init(int nqbz, int npm, int nbnbx, const int* nbnb, const int* n1b, const int* n2b,
                  const int* ihw, const int* nhw, const int* jhw, const double* whw, int nband, int nctot,
                  int nwhis, long cap_coun, long cap_count, int* kc, int* itc, int* itpc, int* jpmc,
                  int* iwini, int* iwend, int* icouini, double* whwc, int* icounkmin, int* icounkmax,
                  int* nkmin, int* nkmax, int* nkqmin, int* nkqmax, long* ncoun_out, long* ncount_out);

/* getwemax (S/m_freq.f90:197-268) + getfreq's real mesh (:110-129): nw and
 * freq_r(0:nw) (Hartree) from the external and BZ band energies (Ry). */
int orc_real_mesh(int neq, const double* eq, long nbz, const double* ebz, double ef, double deltaw, double esmr,
                  double ffac, int nwhis, const double* frhis, double* wemax_out, double* freq_r, int cap);

/* Hooks to an LP64 Fortran-ABI BLAS/LAPACK: zgemm for orc_zgemm, zgetrf +
 * zgetri for orc_zinv (the routines the reference calls). NULL = loops. */
void orc_use_blas(void* zgemm, void* zgetrf, void* zgetri);
int orc_blas_active(void);

/* gauss (S/mate.f90:1-206): n-point Gauss-Legendre on [a,b]. Returns 0 or -1
 * when n is not one of the reference's supported sizes. */
int orc_gauss(int n, double a, double b, double* x, double* w);

/* freq01 (S/m_freq.f90:164-186) + readfreq_r (S/m_readfreq_r.f90:41-50). */
int orc_freq01(int nx, double ua, double* freqx, double* freqw, double* wx, double* expa);

/* Histogram bins (S/m_freq.f90:95-108): returns nwhis; fills frhis[0..nwhis]
 * when frhis != NULL (frhis(1:nwhis+1) in Fortran). */
int orc_frhis(double oratio, double dw, double omg2max, double* frhis);

/* hilbertmat (S/hilbertmat.f90:14-79). rmat has 2*nwhis+1 entries for
 * indices -nwhis..nwhis; his_* likewise. */
void orc_hilbertmat(double complex zz, int nwhis, const double* his_L, const double* his_C,
                    const double* his_R, double complex* rmat);

/* dpsion_init, imaginary branch with npm=1 (S/dpsion5.f90:84-89,121-129).
 * imatt is niw x nwhis column-major. */
void orc_imatt(int niw, const double* freqi, int nwhis, const double* frhis, double* imatt);

/* zmel, MT valence-valence part with identity rotation
 * (S/m_zmel.f90:146-597: :323-371 MT contraction, :546-563 M->E,
 * :590-594 conj).  Atom ia owns basis rows [ias, ias+nv[ia]) and product
 * basis rows [ims, ims+mdim[ia]) (running sums).  ppb holds the per-atom real
 * blocks ppb(i1,i2,I) of size nv x nv x mdim, concatenated.  When m2e is
 * non-NULL (is_m_basis = false) zmel = m2e^H zmelt with m2e ngb x nbb.
 * zmel out: (m2e ? nbb : ngb) x nm x nq. */
int orc_zmel(int natom, const int* nv, const int* mdim, const double complex* phase,
             const double* ppb, int ndima, int nm, const double complex* cphim, int nq,
             const double complex* cphiq, const double complex* m2e, int nbb, int conjg,
             double complex* zmel);

/* get_zmel_init_gemm's MT part with core states and the rotated-atom map
 * iatomp (1-based; NULL = identity) on the reference's ppb(nlnmx, nlnmx,
 * mdimx, natom) (S/m_zmel.f90:199-214, 323-391, 559-560, 590-594).  Middle
 * states ns1..ns2 (core 1..nctot first); cphim holds the valence bands of
 * that range.  Pinned by the valence formula on core-extended vectors and by
 * the row permutation that iatomp stands for (tests/test_zmel_core_oracle.py). */
int orc_zmel_init(int natom, int nlnmx, int mdimx, const int* nlnmc, const int* nlnmv, const int* nblocha,
                  const int* iatomp, const double complex* phase, const double* ppb, int ndima, int ns1, int ns2,
                  int nctot, const double complex* cphim, int nq, const double complex* cphiq,
                  const double complex* m2e, int nbb, int conjg, double complex* zmel);

/* x0gemm (S/x0gemm.f90:2-97) with its module inputs explicit.
 * rcxq: npr x npr_col x nslice, slice s <-> Fortran index s + (1-npm)*nwhis
 *   (npm=1: slices 0..nwhis; npm=2: -nwhis..nwhis).
 * zmel: nbb x (ns2-ns1+1) x nqtot (Fortran zmel(nbb,ns1:ns2,nqtot)).
 * icoun range [icoun_lo, icoun_hi] is 1-based into itc/itpc/jpmc/iwini/iwend/icouini. */
void orc_x0gemm(double complex* rcxq, int npr, int ipr_col, int npr_col, int nwhis, int npm,
                int ns1, int ns2, const double complex* zmel, int nbb, int nqtot,
                int icoun_lo, int icoun_hi, const int* iwini, const int* iwend, const int* itc,
                const int* itpc, const int* jpmc, const int* icouini, const double* whwc);

/* dpsion_init, real branch with npm=1, no chipm (S/dpsion5.f90:91-103).
 * rmatt is (nw+1) x nwhis column-major; freqr(0:nw). */
void orc_rmatt(int nw, const double* freqr, int nwhis, const double* frhis, double* rmatt);
/* dpsion_chiq, real axis with npm=1, no chipm (S/dpsion5.f90:193-200,244-256):
 * scales rcxq in place unless `scaled`, then
 * zxq(:,:,it) = i rcxq(:,:,it) + sum_iw rmatt(it,iw) rcxq(:,:,iw), it = 0..nw. */
void orc_dpsion_real(int npr, int npr_col, int nwhis, int nw, const double* frhis,
                     const double* rmatt, double ecut, int scaled, double complex* rcxq,
                     double complex* zxq);
/* zmel IPW block (S/m_zmel.f90:418-538): rows igc of zmel (leading dimension
 * ldz) for nm middle x ntp0 end states.  G lists 3 x n (int), cubes Fortran
 * order with lower bounds lo and extents n.  branch: 0 = as the reference
 * (nm > ntp0: G1 first), 1 or 2 forced. */
int orc_zmel_ipw(int ngc, const int* ngveccR, const int* nadd, const double complex* phase, int ngpmx, int ngp1,
                 const int* ngvecpB1, const double complex* geigq, int ntp0, int ngp2, const int* ngvecpB2,
                 const double complex* dgeigqk, int nm, const double complex* ggg, const int* glo, const int* gn,
                 const int* igggi, int ngcgp, const int* nvgcgp2, const int* clo, const int* cn,
                 const int* igcgp2i, int branch, double complex* zmel, long ldz);
/* dpsion_chiq, imaginary axis with npm=1 (S/dpsion5.f90:211-232): scales rcxq
 * bins 1..nwhis in place by -wfac/(R-L) and sets
 * zxqi(:,:,iv) = sum_iw imatt(iv,iw) rcxq(:,:,iw).  rcxq has nwhis+1 slices
 * (slice 0 unused); zxqi npr x npr_col x niw. */
void orc_dpsion_imag(int npr, int npr_col, int nwhis, int niw, const double* frhis,
                     const double* imatt, double ecut, double complex* rcxq, double complex* zxqi);

/* Composed direct-frequency chi0: chi(I,J,v) = sum_p w(v,p) conj(Z(I,p)) Z(J,p),
 * i.e. x0gemm followed by dpsion_chiq with w = sum_bins whw*scale*imatt. */
void orc_chi0_freq(int ni, int npairs, int nf, const double complex* z, const double* w,
                   double complex* chi);

/* WVIllwI body for q /= 0 (S/m_llw.f90:299-373) with zminv_h
 * (S/m_lapack.f90:21-47).  chi: ngb x ngb x nw (M basis if is_x0_m_basis),
 * vcousq = sqrt(vcoud) (S/m_readVcoud.f90:40), m2e = zcousq ngb x ngb.
 * wc out: ldw x ngb x nw.  Returns 0, or 2 if an epsilon matrix is singular. */
int orc_wc_build(int ngb, int nw, int nspin, int is_x0_m_basis, int is_wc_m_basis,
                 const double complex* chi, const double* vcousq, const double complex* m2e,
                 double complex* wc, int ldw);

/* In-place general complex inverse (getrf + getri semantics). Returns 0/2. */
int orc_zinv(int n, double complex* a, long lda);

/* Imaginary-axis Sigma_c weights (S/m_sxcf_gemm.f90:432-458).
 * omega[ntq] and ekc[ns1-1 .. ns2-1] (0-based storage of Fortran ekc(ns1:ns2))
 * in Ry; esmr in Ry; x,wt: Gauss points/weights (niw); expa: exp(-ua^2 w^2).
 * wgtim out: (niw+1) x (ns2-ns1+1) x ntq, already multiplied by wkkr. */
void orc_wgtim(int niw, int ns1, int ns2, int ntq, int nctot, const double* omega,
               const double* ekc, double esmr, double ua, double wkkr, const double* x,
               const double* wt, const double* expa, double* wgtim);

/* Imaginary-axis Sigma_c and final contraction (S/m_sxcf_gemm.f90:463-496,
 * 599-608).  zmel: ngb x nm x ntq (M basis, not conjugated).
 * wc: ldw x ngb x (niw+1); only its upper triangle is used, as the keepwv
 * packing does (S/m_sxcf_gemm.f90:337-344,468-475).
 * zsec: ldz x ntq, accumulated (beta = 1), then Im(zsec_nn) clipped to <= 0. */
void orc_sigc_imag(int ngb, int nm, int ntq, int niw, const double complex* zmel,
                   const double complex* wc, int ldw, const double* wgtim,
                   double complex* zsec, int ldz);

/* Real-axis pole term of Sigma_c for one internal k (S/m_sxcf_gemm.f90:501-594,
 * 599-608); see oracle.c.  Returns the number of (iw, column) entries or -1. */
int orc_sigc_real(int ngb, int nm, int ntq, const double complex* zmel, const double complex* wr, int ldw, int nw,
                  const double* freq_r, const double* omega, const double* ekc, double ef, double esmr, int nctot,
                  int ns2r, double wkkr, double complex* zsec, int ldz);

/* get_correlation_block (S/m_sxcf_gemm.f90:423-608): both parts in one czmelwc. */
int orc_sigc_full(int ngb, int nm, int ntq, int niw, const double complex* zmel, const double complex* wc, int ldw,
                  const double* wgtim, const double complex* wr, int ldwr, int nw, const double* freq_r,
                  const double* omega, const double* ekc, double ef, double esmr, int nctot, int ns2r, double wkkr,
                  double complex* zsec, int ldz);

/* wfacx2 / weavx2 (S/wfacx.f90:19-49). */
double orc_wfacx2(double e1, double e2, double ek, double esmr);
double orc_weavx2(double e1, double e2, double ek, double esmr);

/* QP energy (S/main_hqpe.sc.f90:152-161): eqp = elda + (sex+sexcore)*hartree
 * + Re(sec_nn)*hartree - vxc.  Energies in eV, self-energies in Hartree. */
void orc_qp_energy(int ntq, const double* elda, const double* sex, const double* sexcore,
                   const double complex* sec, int ldsec, const double* vxc, double* eqp);

/* Hermitised Sigma (S/main_hqpe.sc.f90:175-181, Hartree):
 * se = sex2 + sexcore2 + (sec + sec^H)/2 - vxcmat, all ntq x ntq. */
void orc_sigma_herm(int ntq, const double complex* sex2, const double complex* sexcore2,
                    const double complex* sec, const double complex* vxcmat, double complex* se);

/* Tetrahedron histogram weights (oracle/tetwt.c): gettetwt + tetwt5x_dtet4
 * (S/m_tetwt.f90:36-256, S/tetwt5.f90:9-486) for npm=1, no core states,
 * matrix_linear=F.  idtetf: 4 x ntetf 1-based indices into qbzw (3 x nqbzw);
 * ib1bz: nqbzw 1-based k indices; ek1/ek2: nband x nqbz energies (Ry) at k
 * and k+q; qlat columns = reciprocal vectors.  Outputs with leading dimension
 * nbnbx_cap: n1b, n2b, ihw, nhw, jhw (1-based values, Fortran layout); whw
 * (nhw_cap).  Returns 0; -1 if a capacity is short (nbnbx_out / nhwtot_out
 * hold the sizes needed); < -1 on the reference's own error exits. */
int orc_tetwt(int nband, int nqbz, int ntetf, const int* idtetf, const double* qbzw, const int* ib1bz,
              const double* ek1, const double* ek2, double ef, double ebmx, const double* qlat,
              const double* frhis, int nwhis, int nbnbx_cap, int nhw_cap, int* nbnb, int* n1b, int* n2b,
              int* ihw, int* nhw, int* jhw, double* whw, int* nbnbx_out, int* nhwtot_out);
/* gettetwt with npm = 1 or 2 (time reversal off: jpm = 2 pairs are unoccupied
 * at k and occupied at k+q) and nctot core states of energies ecore (Ry),
 * appended after the valence bands at every k (S/tetwt5.f90:179-185,
 * 252-264, 349-427).  Per (k, jpm) outputs: nbnb(nqbz, npm); n1b, n2b, ihw,
 * nhw, jhw (nbnbx_cap, nqbz, npm); whw with its pointers running over jpm,
 * then k, then the pairs.  n1b ranges over nband + nctot, n2b over
 * nband + (npm == 2 ? nctot : 0).  orc_tetwt is npm = 1 without core. */
int orc_tetwt_full(int nband, int nqbz, int ntetf, const int* idtetf, const double* qbzw, const int* ib1bz,
                   const double* ek1, const double* ek2, double ef, double ebmx, const double* qlat,
                   const double* frhis, int nwhis, int npm, int nctot, const double* ecore, int matrix_linear,
                   int nbnbx_cap, int nhw_cap, int* nbnb, int* n1b, int* n2b, int* ihw, int* nhw, int* jhw, double* whw,
                   int* nbnbx_out, int* nhwtot_out);
/* lindtet6 (S/tetwt5.f90:514-600): histogram of one tetrahedron; kvec[corner][xyz] */
int orc_lindtet6(const double kvec[4][3], const double ea[4], const double eb[4], const double x[4],
                 double efermia, double efermib, const double* frhis, int nwhis, double* wtthis);
/* the same with tetrahedron_matrix_linear (S/tetwt5.f90:694-709, 751-759): wtthis(nwhis, 4), column c
 * the weight carried by corner c of kvec, summing over c to 4 times orc_lindtet6's */
int orc_lindtet6_ml(const double kvec[4][3], const double ea[4], const double eb[4], const double x[4],
                    double efermia, double efermib, const double* frhis, int nwhis, double* wtthis);
int orc_hisrange(const double* frhis, int nwhis, double demin, double demax, int* ihw, int* nhw);
int orc_chkdgn(const double* ene, int ndat, int* ixini, int* ixend);

/* Gamma point (oracle/gamma.c).  sylm/sylmnc (S/sylm.f90:1-121): lm index
 * 0-based in yl / cy.  orc_gaunt: g[L1 + n12*(L2 + n12*L3)] = int Y_L1 Y_L2
 * Y_L3 for L1, L2 < n12 = (l12+1)^2, L3 < (l3+1)^2 (rotcg with the identity,
 * S/rotcg.f90:54-69). */
void orc_sylm(const double* r, double* yl, int lmx);
void orc_sylmnc(double* cy, int lmx);
int orc_gaunt(int l12, int l3, double* g);
/* getw0 (S/m_w0w0i.f90:146-311): llw(ii:ie, nq0i), dmlx(nq0i,9),
 * epinvq0i(nq0i,nq0i), wklm((lmxax+1)^2), epinv(3,3,nq0i) -> w0(ii:ie),
 * llmat(3,3) at iw = 0 when ii <= 0.  Returns 0 or 2 (singular). */
int orc_getw0(int ii, int ie, int nq0i, int lmxax, const double complex* llw, const double* dmlx,
              const double* epinvq0i, const double* wklm, double wbz, const double* epinv,
              double complex* w0, double complex* llmat);
/* q = 0 body of the W build (S/m_llw.f90:98-178, 299-373 with ix = 1): W in
 * the E basis, row and column 1 zero.  wc: ldw x ngb x nw. */
int orc_wc_build_gamma(int ngb, int nw, int nspin, int is_x0_m_basis, const double complex* chi,
                       const double* vcousq, const double complex* m2e, double complex* wc, int ldw);
/* Offset points q0i (S/m_llw.f90:183-256, 378-440): llw(iw) = 1/epsinv(1,1)
 * with the head 1 - vcou1 chi(1,1); wing(0..ngb-2) = epsinv(1,2:)/epsinv(1,1)
 * at frequency wing_iw (skipped when out of range). */
int orc_eps_head(int ngb, int nw, int nspin, int is_x0_m_basis, const double complex* chi, const double* vcousq,
                 const double complex* m2e, double vcou1, double complex* llw, int wing_iw, double complex* wing);
/* modifyWV0 (S/m_w0w0i.f90:89-140): W(1,1,iw) = w0(iw), then E -> M. */
void orc_wv0_patch(int ngb, int nw, double complex* w, int ldw, const double complex* w0, int is_wc_m_basis,
                   const double complex* m2e);

/* Eigenfunction rotation (oracle/rotwave.c): rotmto (S/rotwave.f90:156-192)
 * with the @MNLA_CPHI tables (1-based ibas/l/k, offset = first row - 1), and
 * rotipw (S/rotwave.f90:223-245).  sym, platt, qlat: 3x3 column-major;
 * dlmm(-lmxax:lmxax, -lmxax:lmxax, 0:lmxax) of the operation. */
int orc_rotmto(const double* qin, int ndima, int nband, int norbt, const int* ibas_tbl, const int* l_tbl,
               const int* k_tbl, const int* offset_tbl, const double* sym, const double* dlmm, int lmxax,
               const int* miat, const double* tiat, int nbas, const double complex* cin, double complex* cout);
int orc_rotipw(const double* qin, const double* qtarget, int ngp, int nband, const double* platt, const double* qlat,
               const double* sym, const int* ngvecp, const int* ngvecprev, int imx, const double* shtvg,
               const double complex* gin, double complex* gout, long ldg);

/* Learned chi0-tail MLP (oracle/mlp.c): L layers of widths dims[0..L],
 * silu between layers, linear output; params W_l (row-major out x in) then
 * b_l per layer; batches row-major.  No reference counterpart (the reference
 * truncates the band sum: nband_chi0, S/m_readgwinput.f90:34-35). */
#include <stdint.h>
int64_t orc_mlp_nparam(int L, const int* dims);
void orc_mlp_forward(int L, const int* dims, const double* p, const double* x, int B, double* y);
/* returns the mean-squared-error loss; g (nparam) receives its gradient when non-NULL */
double orc_mlp_loss_grad(int L, const int* dims, const double* p, const double* x, const double* t, int B,
                         double* g);
double orc_silu(double z);
double orc_dsilu(double z);

#ifdef __cplusplus
}
#endif
#endif
