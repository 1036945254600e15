/*
 * ecalj_gpu.h -- C ABI of the MI355X (gfx950) implementation of ecalj's QSGW
 * chi0 -> W -> Sigma_c path.
 *
 * Conventions (those of the Fortran callers):
 *   - arrays are column-major; complex arrays are interleaved (re, im)
 *     doubles (Fortran complex(8)); index arrays hold 1-based values;
 *   - functions without a suffix take HOST pointers: they copy inputs to the
 *     device, compute, copy results back and synchronise;
 *   - functions with suffix _dev take DEVICE pointers and are asynchronous on
 *     the handle's stream (ecalj_sync waits);
 *   - every function returns an ecalj_status; nothing aborts the process.
 *     Callers map a nonzero status to their own error exit (the reference
 *     calls rx(), S/rx.f90:2-12).
 *   - one handle per host thread; a handle is not re-entrant.  Handles are
 *     independent of each other.
 *
 * Each entry point names the reference routine it replaces (S/ =
 * SRC/subroutines/ of ecalj).
 */
#ifndef ECALJ_GPU_H
#define ECALJ_GPU_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ecalj_ctx* ecalj_handle;

typedef enum {
  ECALJ_OK = 0,
  ECALJ_ERR_DIM = 1,         /* inconsistent or unsupported dimensions       */
  ECALJ_ERR_SINGULAR = 2,    /* epsilon matrix singular (zero pivot)         */
  ECALJ_ERR_HIP = 3,         /* HIP runtime error                            */
  ECALJ_ERR_ARG = 4,         /* NULL pointer or bad enum                     */
  ECALJ_ERR_NOMEM = 5,       /* device allocation failed                     */
  ECALJ_ERR_STATE = 6,       /* required set-up call missing (set_ppb, ...)  */
  ECALJ_ERR_NOT_PD = 7,      /* matrix not positive definite (zhgv's B)      */
  ECALJ_ERR_NOT_CONVERGED = 8, /* iterative eigensolver did not converge     */
  ECALJ_ERR_SCRATCH = 9       /* request exceeds ECALJ_NOALLOC_SCRATCH_BYTES  */
} ecalj_status;

/* Fixed device scratch of the entry point that allocates nothing
 * (ecalj_zminv_noalloc_dev), one array per device held until process exit. */
#define ECALJ_NOALLOC_SCRATCH_BYTES (64u << 20)

/* Arithmetic of the contractions (zmel, chi0, M<->E transforms, Sigma_c).
 * The epsilon inversion, the weights and the QP update are always FP64. */
typedef enum {
  ECALJ_PREC_F64 = 0,     /* complex128 data, FP64 MFMA                          */
  ECALJ_PREC_F32 = 1,     /* complex64 products on FP32 MFMA, FP32 accumulation  */
  ECALJ_PREC_BF16X3 = 2,  /* hi/lo bf16 split, 3 MFMA products, FP32 accumulation */
  ECALJ_PREC_BF16X3_F8 = 3 /* bf16x3, except the Sigma_c frequency contraction:
                              hi*hi on bf16 MFMA, hi*lo + lo*hi on the scaled fp8
                              (E4M3) MFMA, ~2^-12 per product                    */
} ecalj_precision;

/* ---- handle ------------------------------------------------------------ */
/* Replaces gpu_init (S/m_gpu.f90:14-68): binds device `device` (the caller
 * passes local_rank mod ndevs as the reference does) and creates a stream. */
ecalj_status ecalj_create(ecalj_handle* h, int device);
ecalj_status ecalj_destroy(ecalj_handle h);
ecalj_status ecalj_set_precision(ecalj_handle h, int precision);
/* Use an external hipStream_t (e.g. the caller's framework stream); NULL restores the handle's own. */
ecalj_status ecalj_set_stream(ecalj_handle h, void* hip_stream);
ecalj_status ecalj_sync(ecalj_handle h);
/* No counterpart in the reference (stream plumbing): h's stream waits on the
 * device for the work issued so far on other's stream (an event, no host
 * wait), so that independent steps of one q-point on two handles of the same
 * device can overlap.  Not under graph capture (ECALJ_ERR_STATE). */
ecalj_status ecalj_stream_wait(ecalj_handle h, ecalj_handle other);
const char* ecalj_status_string(int status);
/* Kernel timing with HIP events on the handle's stream: when enabled, the
 * launches of kernel class `kclass` are bracketed by events; ecalj_kernel_time
 * returns their summed duration (ms) and count since the last reset.
 * kclass: 0 = Sigma_c frequency GEMM, 1 = chi0 bin accumulation,
 * 2 = zmel GEMMs, 3 = W build (transforms + inversion), 4 = Hilbert. */
ecalj_status ecalj_kernel_timing(ecalj_handle h, int enable);
ecalj_status ecalj_kernel_time(ecalj_handle h, int kclass, double* ms, int64_t* launches);
/* Kernel-variant tuner (no counterpart in the reference, whose GEMMs have one
 * fixed shape each).  mode 0 (default): on the first call of a shape, every
 * variant of a tuned kernel runs once, timed with HIP events (that call
 * waits for the stream), and the fastest is kept for the shape; mode k > 0:
 * always variant k.  Tuned kernels: the bf16x3 Sigma_c contraction (variants
 * 1 = 128x64 tiles, 3 stages, 4 waves; 2 = 128x128 tiles, 2 stages, 4 waves;
 * 3 = 128x128, 2 stages, 8 waves; 4 = variant 3 with the XCD-aware block
 * order), and its grouped form (variants 1-3; k > 3 wraps).
 * ecalj_tuning_report writes the decisions so far, one line per shape, as a
 * NUL-terminated string of at most len bytes. */
ecalj_status ecalj_set_tuning(ecalj_handle h, int mode);

/* HIP graphs of the handle's stream (no counterpart in the reference, which
 * issues every call eagerly).  Between ecalj_graph_begin and ecalj_graph_end
 * the _dev entry points are recorded instead of run; ecalj_graph_launch
 * replays them on the handle's stream.  Capture needs a warm-up: the same
 * calls at the same sizes must have run once before (scratch may not grow
 * under capture: ECALJ_ERR_STATE), and the tuner then reuses its choices.
 * Index lists built on the host at capture time (x0gemm bins, Hilbert
 * coefficients) are fixed in the graph: a replay repeats the captured
 * problem on whatever data the device buffers hold.  Kernel timing is off
 * while capturing.  Fill kernels, not memset nodes, clear memory.
 * A graph keeps the device addresses of the handle's scratch and state as
 * they were at capture.  Any later call that frees or moves them (scratch
 * growth in an eager call, set_ppb, set_coulomb, set_ppovl, set_hamindex)
 * starts a new state generation: ecalj_graph_launch then returns
 * ECALJ_ERR_STATE for graphs of an older generation, and ECALJ_ERR_ARG when
 * called with a handle other than the capturing one.  ecalj_graph_destroy
 * waits for the graph's last launch.  ecalj_graph_begin returns
 * ECALJ_ERR_STATE in a process run under the rocprofv3 profiler (a traced
 * graph replay crashed inside the tracer): profile the eager calls. */
typedef struct ecalj_graph_s* ecalj_graph_t;
ecalj_status ecalj_graph_begin(ecalj_handle h);
ecalj_status ecalj_graph_end(ecalj_handle h, ecalj_graph_t* graph);
ecalj_status ecalj_graph_launch(ecalj_handle h, ecalj_graph_t graph);
ecalj_status ecalj_graph_destroy(ecalj_graph_t graph);
ecalj_status ecalj_tuning_report(ecalj_handle h, char* buf, int64_t len);

/* ---- device memory ------------------------------------------------------
 * The reference keeps its arrays on the device with OpenACC data directives
 * (!$acc enter data create/copyin, !$acc update host; e.g.
 * S/x0kf_v4h.f90:190-207, S/m_sxcf_gemm.f90:362).  A caller without such a
 * runtime allocates and moves device memory through the library, so that the
 * process holds one HIP runtime.  ecalj_dev_zero uses a fill kernel (no
 * hipMemset), so it may be captured in a HIP graph.  ecalj_dev_phase_similarity
 * forms out_k = diag(d_k) A diag(d_k)^H for nk phase vectors d_k (ngb
 * complex each) and nmat matrices A (ngb x ngb): the synthetic W_c(k) inputs
 * of the benchmark workload; it has no reference counterpart. */
ecalj_status ecalj_dev_alloc(ecalj_handle h, size_t bytes, void** ptr);
ecalj_status ecalj_dev_free(ecalj_handle h, void* ptr);
ecalj_status ecalj_memcpy_h2d(ecalj_handle h, void* dst, const void* src, size_t bytes);
ecalj_status ecalj_memcpy_d2h(ecalj_handle h, void* dst, const void* src, size_t bytes);
ecalj_status ecalj_memcpy_d2d(ecalj_handle h, void* dst, const void* src, size_t bytes);
ecalj_status ecalj_dev_zero(ecalj_handle h, void* ptr, size_t bytes);
/* dst = src by a plain vector-copy kernel (16-byte aligned, bytes % 16 == 0):
 * asynchronous on the handle's stream; also the reference kernel of known
 * traffic for HBM-counter calibration. */
ecalj_status ecalj_dev_copy(ecalj_handle h, void* dst, const void* src, size_t bytes);
ecalj_status ecalj_dev_phase_similarity(ecalj_handle h, const void* a, int ngb, int nmat,
                                        const void* d, int nk, void* out);

/* ---- module state (set once per q / rotation, kept on the device) ------ */
/* set_ppb (S/m_zmel.f90:110-145): real radial-angular integrals
 * ppb(i1,i2,I) of each atom, blocks of nv[a] x nv[a] x mdim[a] concatenated.
 * Atom a owns eigenvector rows [sum nv[<a], +nv[a]) and product-basis rows
 * [sum mdim[<a], +mdim[a]).  Host pointer. */
ecalj_status ecalj_set_ppb(ecalj_handle h, int natom, const int* nv, const int* mdim,
                           const double* ppb);
/* Readvcoud + set_m2e_prod_basis (S/m_readVcoud.f90:11-43,
 * S/m_zmel.f90:46-68): eigenvalues vcoud(ngb) and eigenvectors
 * zcousq(ngb,ngb) of the Coulomb matrix; vcousq = sqrt(vcoud).  Host pointers. */
ecalj_status ecalj_set_coulomb(ecalj_handle h, int ngb, const double* vcoud, const double* zcousq);
/* zhgv_h / zhgv_d (S/m_lapack.f90:64-104, 203-229), the Coulomb-matrix
 * diagonalisation of hvccfp0 (S/main_hvccfp0.f90:322-368): the
 * Hermitian-definite generalized eigenproblem a z = lambda b z (itype 1,
 * upper triangles of a and b read), FP64 throughout on the device.
 * Eigenvalues il..iu (1-based, of the ascending list; the reference passes
 * 1..n) go to evl(1:m), m = iu - il + 1, and their b-orthonormal
 * eigenvectors (z^H b z = 1) overwrite a(:, 1:m); b is left unchanged.
 * hvccfp0 passes -vcoul to get the Coulomb eigenvalues in descending order,
 * then stores vcoud = -evl and zcousq = z (feed them to ecalj_set_coulomb).
 * ECALJ_ERR_NOT_PD when b is not positive definite (zhegvx: info > n),
 * ECALJ_ERR_NOT_CONVERGED if the Jacobi sweeps do not converge.  Waits for
 * the stream; not under graph capture.  ecalj_zhgv: host pointers;
 * ecalj_zhgv_dev: device pointers (evl too). */
ecalj_status ecalj_zhgv(ecalj_handle h, double* a, const double* b, int n, double* evl, int il, int iu,
                        int lda, int ldb);
ecalj_status ecalj_zhgv_dev(ecalj_handle h, void* a, const void* b, int n, double* evl, int il, int iu,
                            int lda, int ldb);

/* ---- zmel ---------------------------------------------------------------
 * get_zmel_init_gemm (S/m_zmel.f90:146-597), MT valence part with identity
 * rotation: zmel(I,n,m) = phase_a * sum_{i1,i2} conj(cphim(i2,n)) ppb(i1,i2,I) cphiq(i1,m)
 * over the atom blocks; then zmel = zcousq^H zmel when is_m_basis == 0
 * (:559-560), then conj when zmelconjg (:590-594).
 * cphim: ndima x nm, cphiq: ndima x nq; zmel: ngb x nm x nq.
 * phase: natom complex values (exp(-i 2pi kvec.tr)).
 * nbatch > 1 evaluates nbatch independent problems with cphim, cphiq and
 * zmel advanced by their own sizes (one call for all k-points of a q). */
ecalj_status ecalj_zmel(ecalj_handle h, const double* phase, int ndima, int nm, const double* cphim,
                        int nq, const double* cphiq, int is_m_basis, int zmelconjg, int nbatch,
                        double* zmel);
ecalj_status ecalj_zmel_dev(ecalj_handle h, const double* phase, int ndima, int nm,
                            const void* cphim, int nq, const void* cphiq, int is_m_basis,
                            int zmelconjg, int nbatch, void* zmel);
/* As ecalj_zmel_dev with explicit batch strides (elements) of cphim and
 * cphiq; a stride of 0 shares one block across the batch (the end states at
 * q are the same for every internal k of Sigma_c, S/m_sxcf_gemm.f90:395). */
ecalj_status ecalj_zmel_strided_dev(ecalj_handle h, const double* phase, int ndima, int nm,
                                    const void* cphim, int64_t cphim_stride, int nq,
                                    const void* cphiq, int64_t cphiq_stride, int is_m_basis,
                                    int zmelconjg, int nbatch, void* zmel);
/* set_ppb (S/m_zmel.f90:110-145) from the reference's own array
 * ppb(nlnmx, nlnmx, mdimx, natom): atom a's functions are nlnmc[a] core
 * then nlnmv[a] valence, with nblocha[a] product-basis functions.  Keeps
 * the valence blocks (as ecalj_set_ppb) and the core-valence blocks
 * ppb(nc+1:nc+nv, 1:nc, :, a) for the CoreValence rows of zmel. */
ecalj_status ecalj_set_ppb_full(ecalj_handle h, int natom, int nlnmx, int mdimx, const int* nlnmc,
                                const int* nlnmv, const int* nblocha, const double* ppb);
/* get_zmel_init_gemm's muffin-tin part over a middle-state range with core
 * states (S/m_zmel.f90:146-147, 199-214, 323-391, 559-560, 590-594):
 * middle states ns1..ns2 (1-based) run over the nctot core states first
 * (atom by atom, nlnmc of ecalj_set_ppb_full each) and then the valence
 * bands; cphim holds only the valence bands of the range (nm1v..nm2v of the
 * reference, ndima x nmv), cphiq the end states (ndima x nq).  Core rows:
 * zmel(ims+I, it, m) = phase(ia) sum_b ppb(nc+b, it-ics, I) cphiq(ias+b, m).
 * iatomp (natom, 1-based, NULL = identity; miat(:, invr) of the reference)
 * maps atom ia to the atom whose product basis, ppb block and row offset it
 * uses; ia's own eigenvector rows and core states stay.  zmel: ngb x
 * (ns2-ns1+1) x nq (x nbatch; cphim and cphiq advance by their sizes).
 * ECALJ_ERR_STATE when the range holds core states but nctot differs from
 * the core count of ecalj_set_ppb_full. */
ecalj_status ecalj_zmel_init(ecalj_handle h, const double* phase, const int* iatomp, int ndima,
                             int ns1, int ns2, int nctot, const double* cphim, int nq,
                             const double* cphiq, int is_m_basis, int zmelconjg, int nbatch,
                             double* zmel);
ecalj_status ecalj_zmel_init_dev(ecalj_handle h, const double* phase, const int* iatomp, int ndima,
                                 int ns1, int ns2, int nctot, const void* cphim, int nq,
                                 const void* cphiq, int is_m_basis, int zmelconjg, int nbatch,
                                 void* zmel);

/* ---- zmel: interstitial plane-wave rows ---------------------------------
 * m_read_ppovl tables of the current q (S/m_zmel.f90:419): ggg (nggg complex,
 * Fourier coefficients of the interstitial step function) with its index
 * cube igggi (1-based into ggg, 0 = none; Fortran order over
 * [lo, lo+n) per axis), the combined G list nvgcgp2 (3 x ngcgp) with its
 * index cube igcgp2i.  Host pointers; kept on the device. */
ecalj_status ecalj_set_ppovl(ecalj_handle h, int nggg, const double* ggg, const int* ggg_lo,
                             const int* ggg_n, const int* igggi, int ngcgp, const int* nvgcgp2,
                             const int* gcp_lo, const int* gcp_n, const int* igcgp2i);
/* ZmelIPW block of get_zmel_init_gemm (S/m_zmel.f90:418-538):
 *   zmel(row0+c, n, m) = phase(c) sum_{G1,G2} ggg(G1 - Gc - G2 - nadd) geigq(G1, m) dgeigqk(G2, n)
 * for c < ngc (Gc = ngveccR(:,c)), n < nm, m < ntp0; G1 over ngvecpB1
 * (3 x ngp1, the PW basis of q), G2 over ngvecpB2 (3 x ngp2, of q-k).
 * geigq: ngpmx x ntp0, dgeigqk: ngpmx x nm (complex128; device in _dev);
 * zmel: ldz x nm x ntp0 in the precision's storage type (_dev) or complex128
 * (host), only rows row0..row0+ngc-1 written.  phase: ngc complex (host).
 * Contraction order as the reference: G1 first when nm > ntp0.  Every
 * Gc + G2 (G1 first) or Gc - G1 (G2 first) must be in the combined list
 * (ECALJ_ERR_ARG otherwise). */
ecalj_status ecalj_zmel_ipw(ecalj_handle h, int ngc, const int* ngveccR, const int* nadd,
                            const double* phase, int ngpmx, int ngp1, const int* ngvecpB1,
                            const double* geigq, int ntp0, int ngp2, const int* ngvecpB2,
                            const double* dgeigqk, int nm, double* zmel, int ldz, int row0);
ecalj_status ecalj_zmel_ipw_dev(ecalj_handle h, int ngc, const int* ngveccR, const int* nadd,
                                const double* phase, int ngpmx, int ngp1, const int* ngvecpB1,
                                const void* geigq, int ntp0, int ngp2, const int* ngvecpB2,
                                const void* dgeigqk, int nm, void* zmel, int ldz, int row0);
/* The rows of zmelt (MT rows 0..nbloch-1 from ecalj_zmel_dev with
 * is_m_basis = 1, IPW rows nbloch..nbloch+ngc-1 from ecalj_zmel_ipw_dev) put
 * together, then zmel = zcousq^H zmelt when is_m_basis == 0
 * (S/m_zmel.f90:559-560; ngb of ecalj_set_coulomb = nbloch + ngc) and conj
 * when zmelconjg (:590-594).  mt: nbloch x ncol, ipw: ngc x ncol, zmel:
 * (nbloch+ngc) x ncol, device arrays in the storage type. */
ecalj_status ecalj_zmel_assemble_dev(ecalj_handle h, const void* mt, int nbloch, const void* ipw,
                                     int ngc, int64_t ncol, int is_m_basis, int zmelconjg,
                                     void* zmel);

/* ---- chi0: x0gemm -----------------------------------------------------
 * x0gemm (S/x0gemm.f90:2-97) with its module inputs explicit
 * (m_x0kf: icounkmink/icounkmaxk, iwini, iwend, itc, itpc, jpmc, icouini,
 * whwc; m_zmel: zmel).  For every icoun in [icoun_lo, icoun_hi] with
 * ns1 <= itc <= ns2 and every bin iw in [iwini, iwend]:
 *   rcxq(:,:,iw*(3-2*jpm)) += conj(zmel(1:npr,it,itp)) (x) whw * zmel(ipr_col:ipr_col+npr_col-1,it,itp)
 * rcxq: npr x npr_col x ((npm==1) ? nwhis+1 : 2*nwhis+1), Fortran lower
 * bound (1-npm)*nwhis; zmel: nbb x (ns2-ns1+1) x nqtot.
 * The _batched form accumulates nk such problems (one per k-point, zmel
 * blocks of zmel_stride elements, index ranges icounkmin[k]..icounkmax[k],
 * ns1[k]/ns2[k]) in one pass; this is the whole k loop of x0kf_zxq
 * (S/x0kf_v4h.f90:233-290).  In the _dev form rcxq and zmel are device
 * pointers; the index arrays stay on the host, where x0kf_v4hz_init builds
 * them (S/x0kf_v4h.f90:48-142), and are turned into per-bin gather lists. */
ecalj_status ecalj_x0gemm(ecalj_handle h, double* rcxq, int npr, int ipr_col, int npr_col,
                          int nwhis, int npm, int ns1, int ns2, const double* zmel, int nbb,
                          int nqtot, int icoun_lo, int icoun_hi, const int* iwini,
                          const int* iwend, const int* itc, const int* itpc, const int* jpmc,
                          const int* icouini, const double* whwc);
ecalj_status ecalj_x0gemm_batched_dev(ecalj_handle h, void* rcxq, int npr, int ipr_col,
                                      int npr_col, int nwhis, int npm, int nk, const int* ns1,
                                      const int* ns2, const void* zmel, int64_t zmel_stride,
                                      int nbb, int nqtot, const int* icounkmin,
                                      const int* icounkmax, int ncoun, const int* iwini,
                                      const int* iwend, const int* itc, const int* itpc,
                                      const int* jpmc, const int* icouini, int ncount,
                                      const double* whwc);

/* ---- chi0: tetrahedron histogram weights --------------------------------
 * gettetwt + tetwt5x_dtet4 (S/m_tetwt.f90:36-256, S/tetwt5.f90:9-486) for
 * npm=1 (time reversal), no core states, tetrahedron_matrix_linear=F: the
 * (occupied n at k, unoccupied m at k+q) pairs with tetrahedron weight and
 * their histogram weights whw on the bins frhis(1:nwhis+1) (Hartree), the
 * input of x0kf_v4hz_init.  idtetf: 4 x ntetf 1-based indices into qbzw
 * (3 x nqbzw); ib1bz: nqbzw 1-based k indices; ekxx1, ekxx2: nband x nqbz
 * energies (Ry) at k and k+q; ef (Ry); ebmx: band cutoff (Ry; 1e10 = none);
 * qlat: reciprocal vectors (columns).  Outputs, Fortran layout with leading
 * dimension nbnbx_cap: n1b, n2b (1-based bands), ihw, nhw (first bin,
 * count), jhw (1-based pointer into whw); nbnb(nqbz); whw(nhw_cap).
 * nbnbx_out / nhwtot_out receive the sizes needed; when a capacity is short
 * the call returns ECALJ_ERR_DIM after setting them (call again with
 * enough room; output pointers may then be NULL on the first call).
 * Index maps are exact (same as a sequential loop); whw is summed with FP64
 * atomics (order-dependent rounding).  Host pointers; not capturable. */
ecalj_status ecalj_tetwt(ecalj_handle h, int nband, int nqbz, int ntetf, const int* idtetf,
                         int nqbzw, const double* qbzw, const int* ib1bz, const double* ekxx1,
                         const double* ekxx2, double ef, double ebmx, const double* qlat,
                         const double* frhis, int nwhis, int nbnbx_cap, int64_t nhw_cap,
                         int* nbnb, int* n1b, int* n2b, int* ihw, int* nhw, int* jhw,
                         double* whw, int* nbnbx_out, int64_t* nhwtot_out);
/* gettetwt in full (S/m_tetwt.f90:36-256 with its module inputs npm from
 * m_freq and nctot, ecore(:,is) from m_genallcf_v3; S/tetwt5.f90:179-185,
 * 252-264 for the core states, 349-427 for their degenerate groups):
 * npm = 2 adds the jpm = 2 pairs (a state at k, an occupied one at k+q) of
 * the time-reversal-off chi0; nctot core states of energies ecore (Ry) are
 * appended after the valence bands, flat in k, on the occupied side
 * (n1b up to nband + nctot; with npm = 2, n2b up to nband + nctot for
 * jpm = 2).  Outputs gain a trailing jpm axis: nbnb(nqbz, npm); n1b, n2b,
 * ihw, nhw, jhw (nbnbx_cap, nqbz, npm); whw pointers run over jpm, then k,
 * then the pairs.  matrix_linear != 0 is the GWinput switch
 * tetrahedron_matrix_linear (S/switch.f90:253-263; S/tetwt5.f90:334-336,
 * 694-709, 751-759, mkwcg :788-840): each bin's weight goes to the four
 * corner k-points of a tetrahedron in the barycentric shares of the
 * cross-section's vertex average at the bin centre instead of equally; the
 * lists are unchanged.  ecalj_tetwt is this call with npm = 1, nctot = 0,
 * matrix_linear = 0. */
ecalj_status ecalj_tetwt_full(ecalj_handle h, int nband, int nqbz, int ntetf, const int* idtetf,
                              int nqbzw, const double* qbzw, const int* ib1bz,
                              const double* ekxx1, const double* ekxx2, double ef, double ebmx,
                              const double* qlat, const double* frhis, int nwhis, int npm,
                              int nctot, const double* ecore, int matrix_linear, int nbnbx_cap,
                              int64_t nhw_cap,
                              int* nbnb, int* n1b, int* n2b, int* ihw, int* nhw, int* jhw,
                              double* whw, int* nbnbx_out, int64_t* nhwtot_out);

/* ---- chi0: Hilbert transform to the imaginary axis ---------------------
 * dpsion_chiq, imaginary branch with npm=1 (S/dpsion5.f90:211-232) with
 * dpsion_init's imatt (S/dpsion5.f90:84-89,121-129, hilbertmat
 * S/hilbertmat.f90:14-79) computed in FP64 on the host:
 *   zxqi(:,:,iv) = sum_iw imatt(iv,iw) * (-wfac(iw)/(R-L)) * rcxq(:,:,iw)
 * frhis(1:nwhis+1) bin edges, freqi(1:niw) imaginary frequencies.
 * When scale_rcxq != 0 rcxq is scaled in place as the reference does. */
ecalj_status ecalj_dpsion_imag(ecalj_handle h, double* rcxq, int npr, int npr_col, int nwhis,
                               int niw, const double* frhis, const double* freqi, double ecut,
                               int scale_rcxq, double* zxqi);
ecalj_status ecalj_dpsion_imag_dev(ecalj_handle h, void* rcxq, int npr, int npr_col, int nwhis,
                                   int niw, const double* frhis, const double* freqi,
                                   double ecut, int scale_rcxq, void* zxqi);

/* ---- chi0: Hilbert transform to the real axis --------------------------
 * dpsion_chiq, real branch with npm=1 and no chipm (S/dpsion5.f90:193-200,
 * 244-256) with dpsion_init's rmatt (S/dpsion5.f90:91-103) computed in FP64
 * on the host from hilbertmat at zz = freq_r(it):
 *   zxq(:,:,it) = i*s(it)*rcxq(:,:,it) + sum_iw rmatt(it,iw)*s(iw)*rcxq(:,:,iw),
 *   it = 0..nw,  s(iw) = -wfac(iw)/(R-L) for iw >= 1, s(0) = 1
 * (rcxq slice it is the bin whose centre is freq_r(it); slice 0 is the
 * zero-frequency slot).  rcxq_scaled != 0: rcxq already holds the scaled
 * values (e.g. after ecalj_dpsion_imag with scale_rcxq), s = 1.  rcxq is
 * read only; zxq (npr x npr_col x (nw+1), complex128) is a separate array,
 * where the reference overwrites rcxq(:,:,0:nw).  Requires nw < nwhis and
 * freqr[0] == 0 (the reference's own checks, S/dpsion5.f90:203-204). */
ecalj_status ecalj_dpsion_real(ecalj_handle h, const double* rcxq, int npr, int npr_col,
                               int nwhis, int nw, const double* frhis, const double* freqr,
                               double ecut, int rcxq_scaled, double* zxq);
ecalj_status ecalj_dpsion_real_dev(ecalj_handle h, const void* rcxq, int npr, int npr_col,
                                   int nwhis, int nw, const double* frhis,
                                   const double* freqr, double ecut, int rcxq_scaled,
                                   void* zxq);

/* ---- chi0: composed direct-frequency form ------------------------------
 * chi(I,J,v) = sum_p w(v,p) conj(Z(I,p)) Z(J,p); equal to x0gemm followed by
 * dpsion_chiq when w(v,p) = sum_bins whw * (-wfac/(R-L)) * imatt(v,bin). */
ecalj_status ecalj_chi0_freq(ecalj_handle h, const double* z, const double* w, int ni,
                             int npairs, int nf, double* chi);
ecalj_status ecalj_chi0_freq_dev(ecalj_handle h, const void* z, const void* w, int ni,
                                 int npairs, int nf, void* chi);

/* ---- W_c build ---------------------------------------------------------
 * WVIllwI body for q /= 0 (S/m_llw.f90:299-373), per frequency:
 *   chi *= 2 when nspin == 1; [chi = zcousq^H chi zcousq];
 *   eps = delta - vcousq_i chi_ij vcousq_j; eps^-1 (zminv, S/m_lapack.f90:21-47,108-176);
 *   wc = vcousq_i eps^-1_ij vcousq_j - delta_ij vcoud_i; [wc = zcousq wc zcousq^H].
 * chi: ngb x ngb x nw; wc: ldw x ngb x nw.  The inversion runs without
 * pivoting: epsilon on the imaginary axis has a positive definite Hermitian
 * part (chi is negative semidefinite), for which elimination without
 * pivoting is stable; a pivot below 1e-300 returns ECALJ_ERR_SINGULAR. */
ecalj_status ecalj_wc_build(ecalj_handle h, const double* chi, int ngb, int nw, int nspin,
                            int is_x0_m_basis, int is_wc_m_basis, double* wc, int ldw);
ecalj_status ecalj_wc_build_dev(ecalj_handle h, const void* chi, int ngb, int nw, int nspin,
                                int is_x0_m_basis, int is_wc_m_basis, void* wc, int ldw);

/* WVRllwR body for q /= 0 (S/m_llw.f90:88-180): the same steps on the real
 * frequency mesh (nw = number of frequencies 0..nw_max, i.e. nw_max+1).
 * chi(omega) is not Hermitian there, so epsilon is a general complex
 * matrix: the inversion uses partial pivoting (row exchanges chosen as
 * zgetrf does, first row of largest |re|+|im|); a zero pivot column returns
 * ECALJ_ERR_SINGULAR. */
ecalj_status ecalj_wc_build_real(ecalj_handle h, const double* chi, int ngb, int nw, int nspin,
                                 int is_x0_m_basis, int is_wc_m_basis, double* wc, int ldw);
ecalj_status ecalj_wc_build_real_dev(ecalj_handle h, const void* chi, int ngb, int nw,
                                     int nspin, int is_x0_m_basis, int is_wc_m_basis,
                                     void* wc, int ldw);

/* zminv_h / zminv_d (S/m_lapack.f90:21-47,108-176): in-place inverse of a
 * general complex128 matrix (lda x n), with partial pivoting; the _dev form
 * inverts nbatch matrices `stride` elements apart in device memory.  The
 * plugin seam the reference aliases at compile time (S/m_llw.f90:25-29). */
ecalj_status ecalj_zminv(ecalj_handle h, double* a, int n, int lda);
ecalj_status ecalj_zminv_dev(ecalj_handle h, void* a, int n, int lda, int nbatch,
                             int64_t stride);
/* ecalj_zminv_dev without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (HIP graph capture with
 * no warm-up call, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as ecalj_zminv_dev; its work space comes from
 * the device's fixed array of ECALJ_NOALLOC_SCRATCH_BYTES (64 MiB): about
 * 16 n^2 + 6.3 kB n + 0.3 MB per matrix above n = 128, so one matrix up to
 * n = 1859, 6 of n = 600 or 2 of n = 1200 per call.  Larger requests return
 * ECALJ_ERR_SCRATCH before any work is queued.  All calls on one device share
 * the array: calls that overlap in time on one device (two streams, two
 * handles) race and must be serialised by the caller.  Returns
 * ECALJ_ERR_STATE if the array's device address could not be taken when the
 * first handle on the device was created.  Asynchronous; a zero pivot is
 * reported as for ecalj_zminv_dev. */
ecalj_status ecalj_zminv_noalloc_dev(ecalj_handle h, void* a, int n, int lda, int nbatch,
                                     int64_t stride);

/* ---- the Gamma point of the W build ------------------------------------
 * q = 0 body (iq = 1) of WVIllwI / WVRllwR (S/m_llw.f90:98-178, 299-373 with
 * ix = 1): epsilon restricted to mu, nu >= 2 is inverted (partial pivoting
 * when real_axis), W(1,:) = W(:,1) = 0, and W stays in the E basis (the
 * E -> M step skips iq = 1; modifyWV0 applies it).  Same arguments as
 * ecalj_wc_build otherwise; the Coulomb state is that of q = 0 (vcousq(1)
 * is not used). */
ecalj_status ecalj_wc_build_gamma(ecalj_handle h, const double* chi, int ngb, int nw, int nspin,
                                  int is_x0_m_basis, int real_axis, double* wc, int ldw);
ecalj_status ecalj_wc_build_gamma_dev(ecalj_handle h, const void* chi, int ngb, int nw, int nspin,
                                      int is_x0_m_basis, int real_axis, void* wc, int ldw);
/* Offset points q0i near Gamma (iq > nqibz) of WVRllwR / WVIllwI
 * (S/m_llw.f90:183-256, 378-440): epsilon with the head 1 - vcou1*chi(1,1)
 * (vcou1 = 4 pi/|q|^2, :185, :380), full inverse (pivoted when real_axis),
 * llw(iw) = 1/epsinv(1,1) (complex, nw); when 0 <= wing_iw < nw and wing is
 * given, wing(mu-1) = epsinv(1,mu)/epsinv(1,1), mu = 2..ngb, at that
 * frequency (wmuk, :238-244).  The Coulomb state is that of q0i. */
ecalj_status ecalj_eps_head(ecalj_handle h, const double* chi, int ngb, int nw, int nspin, int is_x0_m_basis,
                            double vcou1, int real_axis, double* llw, int wing_iw, double* wing);
ecalj_status ecalj_eps_head_dev(ecalj_handle h, const void* chi, int ngb, int nw, int nspin,
                                int is_x0_m_basis, double vcou1, int real_axis, void* llw, int wing_iw,
                                void* wing);
/* getw0 (S/m_w0w0i.f90:146-311), host FP64, no handle: the Gamma-cell
 * average of W - v, w0(ii:ie), from llw(ii:ie, nq0i) (complex, leading
 * dimension ie-ii+1), dmlx(nq0i, 9), epinvq0i(nq0i, nq0i),
 * wklm((lmxax+1)^2), wbz = wbz(1), epinv(3, 3, nq0i); llmat(3, 3) complex
 * (iw = 0) when ii <= 0 and llmat is given.  lmxax <= 12. */
ecalj_status ecalj_getw0(int ii, int ie, int nq0i, int lmxax, const double* llw, const double* dmlx,
                         const double* epinvq0i, const double* wklm, double wbz, const double* epinv,
                         double* w0, double* llmat);
/* modifyWV0 (S/m_w0w0i.f90:89-140) on W records held in memory (ldw x ngb x
 * nw, as written by ecalj_wc_build_gamma): W(1,1,iw) = w0(iw) (host array,
 * complex nw), then W <- zcousq W zcousq^H when is_wc_m_basis. */
ecalj_status ecalj_wv0_patch(ecalj_handle h, double* w, int ngb, int ldw, int nw, const double* w0,
                             int is_wc_m_basis);
ecalj_status ecalj_wv0_patch_dev(ecalj_handle h, void* w, int ngb, int ldw, int nw, const double* w0,
                                 int is_wc_m_basis);

/* ---- eigenfunctions on the full mesh (readcphi / readgeig) -------------
 * m_hamindex state (S/readeigen.f90:11-12, from HAMindex0 and __QGpsi):
 * qtt(3,nqtt); iqmap/igmap/iqimap(nqtt) 1-based (the point each q is
 * rotated from, the operation, its index among the irreducible points);
 * ngp(nqtt); ngvecp(3,ngpmx,nqtt) and ngvecprev((2imx+1)^3,nqtt) (1-based
 * indices; may be NULL when ngpmx = 0); symops(3,3,ngrp), shtvg(3,ngrp),
 * miat(nbas,ngrp) 1-based, tiat(3,nbas,ngrp), dlmm(-lmxax:lmxax,
 * -lmxax:lmxax,0:lmxax,ngrp) (rotdlmm's real-harmonic D^l), plat, qlat. */
ecalj_status ecalj_set_hamindex(ecalj_handle h, int nqtt, const double* qtt, const int* iqmap, const int* igmap,
                                const int* iqimap, const int* ngp, int ngpmx, int imx, const int* ngvecp,
                                const int* ngvecprev, int ngrp, const double* symops, const double* shtvg,
                                int nbas, const int* miat, const double* tiat, int lmxax, const double* dlmm,
                                const double* plat, const double* qlat);
/* readmnla_cphi (S/readeigen.f90:487-521): per orbital block (ibas, l, k)
 * 1-based and offset = first row - 1 of the block in cphi. */
ecalj_status ecalj_set_mnla(ecalj_handle h, int norbt, const int* ibas_tbl, const int* l_tbl, const int* k_tbl,
                            const int* offset_tbl);
/* readcphi (S/readeigen.f90:140-181, rotmto S/rotwave.f90:156-192) for nk
 * target points iq(nk) (host, 1-based into qtt) at once: cphi_irr holds the
 * irreducible eigenvectors of one spin, ldc x nband per irreducible point
 * (ldc >= ndima*nspc), cphi_out ldc x nband per target.  Device pointers. */
ecalj_status ecalj_readcphi_dev(ecalj_handle h, int nk, const int* iq, int nspc, int nband, const void* cphi_irr,
                                int ldc, void* cphi_out);
/* readgeig (S/readeigen.f90:71-131, rotipw S/rotwave.f90:223-245): the same
 * for the plane-wave coefficients, ldg >= ngpmx*nspc rows; rows without a
 * rotated G are zero.  A rotated G outside the target's list, or ngp(iq) /=
 * ngp(iqmap(iq)), returns ECALJ_ERR_DIM. */
ecalj_status ecalj_readgeig_dev(ecalj_handle h, int nk, const int* iq, int nspc, int nband, const void* geig_irr,
                                int ldg, void* geig_out);
/* Readers of the front end's files (host, no handle except rec_read_dev).
 * @MNLA_CPHI (S/sugw.f90:261-277): first call with cap = 0 reports norbt
 * (ECALJ_ERR_DIM), then with room.
 * __QGpsi (S/mkqg.f90:298,349-356): hdr6 = nqnum ngpmx nqbz nqi imx nqibz and
 * QpGcut_psi; first call with capacities 0 reports them; then qtt(3,nqnum),
 * ngp, irr (nqnum), ngvecp(3,ngpmx,nqnum), ngvecprev((2imx+1)^3,nqnum).
 * __EValue (S/readeigen.f90:428-440): hdr5 = nband nqi nsp nspc nspx,
 * qtti(3,nqi), evud(nband,nqi,nspx); cap = elements of evud.
 * Fixed-length records (__CPHI recl mrecb, __GEIG recl mrecg,
 * S/sugw.f90:234-236): records rec0, rec0+step, ..., the first nbytes of
 * each, to device memory dst + i*dst_stride bytes. */
ecalj_status ecalj_mnla_read(const char* path, int cap, int* norbt, int* ibas_tbl, int* l_tbl, int* k_tbl,
                             int* offset_tbl);
ecalj_status ecalj_qgpsi_read(const char* path, int* hdr6, double* qpgcut, int nq_cap, int ngpmx_cap, int imx_cap,
                              double* qtt, int* ngp, int* irr, int* ngvecp, int* ngvecprev);
ecalj_status ecalj_evalue_read(const char* path, int* hdr5, int64_t cap, double* qtti, double* evud);
ecalj_status ecalj_rec_read_dev(ecalj_handle h, const char* path, int64_t recl, int rec0, int step, int nrec,
                                int64_t nbytes, void* dst, int64_t dst_stride);

/* ---- on-disk records shared with the Fortran steps ----------------------
 * __Vcoud.iq (S/main_hvccfp0.f90:396-401, read by S/m_readVcoud.f90:36-43):
 * sequential unformatted records ngb; q(3); vcoud(ngb); zcousq(ngb,ngb)
 * (4-byte record markers).  _read: first call with ngb_cap = 0 reports ngb
 * (ECALJ_ERR_DIM), then with room.  No handle needed. */
ecalj_status ecalj_vcoud_write(const char* path, int ngb, const double* q, const double* vcoud,
                               const double* zcousq);
ecalj_status ecalj_vcoud_read(const char* path, int* ngb, double* q, double* vcoud, double* zcousq,
                              int ngb_cap);
/* SEC2U (S/main_hsfp0.sc.f90:339-359, read by S/main_hqpe.sc.f90:70):
 * header8 = nspin nq ntq nqbz nqibz n1 n2 n3, then per q: is, q(3),
 * zsec(ntq,ntq) complex(8).  _read reports the header first (capacities 0). */
ecalj_status ecalj_sec2_write(const char* path, const int* header8, int nq, int ntq, const int* is,
                              const double* q, const double* zsec);
ecalj_status ecalj_sec2_read(const char* path, int* header8, int nq_cap, int ntq_cap, int* is,
                             double* q, double* zsec);
/* __WV.d (S/main_hrcxq.f90:134-136): nprecx mrecl nblochpmx nw+1 niw nq nw_i. */
ecalj_status ecalj_wv_d_write(const char* path, const int64_t* v7);
ecalj_status ecalj_wv_d_read(const char* path, int64_t* v7);
/* __WVI.iq / __WVR.iq (S/m_llw.f90:102,177,308,373; S/m_mpiio.f90:12-74):
 * fixed-length records of W(nblochpmx, nblochpmx) complex(8), record rec at
 * byte (rec-1)*16*ld^2.  nrec records rec0.. to / from device memory (ld x
 * ld each, contiguous), through pinned staging; synchronous. */
ecalj_status ecalj_wv_write_dev(ecalj_handle h, const char* path, int rec0, int nrec, int ld,
                                const void* w);
ecalj_status ecalj_wv_read_dev(ecalj_handle h, const char* path, int rec0, int nrec, int ld, void* w);
/* The complex(4) records of the reference's mixed-precision build (__MP:
 * kindgw = kindrcxq = 4, S/m_kind.f90:3-6; mrecl = 8*nblochpmx**2 bytes,
 * S/rdpp.f90:43; read by S/m_sxcf_gemm.f90:317-318 with that recl).  Device W
 * is complex128 (ld x ld x nrec); elements are rounded to complex(4) on write
 * and widened exactly on read.  Synchronous. */
ecalj_status ecalj_wv4_write_dev(ecalj_handle h, const char* path, int rec0, int nrec, int ld,
                                 const void* w);
ecalj_status ecalj_wv4_read_dev(ecalj_handle h, const char* path, int rec0, int nrec, int ld, void* w);

/* ---- Sigma_c ------------------------------------------------------------
 * Imaginary-axis weights (S/m_sxcf_gemm.f90:432-458), FP64:
 *   wgtim: (niw+1) x (ns2-ns1+1) x ntq, times wkkr.
 * omega(ntq) end-state and ekc(ns1:ns2) middle-state energies in Ry; esmr in
 * Ry; freqx/wwx: Gauss-Legendre points/weights on (0,1); expa = exp(-ua^2 w^2). */
ecalj_status ecalj_sigc_weights(ecalj_handle h, int niw, int ns1, int ns2, int ntq, int nctot,
                                const double* omega, const double* ekc, double esmr, double ua,
                                double wkkr, const double* freqx, const double* wwx,
                                const double* expa, double* wgtim);
ecalj_status ecalj_sigc_weights_dev(ecalj_handle h, int nbatch, int niw, int ns1, int ns2,
                                    int ntq, int nctot, const double* omega, const double* ekc,
                                    double esmr, double ua, const double* wkkr,
                                    const double* freqx, const double* wwx, const double* expa,
                                    void* wgtim);
/* Imaginary-axis contraction and final product (S/m_sxcf_gemm.f90:463-496,
 * 599-608):
 *   czmelwc = sum_iw wc(iw)^H (wgtim(iw,:,:) * zmel), wc(iw) unpacked from its
 *             upper triangle as the keepwv path does (:337-344, 468-475);
 *   zsec(1:ntq,1:ntq) += czmelwc^H zmel  (K = ngb*nm, ldc = ldz);
 *   Im zsec(i,i) clipped to <= 0.
 * zmel: ngb x nm x ntq; wc: ldw x ngb x (niw+1); wgtim (niw+1) x nm x ntq.
 * nbatch > 1 accumulates nbatch internal k-points (zmel, wc, wgtim advanced
 * by their sizes) into the same zsec; the clip is applied once at the end. */
ecalj_status ecalj_sigc(ecalj_handle h, const double* zmel, int ngb, int nm, int ntq, int niw,
                        const double* wc, int ldw, const double* wgtim, double* zsec, int ldz);
ecalj_status ecalj_sigc_dev(ecalj_handle h, int nbatch, const void* zmel, int ngb, int nm,
                            int ntq, int niw, const void* wc, int ldw, const void* wgtim,
                            void* zsec, int ldz);
/* Real-axis pole term of Sigma_c (S/m_sxcf_gemm.f90:501-594, then :599-608),
 * added to zsec with its own final product (the term is linear in czmelwc,
 * so this equals adding it to czmelwc before the final GEMM), then
 * Im(zsec_mm) <= 0.  zmel: ngb x nm x ntq per internal k (middle states
 * 1..nm, the first nctot are core); wr: ldw x ngb x (nw+1) complex128 records
 * W(iw), iw = 0..nw, per k at element stride wr_stride (0 = shared), used
 * through their upper triangle (keepwv); freq_r(0..nw) Hartree; omega (ntq),
 * ekc (nm per k), ef, esmr in Ry; ns2r: last middle state of the real-axis
 * range (sxcf_scz_count's nstte2); wkkr: k weights.  Host arrays except zmel,
 * wr, zsec.  A pole frequency outside freq_r(1..nw-1) returns ECALJ_ERR_DIM. */
ecalj_status ecalj_sigc_real_dev(ecalj_handle h, int nbatch, const void* zmel, int ngb, int nm, int ntq,
                                 const void* wr, int ldw, int64_t wr_stride, int nw, const double* freq_r,
                                 const double* omega, const double* ekc, double ef, double esmr, int nctot,
                                 int ns2r, const double* wkkr, void* zsec, int ldz);
/* get_correlation_block in full (S/m_sxcf_gemm.f90:423-608): the imaginary-
 * axis contraction of ecalj_sigc_dev and the real-axis pole term of
 * ecalj_sigc_real_dev added into the same czmelwc, one final product per
 * internal k, Im(zsec_mm) <= 0 after every k -- the reference's order. */
ecalj_status ecalj_sigc_full_dev(ecalj_handle h, int nbatch, const void* zmel, int ngb, int nm, int ntq,
                                 int niw, const void* wc, int ldw, const void* wgtim, const void* wr,
                                 int ldwr, int64_t wr_stride, int nw, const double* freq_r,
                                 const double* omega, const double* ekc, double ef, double esmr,
                                 int nctot, int ns2r, const double* wkkr, void* zsec, int ldz);
ecalj_status ecalj_sigc_real(ecalj_handle h, const double* zmel, int ngb, int nm, int ntq, const double* wr,
                             int ldw, int nw, const double* freq_r, const double* omega, const double* ekc,
                             double ef, double esmr, int nctot, int ns2r, double wkkr, double* zsec, int ldz);
/* Sigma_c over internal k-points whose middle-state windows differ, the
 * per-k band ranges that sxcf_scz_count assigns (nstti/nstte,
 * S/sxcf_fal2_count.sc.f90:118-236), in one call: k has nm[k] middle states
 * (host array); zmel_k (ngb x nm[k] x ntq, precision storage type) and
 * wgtim_k ((niw+1) x nm[k] x ntq, FP64) are packed one after the other; W_c(k)
 * (ldw x ngb x (niw+1), complex128) are consecutive.  zsec accumulates all
 * k, then Im(zsec_mm) <= 0.  The bf16x3 form runs one grouped persistent
 * kernel over the (k, row tile, column tile) list. */
ecalj_status ecalj_sigc_grouped_dev(ecalj_handle h, int nk, const int* nm, const void* zmel, int ngb,
                                    int ntq, int niw, const void* wc, int ldw, const void* wgtim,
                                    void* zsec, int ldz);

/* ---- QP update (host, FP64) -------------------------------------------
 * S/main_hqpe.sc.f90:152-161: eqp = elda + (sex + sexcore)*hartree
 * + Re(sec(i,i))*hartree - vxc, energies in eV, self-energies in Hartree. */
ecalj_status ecalj_qp_energy(int ntq, const double* elda, const double* sex,
                             const double* sexcore, const double* sec, int ldsec,
                             const double* vxc, double* eqp);
/* S/main_hqpe.sc.f90:175-181: se = sex2 + sexcore2 + (sec+sec^H)/2 - vxcmat. */
ecalj_status ecalj_sigma_herm(int ntq, const double* sex2, const double* sexcore2,
                              const double* sec, const double* vxcmat, double* se);

/* ---- learned chi0 tail ----------------------------------------------------
 * No reference counterpart.  The reference truncates the chi0 band sum
 * (nband_chi0 / emax_chi0, S/m_readgwinput.f90:34-35, applied in
 * tetwt5x_dtet4 at S/m_tetwt.f90:140-145); these entry points predict what
 * that truncation drops.  Network: nlayer layers of widths dims[0..nlayer],
 *   z_l = h_l W_l^T + b_l, h_0 = x, h_{l+1} = silu(z_l) (l < nlayer-1), y = z_{nlayer-1};
 * params (FP32, device): per layer W_l (dims[l+1] x dims[l], row-major) then
 * b_l.  Batches are row-major FP32 device arrays (batch x width).
 * ecalj_mlp_grad_dev: gradient of the mean squared error against target into
 *   grad (same layout as params); loss (device float*, may be NULL) receives
 *   the loss.  Deterministic (fixed-order reductions).
 * ecalj_mlp_adam_dev: one Adam step (step >= 1 counts from 1).
 * ecalj_mlp_to_bf16_dev: bf16 hi (and lo = x - hi, when lo != NULL) copies of params.
 * ecalj_mlp_infer_bf16_dev: forward on bf16 MFMA with FP32 accumulation;
 *   w_lo == NULL gives one bf16 product per MAC, else the hi/lo split
 *   (hi*hi + hi*lo + lo*hi); biases come from params.
 * ecalj_tail_features_dev: features of a truncated chi0 (ni x ni x nw,
 *   complex128, frequencies {0, nu_1, ...}): per basis function I,
 *   d_j = -Re chi(I,I,j), feat(I,:) = ((d_1..d_{nw-1})/d_0, log d_0),
 *   standardised with mu/sd (nw floats each); aux (2 ni doubles) = {d_0, d_iref}.
 * ecalj_tail_apply_dev: chi_out = chi + rho o sqrt(t t^T) per frequency, with
 *   t(I,j) = max(0, d_0(I) (mu_y[j] + sd_y[j] pred(I,j))) and
 *   rho(I,J) = chi(I,J,iref)/sqrt(d_iref(I) d_iref(J)): a Hermitian negative
 *   semidefinite tail.  chi_out may equal chi. */
ecalj_status ecalj_mlp_nparam(int nlayer, const int* dims, int64_t* nparam);
ecalj_status ecalj_mlp_forward_dev(ecalj_handle h, int nlayer, const int* dims, const void* params,
                                   const void* x, int batch, void* y);
ecalj_status ecalj_mlp_grad_dev(ecalj_handle h, int nlayer, const int* dims, const void* params,
                                const void* x, const void* target, int batch, void* grad, void* loss);
ecalj_status ecalj_mlp_adam_dev(ecalj_handle h, int64_t nparam, void* params, const void* grad, void* m,
                                void* v, int step, double lr, double beta1, double beta2, double eps);
ecalj_status ecalj_mlp_to_bf16_dev(ecalj_handle h, int64_t nparam, const void* params, void* hi, void* lo);
ecalj_status ecalj_mlp_infer_bf16_dev(ecalj_handle h, int nlayer, const int* dims, const void* params,
                                      const void* w_hi, const void* w_lo, const void* x, int batch, void* y);
ecalj_status ecalj_tail_features_dev(ecalj_handle h, const void* chi, int ni, int nw, int iref,
                                     const void* mu, const void* sd, void* feat, void* aux);
ecalj_status ecalj_tail_apply_dev(ecalj_handle h, const void* chi, int ni, int nw, int iref,
                                  const void* pred, const void* mu_y, const void* sd_y, const void* aux,
                                  void* chi_out);

#ifdef __cplusplus
}
#endif
#endif
