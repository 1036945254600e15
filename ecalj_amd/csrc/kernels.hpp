// kernels.hpp -- launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include "ctx.hpp"

namespace ecalj {

ecalj_status ppb_relayout(ecalj_ctx* h, const double* src, double* dst, int nv, int md);
ecalj_status x0_herk(ecalj_ctx* h, int mode, const void* z, const int64_t* d_col, const double* d_w,
                     const int64_t* d_bstart, const void* d_items, int nitems, void* rcxq,
                     int64_t slice_elems, int npr, int npr_col, int ipr_col, int mirror);
// d_coef: coef(v, iw) at [v + ldc*(iw-1)], ldc >= hilbert_ldc(nv), zero beyond nv.
// d_dscale (optional, real axis): adds i*dscale(it)*rcxq(:, it) to output it.
inline int hilbert_ldc(int nv) { return ((nv + 31) / 32) * 32; }
ecalj_status hilbert(ecalj_ctx* h, bool wide, void* rcxq, int64_t slice, int nwhis, int nv, int ldc,
                     const double* d_coef, const double* d_scale, const double* d_dscale, cd* out, int64_t n);
ecalj_status chi0_freq(ecalj_ctx* h, int mode, const void* z, const double* w, int ni, int npairs, int nf, cd* chi);
ecalj_status convert(ecalj_ctx* h, const void* src, bool src64, void* dst, bool dst64, int64_t n);
ecalj_status eps_direct(ecalj_ctx* h, const cd* chi, const double* s, double fac, int n, int nw, cd* eps);
ecalj_status wc_form(ecalj_ctx* h, const cd* inv, const int* perm, const double* s, int n, int nw, cd* w,
                     int64_t ldw, int64_t wstride);
// Mixed form of the W build (the reference's __MP split): chi_E and W_E in
// complex64 around the FP64 epsilon and its inverse.
ecalj_status eps_narrow(ecalj_ctx* h, const cf* chi, const double* s, double head, int n, int nw, cd* eps);
ecalj_status wc_form32(ecalj_ctx* h, const cd* inv, const int* perm, const double* s, int n, int nw, cf* w);
ecalj_status widen_ld(ecalj_ctx* h, const cf* src, int n, int nw, cd* dst, int64_t ldw, int64_t wstride);
// In-place inverse of nmat matrices.  perm == nullptr: no row exchanges (the
// matrix must be safe to eliminate in order, e.g. Hermitian positive
// definite).  Otherwise partial pivoting: on return column j of inv(A) is
// column (*perm)[j + n*mat] of a (*perm: device, n ints per matrix).
ecalj_status gj_inverse(ecalj_ctx* h, cd* a, int n, int64_t lda, int64_t astride, int nmat,
                        const int** perm = nullptr);
// a(:, j) <- a(:, perm[j]) for each matrix (column unscramble after gj_inverse)
ecalj_status gj_unscramble(ecalj_ctx* h, cd* a, int n, int64_t lda, int64_t astride, int nmat, const int* perm);
// Scratch bytes that gj_inverse with pivoting followed by gj_unscramble take
// from a fixed arena (slots rounded up to 256 B) for nmat matrices of order n.
size_t zminv_scratch_bytes(int64_t n, int64_t nmat);
struct RealSpec;
// rs (optional): the real-axis pole term, added into czmelwc before the final product
ecalj_status sig_contract(ecalj_ctx* h, int mode, int nbatch, const void* zmel, int n, int nm, int ntq, int niw,
                          const cd* wc, int64_t ldw, const double* wgt, cd* zsec, int64_t ldz,
                          const RealSpec* rs = nullptr);
ecalj_status sig_weights(ecalj_ctx* h, int nbatch, int niw, int nm, int ntq, int ns1, int nctot,
                         const double* omega, const double* ekc, double esmr, double ua, const double* wkkr,
                         const double* x, const double* wt, const double* expa, double* wgt);

ecalj_status dev_copy(ecalj_ctx* h, void* dst, const void* src, size_t bytes);
ecalj_status dev_zero(ecalj_ctx* h, void* p, size_t bytes);
ecalj_status phase_similarity(ecalj_ctx* h, const cd* a, int n, int nmat, const cd* d, int nk, cd* out);

ecalj_status sig_bf3(ecalj_ctx* h, int nb, const cf* zmel, int64_t z_stride, int n, int ncol, int nw1,
                     const cd* wc, int64_t ldw, int64_t w_stride, const double* wgt, int64_t g_stride, cf* cz,
                     int64_t cz_stride, unsigned short* as, unsigned short* bs);
size_t sig_bf3_bytes_per_k(int n, int ncol, int nw1);
// The same contraction with fp8 cross terms (scratch as for sig_bf3)
ecalj_status sig_f8(ecalj_ctx* h, int nb, const cf* zmel, int64_t z_stride, int n, int ncol, int nw1, const cd* wc,
                    int64_t ldw, int64_t w_stride, const double* wgt, int64_t g_stride, cf* cz, int64_t cz_stride,
                    unsigned short* as, unsigned short* bs);
ecalj_status sig_f8_grouped(ecalj_ctx* h, int nk, const int* ncol, const cf* zmel, int n, int nw1, const cd* wc,
                            int64_t ldw, int64_t w_stride, const double* wgt, cf* cz, unsigned short* as,
                            unsigned short* bs);
ecalj_status sig_bf3_grouped(ecalj_ctx* h, int nk, const int* ncol, const cf* zmel, int n, int nw1, const cd* wc,
                             int64_t ldw, int64_t w_stride, const double* wgt, cf* cz, unsigned short* as,
                             unsigned short* bs);
// Inputs of the real-axis pole term (S/m_sxcf_gemm.f90:501-594); host arrays
// except wr.  ekc: nm per internal k.
struct RealSpec {
  const cd* wr = nullptr;
  int64_t ldw = 0, wstride = 0;
  int nw = 0;
  const double *freq_r = nullptr, *omega = nullptr, *ekc = nullptr, *wkkr = nullptr;
  double ef = 0.0, esmr = 0.0;
  int nctot = 0, ns2r = 0;
};
ecalj_status sig_real(ecalj_ctx* h, int mode, int nb, const void* zmel, int n, int nm, int ntq, const RealSpec& rs,
                      cd* zsec, int64_t ldz);
struct RealPlanHandle;
ecalj_status real_plan_make(ecalj_ctx* h, const RealSpec& rs, int nb, int n, int nm, int ntq, RealPlanHandle** out);
void real_plan_free(RealPlanHandle* r);
ecalj_status real_plan_into_cz(ecalj_ctx* h, const RealPlanHandle* r, const RealSpec& rs, int mode,
                               const void* zmel, int n, int nm, int ntq, int k0, int k1, void* cz);
ecalj_status sig_contract_grouped(ecalj_ctx* h, int mode, int nk, const int* nm, const void* zmel, int n, int ntq,
                                  int niw, const cd* wc, int64_t ldw, const double* wgt, cd* zsec, int64_t ldz);
size_t sig_bf3_a_elems_per_k(int n, int nw1);
// The frequency-inner three-product form of the bf16x3 Sigma_c contraction
// (sig3m.hip): operand preparation (six W planes into as, padded complex64
// zmel into zp) and the contraction itself: wm = 2 (8 waves in 2 x 4) or 1
// (8 waves side by side) on the 16x16x32 MFMA, wm = 0 (4 waves side by side)
// on the 32x32x16 MFMA.
size_t sig3m_a_elems_per_k(int n, int nw1);
ecalj_status sig3m_prep(ecalj_ctx* h, int nb, const cf* zmel, int64_t z_stride, int n, int ncol, int nw1,
                        const cd* wc, int64_t ldw, int64_t w_stride, unsigned short* as, cf* zp);
ecalj_status sig3m_run(ecalj_ctx* h, bool xcd, int wm, int nb, int n, int ncol, int nw1, const unsigned short* as,
                       const cf* zp, const double* wgt, int64_t g_stride, cf* cz, int64_t cz_stride);
ecalj_status sigv4_run(ecalj_ctx* h, bool xcd, int nb, int n, int ncol, int nw1, const unsigned short* as,
                       int64_t as_stride, const cf* zp, const double* wgt, int64_t g_stride, cf* cz,
                       int64_t cz_stride);
ecalj_status sig_pad_z(ecalj_ctx* h, int nb, const cf* zmel, int64_t z_stride, int n, int ncol, cf* zp);
ecalj_status sig3m_grouped_prep(ecalj_ctx* h, int nk, const int* ncol, const cf* zmel, int n, int nw1, const cd* wc,
                                int64_t ldw, int64_t w_stride, unsigned short* as, cf* zp);
ecalj_status sig3m_grouped_run(ecalj_ctx* h, int wm, int nk, const int* ncol, int n, int nw1,
                               const unsigned short* as, const cf* zp, const double* wgt, cf* cz);
// x0gemm in the bf16x3 form: per-bin entry lists on the host (bin b: entries
// cnt[b]..cnt[b+1]-1 with zmel column offsets col[] and weights wt[]), output
// slice slice[b] of rcxq (npr x npr_col, complex64, slice stride slice_elems)
ecalj_status x0_bf3(ecalj_ctx* h, const cf* z, int npr, int npr_col, int ipr_col, int nbin, const int64_t* cnt,
                    const int64_t* col, const double* wt, const int* slice, cf* rcxq, int64_t slice_elems,
                    int mirror);
// zmel's second contraction in the bf16x3 form: C_b = s(alpha T_b^T opB(X_b)),
// T_b (K x M, column stride lda) at A + (b % div) a_s0 + (b / div) a_s1, X_b
// (K x N, ldb) at B + (b / div) b_s1 (b_s1 = 0: shared), opB = conj when
// conjB, s = conj when conjC; C_b (M x N, ldc) at C + (b % div) c_s0 + (b / div) c_s1
ecalj_status zmel_z_bf3(ecalj_ctx* h, int M, int N, int K, const cf* A, int64_t lda, int64_t a_s0, int64_t a_s1,
                        const cf* B, int64_t ldb, int64_t b_s1, bool conjB, cd alpha, bool conjC, cf* C,
                        int64_t ldc, int64_t c_s0, int64_t c_s1, int batch, int div);
// zmel's first contraction in the bf16x3 form: T_b = P opB(X_b), P (M x K,
// real, lda) shared, X_b (K x N, K contiguous, ldb) at X + b x_s1, conjugated
// when conjB; T_b (M x N, ldc) at C + b c_s1
ecalj_status zmel_p_bf3(ecalj_ctx* h, int M, int N, int K, const double* P, int64_t lda, const cf* X,
                        int64_t ldb, int64_t x_s1, bool conjB, cf* C, int64_t ldc, int64_t c_s1, int batch);
// Batched complex GEMM on the bf16x3 core of the Sigma_c kernel:
// C_b = alpha op(A_b) op(B_b) (m x ncol, leading dimension m, stride c_stride,
// complex64).  An operand is complex128 (wide) or complex64 with leading
// dimension ld and batch stride (0: shared); herm: the operand is X^H.
struct QOperand {
  const void* p = nullptr;
  int64_t ld = 0, stride = 0;
  bool wide = false, herm = false;
};
ecalj_status qgemm_bf3(ecalj_ctx* h, int nb, int m, int ncol, int k, const QOperand& A, const QOperand& B,
                       double alpha, cf* C, int64_t c_stride);
// The transforms around the epsilon inversion (wbuild.hip), in the handle's
// form.  eps_build: eps (n x n x nw) from chi (complex128, M basis when
// m_basis), head > 0 replaces the (1,1) element by 1 - head fac chi_E(1,1).
// w_form: W = s inv s - diag(s^2) (columns through perm), to the M basis when
// m_basis, into wc (leading dimension ldw); inv is overwritten.  w_e2m_inplace:
// W records (ld ldw) <- zcousq W zcousq^H.  tmp: n*n*nw complex128.
ecalj_status eps_build(ecalj_ctx* h, const void* chi, int n, int nw, double fac, bool m_basis, double head, cd* eps,
                       cd* tmp);
ecalj_status w_form(ecalj_ctx* h, cd* inv, const int* perm, int n, int nw, bool m_basis, cd* wc, int64_t ldw,
                    cd* tmp);
ecalj_status w_e2m_inplace(ecalj_ctx* h, cd* w, int n, int nw, int64_t ldw, cd* tmp);

}  // namespace ecalj
