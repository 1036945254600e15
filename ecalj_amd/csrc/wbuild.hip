// wbuild.hip -- the transforms around the epsilon inversion of the W build,
// shared by WVIllwI / WVRllwR (api.hip), the Gamma point and modifyWV0
// (gamma.hip).
//
// FP64 form: complex128 throughout; epsilon is fused into the epilogue of
// the second M -> E GEMM.  Mixed forms: the reference's __MP split
// (S/m_kind.f90:3-6, S/m_llw.f90:31-39,62-64): the M <-> E transforms on the
// form's GEMM core with chi_E and W_E in complex64 (cmm_d), epsilon and its
// inverse in complex128 (epstilde is complex(8) in every build).  In the
// bf16x3 form the transforms run on the pre-split core of the Sigma_c kernel
// (qgemm_bf3).
#include <algorithm>

#include "ctx.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace ecalj {
namespace {

inline int blocks_for(int64_t tot) { return (int)std::max<int64_t>(1, std::min<int64_t>((tot + 255) / 256, 65536)); }

// eps = delta - fac s_i chi_ij s_j; eps(1,1) = 1 - fac head chi(1,1) when head > 0.
// chi and eps may be the same buffer (elementwise).
__global__ void eps_head_kernel(const cd* chi, const double* __restrict__ s, double fac, double head, int n, int nw,
                                cd* eps) {
  const int64_t nn = (int64_t)n * n, tot = nn * nw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ij = e % nn;
    const int i = (int)(ij % n), j = (int)(ij / n);
    const cd x = chi[e];
    const double sc = (head > 0.0 && ij == 0) ? fac * head : fac * s[i] * s[j];
    eps[e] = cd{(i == j ? 1.0 : 0.0) - sc * x.re, -sc * x.im};
  }
}

// dst (complex64, ld n) = src (complex128, leading dimension ld, record stride ld*n)
__global__ void narrow_ld_kernel(const cd* __restrict__ src, int n, int nw, int64_t ld, cf* __restrict__ dst) {
  const int64_t nn = (int64_t)n * n, tot = nn * nw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int iw = (int)(e / nn);
    const int64_t ij = e % nn;
    const int i = (int)(ij % n), j = (int)(ij / n);
    const cd x = src[ld * n * iw + i + ld * j];
    dst[e] = cf{(float)x.re, (float)x.im};
  }
}

QOperand opnd(const void* p, int64_t ld, int64_t stride, bool w, bool herm) {
  QOperand o;
  o.p = p; o.ld = ld; o.stride = stride; o.wide = w; o.herm = herm;
  return o;
}

// out = zcousq (X zcousq^H) for nw matrices X (form's GEMM core); X and out:
// complex128 with leading dimension ldx / ldo (wide), or complex64 ld n.
// tmp holds nn*nw complex128 (two complex64 halves in the mixed forms).
ecalj_status e2m(ecalj_ctx* h, int n, int nw, const cd* X, int64_t ldx, cd* out, int64_t ldo, cd* tmp, cd* scratch) {
  const int64_t nn = (int64_t)n * n;
  if (h->precision == ECALJ_PREC_F64) {
    GemmCall g3;  // scratch = X zcousq^H
    g3.tb = true; g3.cb = true;
    g3.M = n; g3.N = n; g3.K = n;
    g3.A = X; g3.lda = ldx; g3.bA = {ldx * n, 0};
    g3.B = h->d_zcousq; g3.ldb = n;
    g3.C = scratch; g3.ldc = n; g3.bC = {nn, 0};
    g3.batch = nw; g3.div = nw;
    ECALJ_TRY(gemm(h, g3));
    GemmCall g4;  // out = zcousq scratch
    g4.M = n; g4.N = n; g4.K = n;
    g4.A = h->d_zcousq; g4.lda = n;
    g4.B = scratch; g4.ldb = n; g4.bB = {nn, 0};
    g4.C = out; g4.ldc = ldo; g4.bC = {ldo * n, 0};
    g4.batch = nw; g4.div = nw;
    return gemm(h, g4);
  }
  cf* a32 = (cf*)tmp;
  cf* b32 = a32 + nn * nw;
  if (h->precision == ECALJ_PREC_BF16X3) {
    ECALJ_TRY(qgemm_bf3(h, nw, n, n, n, opnd(X, ldx, ldx * n, true, false), opnd(h->d_zcousq, n, 0, true, true), 1.0,
                        a32, nn));
    ECALJ_TRY(qgemm_bf3(h, nw, n, n, n, opnd(h->d_zcousq, n, 0, true, false), opnd(a32, n, nn, false, false), 1.0,
                        b32, nn));
  } else {
    hipLaunchKernelGGL(narrow_ld_kernel, dim3(blocks_for(nn * nw)), dim3(256), 0, h->stream, X, n, nw, ldx, b32);
    ECALJ_HIP(hipGetLastError());
    GemmCall g3;
    g3.mode = h->precision;
    g3.tb = true; g3.cb = true;
    g3.M = n; g3.N = n; g3.K = n;
    g3.A = b32; g3.lda = n; g3.bA = {nn, 0};
    g3.B = h->d_zcousq32; g3.ldb = n;
    g3.C = a32; g3.ldc = n; g3.bC = {nn, 0};
    g3.batch = nw; g3.div = nw;
    ECALJ_TRY(gemm(h, g3));
    GemmCall g4;
    g4.mode = h->precision;
    g4.M = n; g4.N = n; g4.K = n;
    g4.A = h->d_zcousq32; g4.lda = n;
    g4.B = a32; g4.ldb = n; g4.bB = {nn, 0};
    g4.C = b32; g4.ldc = n; g4.bC = {nn, 0};
    g4.batch = nw; g4.div = nw;
    ECALJ_TRY(gemm(h, g4));
  }
  return widen_ld(h, b32, n, nw, out, ldo, ldo * n);
}

}  // namespace

ecalj_status eps_build(ecalj_ctx* h, const void* chi, int n, int nw, double fac, bool m_basis, double head, cd* eps,
                       cd* tmp) {
  const int64_t nn = (int64_t)n * n;
  if (!m_basis) {
    hipLaunchKernelGGL(eps_head_kernel, dim3(blocks_for(nn * nw)), dim3(256), 0, h->stream, (const cd*)chi,
                       h->d_vcousq, fac, head, n, nw, eps);
    ECALJ_HIP(hipGetLastError());
    return ECALJ_OK;
  }
  // chi_E = fac zcousq^H (chi zcousq) (S/m_llw.f90:122-125, 336-337)
  if (h->precision == ECALJ_PREC_F64) {
    GemmCall g1;
    g1.M = n; g1.N = n; g1.K = n;
    g1.A = chi; g1.lda = n; g1.bA = {nn, 0};
    g1.B = h->d_zcousq; g1.ldb = n;
    g1.C = tmp; g1.ldc = n; g1.bC = {nn, 0};
    g1.batch = nw; g1.div = nw;
    ECALJ_TRY(gemm(h, g1));
    GemmCall g2;
    g2.ta = true; g2.ca = true;
    g2.M = n; g2.N = n; g2.K = n;
    g2.A = h->d_zcousq; g2.lda = n;
    g2.B = tmp; g2.ldb = n; g2.bB = {nn, 0};
    g2.C = eps; g2.ldc = n; g2.bC = {nn, 0};
    g2.batch = nw; g2.div = nw;
    g2.alpha = cd{fac, 0.0};
    if (head > 0.0) {  // plain product, then epsilon with the head
      ECALJ_TRY(gemm(h, g2));
      hipLaunchKernelGGL(eps_head_kernel, dim3(blocks_for(nn * nw)), dim3(256), 0, h->stream, eps, h->d_vcousq, 1.0,
                         head, n, nw, eps);
      ECALJ_HIP(hipGetLastError());
      return ECALJ_OK;
    }
    g2.epi = EPI_EPS; g2.svec = h->d_vcousq;  // epsilon fused into the epilogue
    return gemm(h, g2);
  }
  cf* a32 = (cf*)tmp;
  cf* b32 = a32 + nn * nw;
  if (h->precision == ECALJ_PREC_BF16X3) {
    ECALJ_TRY(qgemm_bf3(h, nw, n, n, n, opnd(chi, n, nn, true, false), opnd(h->d_zcousq, n, 0, true, false), 1.0, b32,
                        nn));
    ECALJ_TRY(qgemm_bf3(h, nw, n, n, n, opnd(h->d_zcousq, n, 0, true, true), opnd(b32, n, nn, false, false), fac, a32,
                        nn));
  } else {
    ECALJ_TRY(convert(h, chi, true, a32, false, nn * nw));
    GemmCall g1;
    g1.mode = h->precision;
    g1.M = n; g1.N = n; g1.K = n;
    g1.A = a32; g1.lda = n; g1.bA = {nn, 0};
    g1.B = h->d_zcousq32; g1.ldb = n;
    g1.C = b32; g1.ldc = n; g1.bC = {nn, 0};
    g1.batch = nw; g1.div = nw;
    ECALJ_TRY(gemm(h, g1));
    GemmCall g2;
    g2.mode = h->precision;
    g2.ta = true; g2.ca = true;
    g2.M = n; g2.N = n; g2.K = n;
    g2.A = h->d_zcousq32; g2.lda = n;
    g2.B = b32; g2.ldb = n; g2.bB = {nn, 0};
    g2.C = a32; g2.ldc = n; g2.bC = {nn, 0};
    g2.batch = nw; g2.div = nw;
    g2.alpha = cd{fac, 0.0};
    ECALJ_TRY(gemm(h, g2));
  }
  return eps_narrow(h, a32, h->d_vcousq, head, n, nw, eps);  // epstilde in complex(8)
}

ecalj_status w_form(ecalj_ctx* h, cd* inv, const int* perm, int n, int nw, bool m_basis, cd* wc, int64_t ldw,
                    cd* tmp) {
  const int64_t nn = (int64_t)n * n;
  if (!m_basis) return wc_form(h, inv, perm, h->d_vcousq, n, nw, wc, ldw, ldw * n);
  if (h->precision == ECALJ_PREC_F64) {
    ECALJ_TRY(wc_form(h, inv, perm, h->d_vcousq, n, nw, tmp, n, nn));
    return e2m(h, n, nw, tmp, n, wc, ldw, nullptr, inv);  // the inverse is free: scratch
  }
  // W_E in complex64 (the second half of tmp), then E -> M on the form's core
  cf* a32 = (cf*)tmp;
  cf* b32 = a32 + nn * nw;
  ECALJ_TRY(wc_form32(h, inv, perm, h->d_vcousq, n, nw, b32));
  if (h->precision == ECALJ_PREC_BF16X3) {
    ECALJ_TRY(qgemm_bf3(h, nw, n, n, n, opnd(b32, n, nn, false, false), opnd(h->d_zcousq, n, 0, true, true), 1.0, a32,
                        nn));
    ECALJ_TRY(qgemm_bf3(h, nw, n, n, n, opnd(h->d_zcousq, n, 0, true, false), opnd(a32, n, nn, false, false), 1.0,
                        b32, nn));
  } else {
    GemmCall g3;
    g3.mode = h->precision;
    g3.tb = true; g3.cb = true;
    g3.M = n; g3.N = n; g3.K = n;
    g3.A = b32; g3.lda = n; g3.bA = {nn, 0};
    g3.B = h->d_zcousq32; g3.ldb = n;
    g3.C = a32; g3.ldc = n; g3.bC = {nn, 0};
    g3.batch = nw; g3.div = nw;
    ECALJ_TRY(gemm(h, g3));
    GemmCall g4;
    g4.mode = h->precision;
    g4.M = n; g4.N = n; g4.K = n;
    g4.A = h->d_zcousq32; g4.lda = n;
    g4.B = a32; g4.ldb = n; g4.bB = {nn, 0};
    g4.C = b32; g4.ldc = n; g4.bC = {nn, 0};
    g4.batch = nw; g4.div = nw;
    ECALJ_TRY(gemm(h, g4));
  }
  return widen_ld(h, b32, n, nw, wc, ldw, ldw * n);
}

ecalj_status w_e2m_inplace(ecalj_ctx* h, cd* w, int n, int nw, int64_t ldw, cd* tmp) {
  // FP64: scratch is tmp itself (X = w is read from its own records)
  return e2m(h, n, nw, w, ldw, w, ldw, tmp, tmp);
}

}  // namespace ecalj
