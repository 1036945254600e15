// zmelipw.hip -- interstitial plane-wave (IPW) rows of zmel
// (get_zmel_init_gemm's ZmelIPW block, S/m_zmel.f90:418-538), and the
// assembly of MT and IPW rows with the M->E transform and conjugation
// (S/m_zmel.f90:559-560, 590-594).
//
//   zmel(nbloch+c, n, m) = phase(c) sum_{G1,G2} ggg(G1 - Gc - G2 - nadd) geigq(G1, m) dgeigqk(G2, n)
//
// computed as the reference does, in the cheaper order: with more middle
// states than end states (nm > ntp0) G1 first, gg(X, m) = sum_G1 ggg(G1 - X -
// nadd) geigq(G1, m) over the combined list X (every Gc + G2), then per end
// state a gather gg(Gc + G2, m) and a GEMM with dgeigqk; otherwise G2 first
// with X = Gc - G1.  The step-function matrix and the gathered panels are
// built by small kernels; the two contractions are batched GEMMs on the
// precision's MFMA core.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace ecalj {
namespace {

__device__ inline int cube_at(int x, int y, int z, const int* lo, const int* n) {
  const int i0 = x - lo[0], i1 = y - lo[1], i2 = z - lo[2];
  if (i0 < 0 || i0 >= n[0] || i1 < 0 || i1 >= n[1] || i2 < 0 || i2 >= n[2]) return -1;
  return i0 + n[0] * (i1 + n[1] * i2);
}

struct IpwGeom {
  int glo[3], gn[3];  // igggi cube
  int clo[3], cn[3];  // igcgp2i cube
  int nadd[3];
};

// mat(X, g) = ggg(igggi(s*gp(g) - X - nadd)), 0 outside the cube or for index 0
template <class T>
__global__ void ipw_ggg_kernel(const cd* __restrict__ ggg, const int* __restrict__ igggi,
                               const int* __restrict__ nvgcgp2, int ngcgp, const int* __restrict__ gp, int ngp, int s,
                               IpwGeom geo, T* __restrict__ mat) {
  const int64_t tot = (int64_t)ngcgp * ngp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % ngcgp), c = (int)(e / ngcgp);
    const int x = s * gp[3 * c] - nvgcgp2[3 * r] - geo.nadd[0];
    const int y = s * gp[3 * c + 1] - nvgcgp2[3 * r + 1] - geo.nadd[1];
    const int z = s * gp[3 * c + 2] - nvgcgp2[3 * r + 2] - geo.nadd[2];
    const int at = cube_at(x, y, z, geo.glo, geo.gn);
    double re = 0.0, im = 0.0;
    if (at >= 0) {
      const int i = igggi[at];
      if (i > 0) {
        re = ggg[i - 1].re;
        im = ggg[i - 1].im;
      }
    }
    st_c(mat + e, re, im);
  }
}

// work(c, q, j) = phase(c) * gg(igcgp2i(gc(c) + t*gq(q)) - 1, j)
template <class T>
__global__ void ipw_gather_kernel(const T* __restrict__ gg, int ngcgp, const int* __restrict__ igcgp2i,
                                  const int* __restrict__ gc, int ngc, const int* __restrict__ gq, int ngq, int t,
                                  const cd* __restrict__ phase, int ncol, IpwGeom geo, T* __restrict__ work) {
  const int64_t tot = (int64_t)ngc * ngq * ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % ngc);
    const int64_t rest = e / ngc;
    const int q = (int)(rest % ngq), j = (int)(rest / ngq);
    const int at = cube_at(gc[3 * c] + t * gq[3 * q], gc[3 * c + 1] + t * gq[3 * q + 1], gc[3 * c + 2] + t * gq[3 * q + 2],
                           geo.clo, geo.cn);
    const int r = igcgp2i[at] - 1;  // checked on the host
    double xr, xi;
    ld_c2(gg + r + (int64_t)ngcgp * j, xr, xi);
    const cd p = phase[c];
    st_c(work + e, p.re * xr - p.im * xi, p.re * xi + p.im * xr);
  }
}

// out(r, c) = r < nbloch ? mt(r, c) : ipw(r - nbloch, c), conjugated when asked
template <class T>
__global__ void zmel_stack_kernel(const T* __restrict__ mt, int nbloch, const T* __restrict__ ipw, int ngc,
                                  int64_t ncol, int conj, T* __restrict__ out) {
  const int nbb = nbloch + ngc;
  const int64_t tot = (int64_t)nbb * ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % nbb);
    const int64_t c = e / nbb;
    double xr, xi;
    if (r < nbloch)
      ld_c2(mt + r + (int64_t)nbloch * c, xr, xi);
    else
      ld_c2(ipw + (r - nbloch) + (int64_t)ngc * c, xr, xi);
    st_c(out + e, xr, conj ? -xi : xi);
  }
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65536); }

template <class T>
ecalj_status zmel_ipw_t(ecalj_ctx* h, int ngc, const int* ngveccR, const int* nadd, const double* phase, int ngpmx,
                        int ngp1, const int* g1, const void* geigq, int ntp0, int ngp2, const int* g2,
                        const void* dgeigqk, int nm, T* zmel, int64_t ldz, int row0) {
  const bool g1first = nm > ntp0;  // S/m_zmel.f90:435
  const int ngp = g1first ? ngp1 : ngp2, ngq = g1first ? ngp2 : ngp1;
  const int s = g1first ? 1 : -1, t = g1first ? 1 : -1;
  const int* gp = g1first ? g1 : g2;
  const int* gq = g1first ? g2 : g1;
  const int nB = g1first ? ntp0 : nm, nC = g1first ? nm : ntp0;
  IpwGeom geo;
  for (int d = 0; d < 3; ++d) {
    geo.glo[d] = h->ggg_lo[d];
    geo.gn[d] = h->ggg_n[d];
    geo.clo[d] = h->gcp_lo[d];
    geo.cn[d] = h->gcp_n[d];
    geo.nadd[d] = nadd[d];
  }
  // every second-stage lookup must hit the combined list (the reference assumes it)
  for (int q = 0; q < ngq; ++q)
    for (int c = 0; c < ngc; ++c) {
      int v[3], i[3];
      for (int d = 0; d < 3; ++d) {
        v[d] = ngveccR[3 * c + d] + t * gq[3 * q + d];
        i[d] = v[d] - geo.clo[d];
        if (i[d] < 0 || i[d] >= geo.cn[d]) return ECALJ_ERR_ARG;
      }
      if (h->h_igcgp2i[i[0] + (size_t)geo.cn[0] * (i[1] + (size_t)geo.cn[1] * i[2])] <= 0) return ECALJ_ERR_ARG;
    }
  // G lists and phases
  std::vector<unsigned char> blob;
  auto put = [&](const void* src, size_t bytes) {
    const size_t o = blob.size();
    blob.resize(o + ((bytes + 15) / 16) * 16);
    memcpy(blob.data() + o, src, bytes);
    return o;
  };
  const size_t o_gc = put(ngveccR, sizeof(int) * 3 * ngc), o_gp = put(gp, sizeof(int) * 3 * ngp),
               o_gq = put(gq, sizeof(int) * 3 * ngq), o_ph = put(phase, sizeof(cd) * ngc);
  unsigned char* d;
  ECALJ_TRY(upload_blob(h, ecalj_ctx::S_IPWT, blob, &d));
  // eigenvector coefficients in the storage type
  const void* B = g1first ? geigq : dgeigqk;
  const void* C2 = g1first ? dgeigqk : geigq;
  if (sizeof(T) == sizeof(cf)) {
    void *b32, *c32;
    ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_CVT1, sizeof(cf) * (size_t)ngpmx * nB, &b32));
    ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_CVT2, sizeof(cf) * (size_t)ngpmx * nC, &c32));
    ECALJ_TRY(convert(h, B, true, b32, false, (int64_t)ngpmx * nB));
    ECALJ_TRY(convert(h, C2, true, c32, false, (int64_t)ngpmx * nC));
    B = b32;
    C2 = c32;
  }
  T *mat, *gg, *work;
  const int ngcgp = h->ngcgp;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_IPW1, sizeof(T) * (size_t)ngcgp * ngp, (void**)&mat));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_IPW2, sizeof(T) * (size_t)ngcgp * nB, (void**)&gg));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_IPW3, sizeof(T) * (size_t)ngc * ngq * nB, (void**)&work));
  hipLaunchKernelGGL((ipw_ggg_kernel<T>), dim3(grid_for((int64_t)ngcgp * ngp)), dim3(256), 0, h->stream, h->d_ggg,
                     h->d_igggi, h->d_nvgcgp2, ngcgp, (const int*)(d + o_gp), ngp, s, geo, mat);
  ECALJ_HIP(hipGetLastError());
  GemmCall a;  // gg = mat * B
  a.mode = h->precision;
  a.M = ngcgp; a.N = nB; a.K = ngp;
  a.A = mat; a.lda = ngcgp;
  a.B = B; a.ldb = ngpmx;
  a.C = gg; a.ldc = ngcgp;
  ECALJ_TRY(gemm(h, a));
  hipLaunchKernelGGL((ipw_gather_kernel<T>), dim3(grid_for((int64_t)ngc * ngq * nB)), dim3(256), 0, h->stream, gg,
                     ngcgp, h->d_igcgp2i, (const int*)(d + o_gc), ngc, (const int*)(d + o_gq), ngq, t,
                     (const cd*)(d + o_ph), nB, geo, work);
  ECALJ_HIP(hipGetLastError());
  GemmCall z;  // per column j of the first contraction: zmel rows = work_j * C2
  z.mode = h->precision;
  z.M = ngc; z.N = nC; z.K = ngq;
  z.A = work; z.lda = ngc; z.bA = {(int64_t)ngc * ngq, 0};
  z.B = C2; z.ldb = ngpmx;
  z.C = zmel + row0;
  if (g1first) {  // j = end state: C_j(c, n) at zmel(row0+c, n, j)
    z.ldc = ldz;
    z.bC = {ldz * nm, 0};
  } else {  // j = middle state: C_j(c, m) at zmel(row0+c, j, m)
    z.ldc = ldz * nm;
    z.bC = {ldz, 0};
  }
  z.batch = nB; z.div = nB;
  ECALJ_TRY(gemm(h, z));
  return ECALJ_OK;
}

inline bool wide_h(const ecalj_ctx* h) { return h->precision == ECALJ_PREC_F64; }

}  // namespace
}  // namespace ecalj

using namespace ecalj;

extern "C" ecalj_status ecalj_set_ppovl(ecalj_handle h, int nggg, const double* ggg, const int* ggg_lo,
                                        const int* ggg_n, const int* igggi, int ngcgp, const int* nvgcgp2,
                                        const int* gcp_lo, const int* gcp_n, const int* igcgp2i) {
  if (!h || !ggg || !ggg_lo || !ggg_n || !igggi || !nvgcgp2 || !gcp_lo || !gcp_n || !igcgp2i) return ECALJ_ERR_ARG;
  if (nggg < 1 || ngcgp < 1) return ECALJ_ERR_DIM;
  if (h->capturing) return ECALJ_ERR_STATE;
  size_t ng = 1, nc = 1;
  for (int d = 0; d < 3; ++d) {
    if (ggg_n[d] < 1 || gcp_n[d] < 1) return ECALJ_ERR_DIM;
    ng *= ggg_n[d];
    nc *= gcp_n[d];
  }
  for (size_t i = 0; i < ng; ++i)
    if (igggi[i] < 0 || igggi[i] > nggg) return ECALJ_ERR_ARG;
  for (size_t i = 0; i < nc; ++i)
    if (igcgp2i[i] < 0 || igcgp2i[i] > ngcgp) return ECALJ_ERR_ARG;
  ECALJ_HIP(hipSetDevice(h->device));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  ++h->state_gen;
  for (void** p : {(void**)&h->d_ggg, (void**)&h->d_igggi, (void**)&h->d_nvgcgp2, (void**)&h->d_igcgp2i})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  if (hipMalloc((void**)&h->d_ggg, sizeof(cd) * nggg) != hipSuccess ||
      hipMalloc((void**)&h->d_igggi, sizeof(int) * ng) != hipSuccess ||
      hipMalloc((void**)&h->d_nvgcgp2, sizeof(int) * 3 * (size_t)ngcgp) != hipSuccess ||
      hipMalloc((void**)&h->d_igcgp2i, sizeof(int) * nc) != hipSuccess)
    return ECALJ_ERR_NOMEM;
  ECALJ_HIP(hipMemcpy(h->d_ggg, ggg, sizeof(cd) * nggg, hipMemcpyHostToDevice));
  ECALJ_HIP(hipMemcpy(h->d_igggi, igggi, sizeof(int) * ng, hipMemcpyHostToDevice));
  ECALJ_HIP(hipMemcpy(h->d_nvgcgp2, nvgcgp2, sizeof(int) * 3 * (size_t)ngcgp, hipMemcpyHostToDevice));
  ECALJ_HIP(hipMemcpy(h->d_igcgp2i, igcgp2i, sizeof(int) * nc, hipMemcpyHostToDevice));
  h->h_igcgp2i.assign(igcgp2i, igcgp2i + nc);
  for (int d = 0; d < 3; ++d) {
    h->ggg_lo[d] = ggg_lo[d];
    h->ggg_n[d] = ggg_n[d];
    h->gcp_lo[d] = gcp_lo[d];
    h->gcp_n[d] = gcp_n[d];
  }
  h->ngcgp = ngcgp;
  return ECALJ_OK;
}

extern "C" ecalj_status ecalj_zmel_ipw_dev(ecalj_handle h, int ngc, const int* ngveccR, const int* nadd,
                                           const double* phase, int ngpmx, int ngp1, const int* ngvecpB1,
                                           const void* geigq, int ntp0, int ngp2, const int* ngvecpB2,
                                           const void* dgeigqk, int nm, void* zmel, int ldz, int row0) {
  if (!h || !ngveccR || !nadd || !phase || !ngvecpB1 || !geigq || !ngvecpB2 || !dgeigqk || !zmel)
    return ECALJ_ERR_ARG;
  if (!h->d_ggg) return ECALJ_ERR_STATE;
  if (ngc < 0 || ngp1 < 1 || ngp2 < 1 || ngpmx < std::max(ngp1, ngp2) || nm < 0 || ntp0 < 0 || row0 < 0 ||
      ldz < row0 + ngc)
    return ECALJ_ERR_DIM;
  if (ngc == 0 || nm == 0 || ntp0 == 0) return ECALJ_OK;
  if (wide_h(h))
    return zmel_ipw_t<cd>(h, ngc, ngveccR, nadd, phase, ngpmx, ngp1, ngvecpB1, geigq, ntp0, ngp2, ngvecpB2, dgeigqk,
                          nm, (cd*)zmel, ldz, row0);
  return zmel_ipw_t<cf>(h, ngc, ngveccR, nadd, phase, ngpmx, ngp1, ngvecpB1, geigq, ntp0, ngp2, ngvecpB2, dgeigqk,
                        nm, (cf*)zmel, ldz, row0);
}

extern "C" ecalj_status ecalj_zmel_ipw(ecalj_handle h, int ngc, const int* ngveccR, const int* nadd,
                                       const double* phase, int ngpmx, int ngp1, const int* ngvecpB1,
                                       const double* geigq, int ntp0, int ngp2, const int* ngvecpB2,
                                       const double* dgeigqk, int nm, double* zmel, int ldz, int row0) {
  if (!h || !geigq || !dgeigqk || !zmel) return ECALJ_ERR_ARG;
  if (h->capturing) return ECALJ_ERR_STATE;
  if (ngpmx < 1 || nm < 0 || ntp0 < 0 || ldz < 1) return ECALJ_ERR_DIM;
  const bool w = wide_h(h);
  const size_t es = w ? sizeof(cd) : sizeof(cf);
  const int64_t nz = (int64_t)ldz * nm * ntp0;
  void *dq = nullptr, *dk = nullptr, *dz = nullptr, *t64 = nullptr;
  ecalj_status st = ECALJ_OK;
  auto fail = [&](ecalj_status s) {
    (void)hipStreamSynchronize(h->stream);
    for (void* p : {dq, dk, dz, t64})
      if (p) (void)hipFree(p);
    return s;
  };
  if (hipMalloc(&dq, sizeof(cd) * (size_t)ngpmx * std::max(ntp0, 1)) != hipSuccess ||
      hipMalloc(&dk, sizeof(cd) * (size_t)ngpmx * std::max(nm, 1)) != hipSuccess ||
      hipMalloc(&dz, es * std::max<int64_t>(nz, 1)) != hipSuccess ||
      hipMalloc(&t64, sizeof(cd) * std::max<int64_t>(nz, 1)) != hipSuccess)
    return fail(ECALJ_ERR_NOMEM);
  if (hipMemcpyAsync(dq, geigq, sizeof(cd) * (size_t)ngpmx * ntp0, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(dk, dgeigqk, sizeof(cd) * (size_t)ngpmx * nm, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(t64, zmel, sizeof(cd) * nz, hipMemcpyHostToDevice, h->stream) != hipSuccess)
    return fail(ECALJ_ERR_HIP);
  if (!w && (st = convert(h, t64, true, dz, false, nz)) != ECALJ_OK) return fail(st);
  st = ecalj_zmel_ipw_dev(h, ngc, ngveccR, nadd, phase, ngpmx, ngp1, ngvecpB1, dq, ntp0, ngp2, ngvecpB2, dk, nm,
                          w ? t64 : dz, ldz, row0);
  if (st != ECALJ_OK) return fail(st);
  if (!w && (st = convert(h, dz, false, t64, true, nz)) != ECALJ_OK) return fail(st);
  if (hipMemcpyAsync(zmel, t64, sizeof(cd) * nz, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return fail(ECALJ_ERR_HIP);
  return fail(ECALJ_OK);
}

extern "C" ecalj_status ecalj_zmel_assemble_dev(ecalj_handle h, const void* mt, int nbloch, const void* ipw, int ngc,
                                                int64_t ncol, int is_m_basis, int zmelconjg, void* zmel) {
  if (!h || (!mt && nbloch) || (!ipw && ngc) || !zmel) return ECALJ_ERR_ARG;
  if (nbloch < 0 || ngc < 0 || ncol < 0) return ECALJ_ERR_DIM;
  const int nbb = nbloch + ngc;
  if (!is_m_basis && (!h->d_zcousq || h->ngb != nbb)) return ECALJ_ERR_STATE;
  if (nbb == 0 || ncol == 0) return ECALJ_OK;
  const bool w = wide_h(h);
  const int64_t tot = (int64_t)nbb * ncol;
  void* dst = zmel;
  if (!is_m_basis) ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_IPWB, (w ? sizeof(cd) : sizeof(cf)) * tot, &dst));
  const int conj_now = (zmelconjg && is_m_basis) ? 1 : 0;
  if (w)
    hipLaunchKernelGGL((zmel_stack_kernel<cd>), dim3(grid_for(tot)), dim3(256), 0, h->stream, (const cd*)mt, nbloch,
                       (const cd*)ipw, ngc, ncol, conj_now, (cd*)dst);
  else
    hipLaunchKernelGGL((zmel_stack_kernel<cf>), dim3(grid_for(tot)), dim3(256), 0, h->stream, (const cf*)mt, nbloch,
                       (const cf*)ipw, ngc, ncol, conj_now, (cf*)dst);
  ECALJ_HIP(hipGetLastError());
  if (!is_m_basis) {  // zmel = zcousq^H zmel_M (S/m_zmel.f90:559-560), then conj (:590-594)
    GemmCall e;
    e.mode = h->precision;
    e.ta = true; e.ca = true;
    e.M = nbb; e.N = (int)ncol; e.K = nbb;
    e.A = w ? (const void*)h->d_zcousq : (const void*)h->d_zcousq32; e.lda = nbb;
    e.B = dst; e.ldb = nbb;
    e.C = zmel; e.ldc = nbb;
    e.conjC = zmelconjg;
    ECALJ_TRY(gemm(h, e));
  }
  return ECALJ_OK;
}
