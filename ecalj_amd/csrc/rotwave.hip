// rotwave.hip -- eigenfunctions at any q of the full mesh from those at the
// irreducible points, as readcphi / readgeig return them
// (S/readeigen.f90:71-181, the device forms :183-393):
//   muffin-tin part (rotmto, S/rotwave.f90:156-192)
//     cphi(offset_rev(miat(ibas),l,k) + m, :) =
//       sum_m' dlmm(m, m', l) cphi_irr(offset(iorb) + m', :) exp(-2 pi i (g q).tiat(:,ibas))
//   plane-wave part (rotipw, S/rotwave.f90:223-245)
//     geig(ngvecprev(g (q+G) - qtarget), :) = geig_irr(G, :) exp(-2 pi i (g (q+G)).shtvg)
// for a batch of target points in one launch each.  The handle keeps the
// m_hamindex tables (the q list, its maps to the irreducible points and
// operations, G lists and their index cubes, the operations with miat, tiat,
// shtvg and dlmm) and the @MNLA_CPHI orbital tables (readmnla_cphi,
// S/readeigen.f90:487-521).  The per-target index maps and phases are built
// on the host (O(targets x G)), the band-wide work on the device.
// Host readers of the files these tables come from: @MNLA_CPHI, __QGpsi,
// __EValue, and the fixed-length __CPHI / __GEIG records straight to device.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "kernels.hpp"

struct RotState {
  // m_hamindex
  int nqtt = 0, ngpmx = 0, imx = -1, ngrp = 0, nbas = 0, lmxax = -1, nqi = 0;
  std::vector<double> qtt, symops, shtvg, tiat, plat, qlat;
  std::vector<int> iqmap, igmap, iqimap, ngp, ngvecp, ngvecprev, miat;
  double* d_dlmm = nullptr;
  // @MNLA_CPHI
  int norbt = 0, ndima = 0, mb = 0, ml = 0, mk = 0;
  std::vector<int> ibas, l, k, off, rev;
  // device index tables of the last target lists (reused when a call repeats them)
  std::vector<int> mto_iq, ipw_iq;
  int mto_nspc = 0;
};

namespace ecalj {

void rot_free(ecalj_ctx* h) {
  if (!h->rot) return;
  if (h->rot->d_dlmm) (void)hipFree(h->rot->d_dlmm);
  delete h->rot;
  h->rot = nullptr;
}

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

struct MtoItem {
  int src, dst, l, dl;  // rows (spin offset included), l, element offset of dlmm(., ., l, igg) at m = m' = 0
  double pr, pi;        // exp(-2 pi i (g q).tiat)
};

// one thread per (target, band, orbital block x spinor component): 2l+1 outputs
__global__ void rot_mto_kernel(const MtoItem* __restrict__ items, int nper, const int* __restrict__ kin, int nk,
                               int nband, const cd* __restrict__ in, int64_t ldc, int64_t in_stride,
                               const double* __restrict__ dlmm, int ldd, cd* __restrict__ out, int64_t out_stride) {
  const int64_t tot = (int64_t)nk * nband * nper;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int it = (int)(e % nper);
    const int64_t rest = e / nper;
    const int b = (int)(rest % nband);
    const int kt = (int)(rest / nband);
    const MtoItem m = items[(int64_t)kt * nper + it];
    const cd* x = in + kin[kt] * in_stride + ldc * b + m.src;
    cd* y = out + kt * out_stride + ldc * b + m.dst;
    const double* D = dlmm + m.dl;
    for (int a = -m.l; a <= m.l; ++a) {
      double sr = 0.0, si = 0.0;
      for (int c = -m.l; c <= m.l; ++c) {
        const double d = D[a + (int64_t)ldd * c];
        const cd v = x[c + m.l];
        sr += d * v.re;
        si += d * v.im;
      }
      y[a + m.l] = cd{sr * m.pr - si * m.pi, sr * m.pi + si * m.pr};
    }
  }
}

// one thread per (target, band, spinor component, source G)
__global__ void rot_ipw_kernel(const int* __restrict__ map, const cd* __restrict__ ph, const int* __restrict__ kin,
                               int nk, int ngpmx, int nspc, int nband, const cd* __restrict__ in, int64_t ldg,
                               int64_t in_stride, cd* __restrict__ out, int64_t out_stride) {
  const int64_t tot = (int64_t)nk * nband * nspc * ngpmx;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int ig = (int)(e % ngpmx);
    int64_t rest = e / ngpmx;
    const int isc = (int)(rest % nspc);
    rest /= nspc;
    const int b = (int)(rest % nband);
    const int kt = (int)(rest / nband);
    const int t = kt * ngpmx + ig;
    const int ig2 = map[t];
    if (ig2 < 0) continue;
    const int64_t ioff = (int64_t)isc * ngpmx;
    const cd v = in[kin[kt] * in_stride + ldg * b + ioff + ig];
    const cd p = ph[t];
    out[kt * out_stride + ldg * b + ioff + ig2] = cd{v.re * p.re - v.im * p.im, v.re * p.im + v.im * p.re};
  }
}

inline int blocks_for(int64_t tot) { return (int)std::max<int64_t>(1, std::min<int64_t>((tot + 255) / 256, 65536)); }

template <class T>
std::vector<unsigned char> as_blob(const std::vector<T>& v) {
  std::vector<unsigned char> b(sizeof(T) * v.size());
  if (!v.empty()) std::memcpy(b.data(), v.data(), b.size());
  return b;
}

// target iq (1-based): checks and the indices of the source point and operation
ecalj_status target(const RotState& r, int iq, int* iqq, int* igg, int* iqi) {
  if (iq < 1 || iq > r.nqtt) return ECALJ_ERR_DIM;
  *iqq = r.iqmap[iq - 1];
  *igg = r.igmap[iq - 1];
  *iqi = r.iqimap[iq - 1];
  return ECALJ_OK;
}

void matvec(const double* m, const double* v, double* o) {  // 3x3 column-major
  for (int i = 0; i < 3; ++i) o[i] = m[i] * v[0] + m[i + 3] * v[1] + m[i + 6] * v[2];
}

struct File {
  FILE* f = nullptr;
  File(const char* path, const char* mode) : f(fopen(path, mode)) {}
  ~File() {
    if (f) fclose(f);
  }
};

// one sequential unformatted record (4-byte markers); payload length or -1
long get_rec(FILE* f, std::vector<unsigned char>& buf) {
  int32_t m1 = 0, m2 = 0;
  if (fread(&m1, 4, 1, f) != 1 || m1 < 0) return -1;
  buf.resize((size_t)m1);
  if (m1 && fread(buf.data(), 1, (size_t)m1, f) != (size_t)m1) return -1;
  if (fread(&m2, 4, 1, f) != 1 || m2 != m1) return -1;
  return m1;
}

}  // namespace
}  // namespace ecalj

using namespace ecalj;

extern "C" {

ecalj_status ecalj_set_hamindex(ecalj_handle h, int nqtt, const double* qtt, const int* iqmap, const int* igmap,
                                const int* iqimap, const int* ngp, int ngpmx, int imx, const int* ngvecp,
                                const int* ngvecprev, int ngrp, const double* symops, const double* shtvg, int nbas,
                                const int* miat, const double* tiat, int lmxax, const double* dlmm,
                                const double* plat, const double* qlat) {
  if (!h || !qtt || !iqmap || !igmap || !iqimap || !ngp || !symops || !shtvg || !miat || !tiat || !dlmm || !plat ||
      !qlat)
    return ECALJ_ERR_ARG;
  if (nqtt < 1 || ngrp < 1 || nbas < 1 || lmxax < 0 || lmxax > 12 || ngpmx < 0 || imx < -1) return ECALJ_ERR_DIM;
  if (ngpmx > 0 && (!ngvecp || !ngvecprev || imx < 0)) return ECALJ_ERR_ARG;
  if (h->capturing) return ECALJ_ERR_STATE;
  RotState r;
  r.nqtt = nqtt; r.ngpmx = ngpmx; r.imx = imx; r.ngrp = ngrp; r.nbas = nbas; r.lmxax = lmxax;
  r.qtt.assign(qtt, qtt + 3 * (size_t)nqtt);
  r.iqmap.assign(iqmap, iqmap + nqtt);
  r.igmap.assign(igmap, igmap + nqtt);
  r.iqimap.assign(iqimap, iqimap + nqtt);
  r.ngp.assign(ngp, ngp + nqtt);
  for (int i = 0; i < nqtt; ++i) {
    if (r.iqmap[i] < 1 || r.iqmap[i] > nqtt || r.igmap[i] < 1 || r.igmap[i] > ngrp || r.iqimap[i] < 1 ||
        r.ngp[i] < 0 || r.ngp[i] > ngpmx)
      return ECALJ_ERR_DIM;
    r.nqi = std::max(r.nqi, r.iqimap[i]);
  }
  if (ngpmx > 0) {
    const size_t nc = 2 * (size_t)imx + 1;
    r.ngvecp.assign(ngvecp, ngvecp + 3 * (size_t)ngpmx * nqtt);
    r.ngvecprev.assign(ngvecprev, ngvecprev + nc * nc * nc * nqtt);
  }
  r.symops.assign(symops, symops + 9 * (size_t)ngrp);
  r.shtvg.assign(shtvg, shtvg + 3 * (size_t)ngrp);
  r.miat.assign(miat, miat + (size_t)nbas * ngrp);
  for (int v : r.miat)
    if (v < 1 || v > nbas) return ECALJ_ERR_DIM;
  r.tiat.assign(tiat, tiat + 3 * (size_t)nbas * ngrp);
  r.plat.assign(plat, plat + 9);
  r.qlat.assign(qlat, qlat + 9);
  const size_t ldd = 2 * (size_t)lmxax + 1, nd = ldd * ldd * (lmxax + 1) * ngrp;
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  ++h->state_gen;
  RotState* keep = h->rot ? h->rot : new RotState;
  if (keep->d_dlmm) (void)hipFree(keep->d_dlmm);
  keep->d_dlmm = nullptr;
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * nd) != hipSuccess) {
    if (!h->rot) delete keep;
    return ECALJ_ERR_NOMEM;
  }
  if (hipMemcpy(d, dlmm, sizeof(double) * nd, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    if (!h->rot) delete keep;
    return ECALJ_ERR_HIP;
  }
  // keep the orbital tables of an earlier ecalj_set_mnla
  r.norbt = keep->norbt; r.ndima = keep->ndima; r.mb = keep->mb; r.ml = keep->ml; r.mk = keep->mk;
  r.ibas = keep->ibas; r.l = keep->l; r.k = keep->k; r.off = keep->off; r.rev = keep->rev;
  r.d_dlmm = d;
  r.mto_iq.clear();
  r.ipw_iq.clear();
  *keep = r;
  h->rot = keep;
  return ECALJ_OK;
}

// readmnla_cphi (S/readeigen.f90:487-521): offset = first row of the block - 1
ecalj_status ecalj_set_mnla(ecalj_handle h, int norbt, const int* ibas_tbl, const int* l_tbl, const int* k_tbl,
                            const int* offset_tbl) {
  if (!h || !ibas_tbl || !l_tbl || !k_tbl || !offset_tbl) return ECALJ_ERR_ARG;
  if (norbt < 1) return ECALJ_ERR_DIM;
  RotState* r = h->rot ? h->rot : (h->rot = new RotState);
  r->norbt = norbt;
  r->ibas.assign(ibas_tbl, ibas_tbl + norbt);
  r->l.assign(l_tbl, l_tbl + norbt);
  r->k.assign(k_tbl, k_tbl + norbt);
  r->off.assign(offset_tbl, offset_tbl + norbt);
  r->mb = r->ml = r->mk = 0;
  r->ndima = 0;
  for (int i = 0; i < norbt; ++i) {
    if (r->ibas[i] < 1 || r->l[i] < 0 || r->k[i] < 1 || r->off[i] < 0) return ECALJ_ERR_DIM;
    r->mb = std::max(r->mb, r->ibas[i]);
    r->ml = std::max(r->ml, r->l[i]);
    r->mk = std::max(r->mk, r->k[i]);
    r->ndima = std::max(r->ndima, r->off[i] + 2 * r->l[i] + 1);
  }
  r->mto_iq.clear();
  r->rev.assign((size_t)r->mb * (r->ml + 1) * r->mk, -1);
  for (int i = 0; i < norbt; ++i)
    r->rev[(r->ibas[i] - 1) + (size_t)r->mb * (r->l[i] + (size_t)(r->ml + 1) * (r->k[i] - 1))] = r->off[i];
  return ECALJ_OK;
}

ecalj_status ecalj_readcphi_dev(ecalj_handle h, int nk, const int* iq, int nspc, int nband, const void* cphi_irr,
                                int ldc, void* cphi_out) {
  if (!h || !iq || !cphi_irr || !cphi_out) return ECALJ_ERR_ARG;
  RotState* r = h->rot;
  if (!r || !r->d_dlmm || r->norbt == 0) return ECALJ_ERR_STATE;
  if (nk < 1 || nband < 1 || (nspc != 1 && nspc != 2) || ldc < r->ndima * nspc) return ECALJ_ERR_DIM;
  if (r->ml > r->lmxax || r->mb > r->nbas) return ECALJ_ERR_DIM;
  const int nper = r->norbt * nspc;
  const int ldd = 2 * r->lmxax + 1;
  const bool reuse = !h->capturing && r->mto_nspc == nspc && r->mto_iq.size() == (size_t)nk &&
                     std::equal(iq, iq + nk, r->mto_iq.begin()) && h->slot[ecalj_ctx::S_ROT1] &&
                     h->slot[ecalj_ctx::S_ROT2];
  std::vector<MtoItem> items(reuse ? 0 : (size_t)nk * nper);
  std::vector<int> kin(reuse ? 0 : nk);
  for (int t = 0; t < (reuse ? 0 : nk); ++t) {
    int iqq, igg, iqi;
    ECALJ_TRY(target(*r, iq[t], &iqq, &igg, &iqi));
    kin[t] = iqi - 1;
    double qrot[3];
    matvec(&r->symops[9 * (size_t)(igg - 1)], &r->qtt[3 * (size_t)(iqq - 1)], qrot);
    for (int io = 0; io < r->norbt; ++io) {
      const int ib = r->ibas[io], l = r->l[io], k = r->k[io];
      const int jb = r->miat[(ib - 1) + (size_t)r->nbas * (igg - 1)];
      if (jb > r->mb) return ECALJ_ERR_DIM;
      const int dst = r->rev[(jb - 1) + (size_t)r->mb * (l + (size_t)(r->ml + 1) * (k - 1))];
      if (dst < 0) return ECALJ_ERR_DIM;  // the operation maps the block onto none of the tables
      const double* ti = &r->tiat[3 * ((ib - 1) + (size_t)r->nbas * (igg - 1))];
      const double ph = -kTwoPi * (qrot[0] * ti[0] + qrot[1] * ti[1] + qrot[2] * ti[2]);
      for (int s = 0; s < nspc; ++s) {
        MtoItem& m = items[(size_t)t * nper + io + (size_t)r->norbt * s];
        m.src = r->off[io] + r->ndima * s;
        m.dst = dst + r->ndima * s;
        m.l = l;
        m.dl = ldd * ldd * (l + (r->lmxax + 1) * (igg - 1)) + r->lmxax * (1 + ldd);
        m.pr = std::cos(ph);
        m.pi = std::sin(ph);
      }
    }
  }
  unsigned char *d_items = (unsigned char*)h->slot[ecalj_ctx::S_ROT1], *d_kin = (unsigned char*)h->slot[ecalj_ctx::S_ROT2];
  if (!reuse) {
    r->mto_iq.clear();
    ECALJ_TRY(upload_blob(h, ecalj_ctx::S_ROT1, as_blob(items), &d_items));
    ECALJ_TRY(upload_blob(h, ecalj_ctx::S_ROT2, as_blob(kin), &d_kin));
    r->mto_iq.assign(iq, iq + nk);
    r->mto_nspc = nspc;
  }
  const int64_t stride = (int64_t)ldc * nband;
  hipLaunchKernelGGL(rot_mto_kernel, dim3(blocks_for((int64_t)nk * nband * nper)), dim3(256), 0, h->stream,
                     (const MtoItem*)d_items, nper, (const int*)d_kin, nk, nband, (const cd*)cphi_irr, (int64_t)ldc,
                     stride, r->d_dlmm, ldd, (cd*)cphi_out, stride);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status ecalj_readgeig_dev(ecalj_handle h, int nk, const int* iq, int nspc, int nband, const void* geig_irr,
                                int ldg, void* geig_out) {
  if (!h || !iq || !geig_irr || !geig_out) return ECALJ_ERR_ARG;
  RotState* r = h->rot;
  if (!r || r->nqtt == 0) return ECALJ_ERR_STATE;
  if (nk < 1 || nband < 1 || (nspc != 1 && nspc != 2) || ldg < r->ngpmx * nspc) return ECALJ_ERR_DIM;
  const int64_t stride = (int64_t)ldg * nband;
  ECALJ_TRY(dev_zero(h, geig_out, sizeof(cd) * (size_t)stride * nk));  // S/readeigen.f90:68, 200-202
  if (r->ngpmx == 0) return ECALJ_OK;
  const int nc = 2 * r->imx + 1, ngpmx = r->ngpmx;
  double platt[9];  // transpose(plat), the inverse of qlat
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) platt[i + 3 * j] = r->plat[j + 3 * i];
  const bool reuse = !h->capturing && r->ipw_iq.size() == (size_t)nk && std::equal(iq, iq + nk, r->ipw_iq.begin()) &&
                     h->slot[ecalj_ctx::S_ROT3];
  std::vector<int> map(reuse ? 0 : (size_t)nk * ngpmx, -1), kin(reuse ? 0 : nk);
  std::vector<cd> ph(reuse ? 0 : (size_t)nk * ngpmx, cd{0.0, 0.0});
  for (int t = 0; t < (reuse ? 0 : nk); ++t) {
    int iqq, igg, iqi;
    ECALJ_TRY(target(*r, iq[t], &iqq, &igg, &iqi));
    kin[t] = iqi - 1;
    const int n = r->ngp[iqq - 1];
    if (r->ngp[iq[t] - 1] != n) return ECALJ_ERR_DIM;  // S/readeigen.f90:83-89
    const double* qin = &r->qtt[3 * (size_t)(iqq - 1)];
    const double* qt = &r->qtt[3 * (size_t)(iq[t] - 1)];
    const double* g = &r->symops[9 * (size_t)(igg - 1)];
    const double* sh = &r->shtvg[3 * (size_t)(igg - 1)];
    const int* cube = &r->ngvecprev[(size_t)nc * nc * nc * (iq[t] - 1)];
    for (int ig = 0; ig < n; ++ig) {
      const int* G = &r->ngvecp[3 * (ig + (size_t)ngpmx * (iqq - 1))];
      double qpg[3], qpgr[3], d[3];
      for (int i = 0; i < 3; ++i)
        qpg[i] = qin[i] + r->qlat[i] * G[0] + r->qlat[i + 3] * G[1] + r->qlat[i + 6] * G[2];
      matvec(g, qpg, qpgr);
      for (int i = 0; i < 3; ++i) d[i] = qpgr[i] - qt[i];
      int nn[3];
      for (int i = 0; i < 3; ++i) {
        nn[i] = (int)std::lround(platt[i] * d[0] + platt[i + 3] * d[1] + platt[i + 6] * d[2]);
        if (nn[i] < -r->imx || nn[i] > r->imx) return ECALJ_ERR_DIM;
      }
      const int ig2 = cube[(nn[0] + r->imx) + nc * ((nn[1] + r->imx) + nc * (nn[2] + r->imx))];
      if (ig2 < 1 || ig2 > n) return ECALJ_ERR_DIM;  // rotated G not in the list of qtarget
      const double a = -kTwoPi * (qpgr[0] * sh[0] + qpgr[1] * sh[1] + qpgr[2] * sh[2]);
      map[(size_t)t * ngpmx + ig] = ig2 - 1;
      ph[(size_t)t * ngpmx + ig] = cd{std::cos(a), std::sin(a)};
    }
  }
  // one table: map (int, nk*ngpmx), phases (16-byte aligned), source blocks
  const size_t o2 = (sizeof(int) * (size_t)nk * ngpmx + 15) / 16 * 16, o3 = o2 + sizeof(cd) * (size_t)nk * ngpmx;
  unsigned char* d = (unsigned char*)h->slot[ecalj_ctx::S_ROT3];
  if (!reuse) {
    r->ipw_iq.clear();
    std::vector<unsigned char> blob(o3 + sizeof(int) * (size_t)nk);
    std::memcpy(blob.data(), map.data(), sizeof(int) * map.size());
    std::memcpy(blob.data() + o2, ph.data(), sizeof(cd) * ph.size());
    std::memcpy(blob.data() + o3, kin.data(), sizeof(int) * kin.size());
    ECALJ_TRY(upload_blob(h, ecalj_ctx::S_ROT3, blob, &d));
    r->ipw_iq.assign(iq, iq + nk);
  }
  hipLaunchKernelGGL(rot_ipw_kernel, dim3(blocks_for((int64_t)nk * nband * nspc * ngpmx)), dim3(256), 0, h->stream,
                     (const int*)d, (const cd*)(d + o2), (const int*)(d + o3), nk, ngpmx, nspc, nband,
                     (const cd*)geig_irr, (int64_t)ldg, stride, (cd*)geig_out, stride);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// ---- readers of the front end's files
// @MNLA_CPHI (written by sugw, S/sugw.f90:261-277): a header line, then
// "m n l ibas iord iorb" per row; tables from the first row of each block.
ecalj_status ecalj_mnla_read(const char* path, int cap, int* norbt, int* ibas_tbl, int* l_tbl, int* k_tbl,
                             int* offset_tbl) {
  if (!path || !norbt) return ECALJ_ERR_ARG;
  File fl(path, "r");
  if (!fl.f) return ECALJ_ERR_ARG;
  char line[512];
  if (!fgets(line, sizeof line, fl.f)) return ECALJ_ERR_DIM;
  int n = 0, prev = 0;
  int m, nn, l, ib, iord, io;
  while (fscanf(fl.f, "%d %d %d %d %d %d", &m, &nn, &l, &ib, &iord, &io) == 6) {
    if (io < 1) return ECALJ_ERR_DIM;
    n = std::max(n, io);
    if (io != prev) {
      if (io <= cap && ibas_tbl && l_tbl && k_tbl && offset_tbl) {
        k_tbl[io - 1] = nn;
        l_tbl[io - 1] = l;
        ibas_tbl[io - 1] = ib;
        offset_tbl[io - 1] = iord - 1;
      }
      prev = io;
    }
  }
  *norbt = n;
  if (n < 1) return ECALJ_ERR_DIM;
  return (n > cap || !ibas_tbl || !l_tbl || !k_tbl || !offset_tbl) ? ECALJ_ERR_DIM : ECALJ_OK;
}

// __QGpsi (mkqg, S/mkqg.f90:298,349-356; read by init_readeigen,
// S/readeigen.f90:404-424): header nqnum ngpmx QpGcut_psi nqbz nqi imx nqibz;
// per q: (q(3), ngp, irr), (ngvecp(3, max(ngp,1)), ngvecprev(-imx:imx)^3).
// hdr6 = nqnum ngpmx nqbz nqi imx nqibz.  First call with capacities 0
// reports hdr6 (ECALJ_ERR_DIM); then ngvecp is 3 x ngpmx x nqnum and
// ngvecprev (2 imx+1)^3 x nqnum.
ecalj_status ecalj_qgpsi_read(const char* path, int* hdr6, double* qpgcut, int nq_cap, int ngpmx_cap, int imx_cap,
                              double* qtt, int* ngp, int* irr, int* ngvecp, int* ngvecprev) {
  if (!path || !hdr6) return ECALJ_ERR_ARG;
  File fl(path, "rb");
  if (!fl.f) return ECALJ_ERR_ARG;
  std::vector<unsigned char> b;
  if (get_rec(fl.f, b) < 32) return ECALJ_ERR_DIM;
  int nqnum, ngpmx, nqbz, nqi, imx, nqibz;
  double cut;
  std::memcpy(&nqnum, &b[0], 4);
  std::memcpy(&ngpmx, &b[4], 4);
  std::memcpy(&cut, &b[8], 8);
  std::memcpy(&nqbz, &b[16], 4);
  std::memcpy(&nqi, &b[20], 4);
  std::memcpy(&imx, &b[24], 4);
  std::memcpy(&nqibz, &b[28], 4);
  const int h6[6] = {nqnum, ngpmx, nqbz, nqi, imx, nqibz};
  std::memcpy(hdr6, h6, sizeof h6);
  if (qpgcut) *qpgcut = cut;
  if (nqnum < 0 || ngpmx < 0 || imx < 0) return ECALJ_ERR_DIM;
  if (nqnum > nq_cap || ngpmx > ngpmx_cap || imx > imx_cap || !qtt || !ngp || !irr || !ngvecp || !ngvecprev)
    return ECALJ_ERR_DIM;
  const size_t nc = 2 * (size_t)imx + 1, ncube = nc * nc * nc;
  for (int iq = 0; iq < nqnum; ++iq) {
    if (get_rec(fl.f, b) < 32) return ECALJ_ERR_DIM;
    std::memcpy(&qtt[3 * (size_t)iq], &b[0], 24);
    std::memcpy(&ngp[iq], &b[24], 4);
    std::memcpy(&irr[iq], &b[28], 4);
    const int n = ngp[iq];
    if (n < 0 || n > ngpmx) return ECALJ_ERR_DIM;
    const size_t ng = 3 * (size_t)std::max(n, 1);
    if (get_rec(fl.f, b) != (long)(4 * (ng + ncube))) return ECALJ_ERR_DIM;
    int* gv = &ngvecp[3 * (size_t)ngpmx * iq];
    std::fill(gv, gv + 3 * (size_t)ngpmx, 0);
    std::memcpy(gv, b.data(), 4 * 3 * (size_t)n);
    std::memcpy(&ngvecprev[ncube * iq], b.data() + 4 * ng, 4 * ncube);
  }
  return ECALJ_OK;
}

// __EValue (sugw; read by init_readeigen, S/readeigen.f90:428-440):
// (nband, nqi, nsp, nspc), qtti(3, nqi), evud(nband, nqi, nspx).  hdr5 =
// nband nqi nsp nspc nspx; first call with capacity 0 reports hdr5.
ecalj_status ecalj_evalue_read(const char* path, int* hdr5, int64_t cap, double* qtti, double* evud) {
  if (!path || !hdr5) return ECALJ_ERR_ARG;
  File fl(path, "rb");
  if (!fl.f) return ECALJ_ERR_ARG;
  std::vector<unsigned char> b;
  if (get_rec(fl.f, b) != 16) return ECALJ_ERR_DIM;
  int h4[4];
  std::memcpy(h4, b.data(), 16);
  std::vector<unsigned char> bq, be;
  if (get_rec(fl.f, bq) != (long)(24 * (size_t)std::max(h4[1], 0))) return ECALJ_ERR_DIM;
  const long ne = get_rec(fl.f, be);
  const int64_t per = (int64_t)h4[0] * h4[1];
  if (ne < 0 || per <= 0 || ne % (8 * per) != 0) return ECALJ_ERR_DIM;
  const int nspx = (int)(ne / (8 * per));
  const int h5[5] = {h4[0], h4[1], h4[2], h4[3], nspx};
  std::memcpy(hdr5, h5, sizeof h5);
  if (per * nspx > cap || !qtti || !evud) return ECALJ_ERR_DIM;
  std::memcpy(qtti, bq.data(), bq.size());
  std::memcpy(evud, be.data(), be.size());
  return ECALJ_OK;
}

// Fixed-length records (openm/readm, S/m_mpiio.f90; __CPHI recl mrecb and
// __GEIG recl mrecg, S/sugw.f90:234-236): records rec0, rec0+step, ... (nrec
// of them), the first nbytes of each, to dst + i*dst_stride bytes in device
// memory, through pinned staging.  For the irreducible eigenvectors of spin
// isp: rec0 = isp, step = nsp (ikpisp = isp + nsp*(iqi-1)).
ecalj_status ecalj_rec_read_dev(ecalj_handle h, const char* path, int64_t recl, int rec0, int step, int nrec,
                                int64_t nbytes, void* dst, int64_t dst_stride) {
  if (!h || !path || !dst) return ECALJ_ERR_ARG;
  if (recl < 1 || rec0 < 1 || step < 1 || nrec < 0 || nbytes < 0 || nbytes > recl || dst_stride < nbytes)
    return ECALJ_ERR_DIM;
  if (h->capturing) return ECALJ_ERR_STATE;
  File fl(path, "rb");
  if (!fl.f) return ECALJ_ERR_ARG;
  void* pin = nullptr;
  if (hipHostMalloc(&pin, std::max<int64_t>(nbytes, 1)) != hipSuccess) return ECALJ_ERR_NOMEM;
  ecalj_status st = ECALJ_OK;
  for (int i = 0; i < nrec && st == ECALJ_OK; ++i) {
    const int64_t rec = rec0 + (int64_t)step * i;
    if (fseeko(fl.f, (off_t)(recl * (rec - 1)), SEEK_SET) != 0 || fread(pin, 1, (size_t)nbytes, fl.f) != (size_t)nbytes) {
      st = ECALJ_ERR_DIM;  // short file
      break;
    }
    if (hipMemcpyAsync((char*)dst + dst_stride * i, pin, (size_t)nbytes, hipMemcpyHostToDevice, h->stream) !=
            hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      st = ECALJ_ERR_HIP;
  }
  (void)hipHostFree(pin);
  return st;
}

}  // extern "C"
