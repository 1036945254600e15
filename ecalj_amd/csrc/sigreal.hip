// sigreal.hip -- real-axis pole term of Sigma_c (S/m_sxcf_gemm.f90:501-594, 599-608).
//
// For every (middle state it, external state itp) whose pole lies between the
// QP energy and the Fermi level, three real frequencies around
// we = |omega - weavx2|/2 carry 3-point omega^2 interpolation weights, and
//   czmelwc(:,it,itp) += sum_j wgt_j Wt(iw_j)^H zmel(:,it,itp),
//   Wt(iw)(I,J) = (I <= J) ? W(iw)(I,J) : conj(W(iw)(J,I))   (keepwv packing).
// The pole lists are built on the host, as the reference builds them.  On the
// device the term is a gather - skinny GEMM - scatter: one work item per
// (k, iw, group of <= 16 entries, 64-row tile) streams the rows of Wt(iw)
// once and writes one output column per entry; a second kernel sums each
// pole's three columns in frequency order and adds v^H zmel to zsec.  Both
// run in FP64 whatever the storage type of zmel.  The term reads each used
// W(iw) once per k: it is HBM-bound.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace ecalj {

constexpr int RT = 64, RJ = 32, RE = 16;  // rows per item, J tile, entries per item

struct RealItem {
  int k, iw, e0, ne, i0, pad;
};

// pk (optional): the internal k of each pole, for items whose entries come
// from several k-points (W shared by every k: items grouped by frequency only).
template <class T>
__global__ __launch_bounds__(256) void sig_real_entries_kernel(const RealItem* __restrict__ items, int item0,
                                                               const int* __restrict__ perm,
                                                               const int* __restrict__ ecol,
                                                               const double* __restrict__ ew,
                                                               const int* __restrict__ pk,
                                                               const T* __restrict__ zmel, int64_t zstride,
                                                               const cd* __restrict__ wr, int64_t ldw,
                                                               int64_t wstride, int n, cd* __restrict__ slot) {
  __shared__ cd wt[RT][RJ + 1];  // Wt(I0+r, J0+jj)
  __shared__ cd zb[RJ][RE];      // w_e * zmel(J0+jj, col_e)
  const RealItem it = items[item0 + blockIdx.x];
  const int t = threadIdx.x;
  const cd* W = wr + wstride * it.k + ldw * n * (int64_t)it.iw;
  const T* Z = zmel + zstride * it.k;
  const int r = t % RT, eg = t / RT;  // output row, entry group of 4
  double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
  for (int J0 = 0; J0 < n; J0 += RJ) {
    // Wt tile: upper part (I <= J) read along I, lower part (I > J) read along J
    for (int q = t; q < RT * RJ; q += 256) {
      const int rr = q % RT, jj = q / RT, I = it.i0 + rr, J = J0 + jj;
      if (I <= J && I < n && J < n) wt[rr][jj] = W[I + ldw * J];
    }
    for (int q = t; q < RT * RJ; q += 256) {
      const int jj = q % RJ, rr = q / RJ, I = it.i0 + rr, J = J0 + jj;
      if (I < n && J < n) {
        if (I > J) {
          const cd v = W[J + ldw * I];
          wt[rr][jj] = cd{v.re, -v.im};
        }
      } else {
        wt[rr][jj] = cd{0.0, 0.0};
      }
    }
    for (int q = t; q < RJ * RE; q += 256) {
      const int jj = q % RJ, e = q / RJ, J = J0 + jj;
      cd v{0.0, 0.0};
      if (e < it.ne && J < n) {
        const int id = perm[it.e0 + e];
        const T* Ze = pk ? zmel + zstride * pk[id / 3] : Z;
        double xr, xi;
        ld_c2(Ze + (int64_t)n * ecol[id] + J, xr, xi);
        v = cd{ew[id] * xr, ew[id] * xi};
      }
      zb[jj][e] = v;
    }
    __syncthreads();
    const int jmax = min(RJ, n - J0);
    for (int jj = 0; jj < jmax; ++jj) {
      const cd a = wt[r][jj];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const cd b = zb[jj][4 * eg + q];
        ar[q] += a.re * b.re - a.im * b.im;
        ai[q] += a.re * b.im + a.im * b.re;
      }
    }
    __syncthreads();
  }
  const int I = it.i0 + r;
  if (I < n) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = 4 * eg + q;
      if (e < it.ne) slot[(int64_t)n * perm[it.e0 + e] + I] = cd{ar[q], ai[q]};
    }
  }
}

// zsec(m, m') += sum over poles p of external state m, k by k, of
// sum_I conj(v_p(I)) zmel_k(I, it_p, m'), v_p = slot[3p] + slot[3p+1] + slot[3p+2]
// (frequency order); Im(zsec_mm) <= 0 after every k, as the reference clips
// after every (k, band block) call.  plist holds each m's poles in k order.
// One block per (m, m').
template <class T>
__global__ __launch_bounds__(256) void sig_real_final_kernel(const int* __restrict__ pstart,
                                                             const int* __restrict__ plist,
                                                             const int* __restrict__ pk,
                                                             const int* __restrict__ pit, const cd* __restrict__ slot,
                                                             const T* __restrict__ zmel, int64_t zstride, int n,
                                                             int nm, int ntq, int nk, cd* __restrict__ zsec,
                                                             int64_t ldz) {
  __shared__ double sr[256], si[256];
  const int m = blockIdx.x % ntq, mp = blockIdx.x / ntq;
  cd* zs = zsec + m + ldz * mp;
  const cd z0 = *zs;
  double tr = z0.re, ti = z0.im;
  int q = pstart[m];
  const int qend = pstart[m + 1];
  for (int k = 0; k < nk; ++k) {
    for (; q < qend && pk[plist[q]] == k; ++q) {
      const int p = plist[q];
      const cd* s0 = slot + (int64_t)n * 3 * p;
      const T* Z = zmel + zstride * k + (int64_t)n * (pit[p] + (int64_t)nm * mp);
      double a = 0.0, b = 0.0;
      for (int I = threadIdx.x; I < n; I += 256) {
        const cd v0 = s0[I], v1 = s0[n + I], v2 = s0[2 * n + I];
        const double vr = (v0.re + v1.re) + v2.re, vi = (v0.im + v1.im) + v2.im;
        double zr, zi;
        ld_c2(Z + I, zr, zi);
        a += vr * zr + vi * zi;  // conj(v) z
        b += vr * zi - vi * zr;
      }
      sr[threadIdx.x] = a;
      si[threadIdx.x] = b;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
          sr[threadIdx.x] += sr[threadIdx.x + w];
          si[threadIdx.x] += si[threadIdx.x + w];
        }
        __syncthreads();
      }
      tr += sr[0];
      ti += si[0];
      __syncthreads();
    }
    if (m == mp && ti > 0.0) ti = 0.0;
  }
  if (threadIdx.x == 0) *zs = cd{tr, ti};
}

// czmelwc_k(:, it_p + nm*itp_p) += slot[3p] + slot[3p+1] + slot[3p+2] (frequency
// order) for the poles p of k in [k0, k1); cz of k at cz + czstride*(k - k0).
// No two poles of one k share a column.
template <class T>
__global__ void sig_real_add_kernel(const int* __restrict__ pk, const int* __restrict__ pcol, int p0, int p1, int k0,
                                    const cd* __restrict__ slot, int n, T* __restrict__ cz, int64_t czstride) {
  const int64_t tot = (int64_t)(p1 - p0) * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int p = p0 + (int)(e / n), I = (int)(e % n);
    const cd* s0 = slot + (int64_t)n * 3 * p;
    const cd v0 = s0[I], v1 = s0[n + I], v2 = s0[2 * n + I];
    T* c = cz + czstride * (pk[p] - k0) + (int64_t)n * pcol[p] + I;
    double xr, xi;
    ld_c2(c, xr, xi);
    xr += (v0.re + v1.re) + v2.re;
    xi += (v0.im + v1.im) + v2.im;
    st_c(c, xr, xi);
  }
}

// ---------------------------------------------------------------- host
static double wfacx2_h(double e1, double e2, double ek, double esmr) {  // S/wfacx.f90:19-35
  const double el = std::min(e1, e2), eh = std::max(e1, e2);
  if (eh - el < 1e-6) return 0.0;
  if (esmr == 0.0) return (el <= ek && ek < eh) ? 1.0 : 0.0;
  return 0.5 * std::erfc(-(eh - ek) / std::sqrt(2.0) / esmr) - 0.5 * std::erfc(-(el - ek) / std::sqrt(2.0) / esmr);
}

static double weavx2_h(double e1, double e2, double ek, double esmr) {  // S/wfacx.f90:37-49
  const double pi = 3.14159265358979323846;
  const double el = std::min(e1, e2), eh = std::max(e1, e2);
  const double wtt = 0.5 * std::erfc(-(eh - ek) / std::sqrt(2.0) / esmr) -
                     0.5 * std::erfc(-(el - ek) / std::sqrt(2.0) / esmr);
  const double s2 = 2.0 * esmr * esmr;
  return ek + esmr / std::sqrt(2.0 * pi) * (-std::exp(-(eh - ek) * (eh - ek) / s2) +
                                            std::exp(-(el - ek) * (el - ek) / s2)) / wtt;
}

// 3x3 inverse by cofactors; a[i][j] = row i, column j
static void inv3(const double a[3][3], double r[3][3]) {
  double c[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
      c[i][j] = a[i1][j1] * a[i2][j2] - a[i1][j2] * a[i2][j1];  // cofactor of (j, i): adjugate entry (i, j)
    }
  const double det = a[0][0] * c[0][0] + a[0][1] * c[1][0] + a[0][2] * c[2][0];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = c[i][j] / det;
}

// Pole lists of nb internal k-points (host, as S/m_sxcf_gemm.f90:512-552
// builds them) and their device table.  Poles are generated k by k; the
// entries of pole p are 3p..3p+2 (frequency order).
struct RealPlan {
  int nb = 0, np = 0, ne = 0;
  std::vector<int> pk, pit, pitp, pcol, eiw, ecol;
  std::vector<double> ew;
  std::vector<int> perm, pstart, plist;
  std::vector<RealItem> items;
  std::vector<int> kitem, kpole;  // first item / pole of each k (nb+1)
  bool shared = false;            // one W for every k: items grouped by frequency across k
  mutable bool entries_done = false;
  unsigned char* d = nullptr;     // device table
  size_t o_it = 0, o_pm = 0, o_ec = 0, o_ps = 0, o_pl = 0, o_pk = 0, o_pt = 0, o_pc = 0, o_ew = 0;
  cd* slot = nullptr;
};

static ecalj_status real_plan(ecalj_ctx* h, const RealSpec& rs, int nb, int n, int nm, int ntq, RealPlan& P) {
  const double ddw = 10.0, wfaccut = 1e-8;  // S/m_sxcf_gemm.f90:264,408
  const int nw = rs.nw;
  P.nb = nb;
  P.kpole.assign(nb + 1, 0);
  for (int k = 0; k < nb; ++k) {
    P.kpole[k] = (int)P.pk.size();
    const double* e = rs.ekc + (int64_t)nm * k;
    int nt0p = 0, nt0m = 0;
    for (int i = 0; i < nm; ++i) {
      nt0p += e[i] < rs.ef + ddw * rs.esmr;
      nt0m += e[i] < rs.ef - ddw * rs.esmr;
    }
    for (int itp = 0; itp < ntq; ++itp) {
      const double omg = rs.omega[itp];
      const int itini = (omg >= rs.ef) ? std::max(1, nt0m + 1) : 1;
      const int itend = (omg >= rs.ef) ? rs.ns2r : std::min(nt0p, rs.ns2r);
      for (int it = itini; it <= itend; ++it) {
        double wf = wfacx2_h(omg, rs.ef, e[it - 1], it <= rs.nctot ? 0.0 : rs.esmr);
        if (wf < wfaccut) continue;
        wf *= rs.wkkr[k] * ((omg - rs.ef) >= 0.0 ? 1.0 : -1.0);
        const double we = 0.5 * std::fabs(omg - weavx2_h(omg, rs.ef, e[it - 1], rs.esmr));
        int ixs = 0;
        for (int i = 1; i <= nw; ++i)
          if (rs.freq_r[i] > we) {
            ixs = i;
            break;
          }
        if (ixs < 1 || ixs + 1 > nw) return ECALJ_ERR_DIM;  // pole outside the real mesh
        double a[3][3], ai[3][3];
        for (int i = 0; i < 3; ++i) {
          const double x = rs.freq_r[ixs - 1 + i];
          a[i][0] = 1.0;
          a[i][1] = x * x;
          a[i][2] = x * x * x * x;
        }
        inv3(a, ai);
        const double xv[3] = {1.0, we * we, we * we * we * we};
        P.pk.push_back(k);
        P.pit.push_back(it - 1);
        P.pitp.push_back(itp);
        P.pcol.push_back((it - 1) + nm * itp);
        for (int j = 0; j < 3; ++j) {
          P.eiw.push_back(ixs - 1 + j);
          P.ecol.push_back((it - 1) + nm * itp);
          P.ew.push_back(wf * (xv[0] * ai[0][j] + xv[1] * ai[1][j] + xv[2] * ai[2][j]));
        }
      }
    }
  }
  P.kpole[nb] = P.np = (int)P.pk.size();
  P.ne = 3 * P.np;
  if (P.np == 0) return ECALJ_OK;
  // Entries ordered by (k, iw), items of <= RE entries of one (k, iw) x row
  // tiles, items of k contiguous.  With one W for every k (wstride 0) the
  // entries of all k are ordered by iw and an item takes entries of several
  // k: each W(iw) row tile is then read once per RE entries of the whole
  // call, not once per k.
  P.shared = rs.wstride == 0;
  P.perm.resize(P.ne);
  for (int e = 0; e < P.ne; ++e) P.perm[e] = e;
  std::stable_sort(P.perm.begin(), P.perm.end(), [&](int x, int y) {
    const int kx = P.shared ? 0 : P.pk[x / 3], ky = P.shared ? 0 : P.pk[y / 3];
    return kx != ky ? kx < ky : P.eiw[x] < P.eiw[y];
  });
  P.kitem.assign(nb + 1, 0);
  for (int s0 = 0; s0 < P.ne;) {
    const int k = P.shared ? 0 : P.pk[P.perm[s0] / 3], iw = P.eiw[P.perm[s0]];
    int e = s0;
    while (e < P.ne && (P.shared || P.pk[P.perm[e] / 3] == k) && P.eiw[P.perm[e]] == iw && e - s0 < RE) ++e;
    for (int i0 = 0; i0 < n; i0 += RT) P.items.push_back({k, iw, s0, e - s0, i0, 0});
    s0 = e;
  }
  if (!P.shared) {
    int q = 0;
    for (int k = 0; k <= nb; ++k) {
      while (q < (int)P.items.size() && P.items[q].k < k) ++q;
      P.kitem[k] = q;
    }
  }
  // poles grouped by external state, k order kept
  P.pstart.assign(ntq + 1, 0);
  P.plist.resize(P.np);
  for (int p = 0; p < P.np; ++p) P.pstart[P.pitp[p] + 1]++;
  for (int m = 0; m < ntq; ++m) P.pstart[m + 1] += P.pstart[m];
  {
    std::vector<int> fill(P.pstart.begin(), P.pstart.end() - 1);
    for (int p = 0; p < P.np; ++p) P.plist[fill[P.pitp[p]]++] = p;
  }
  std::vector<unsigned char> blob;
  auto put = [&](const void* src, size_t bytes) {
    const size_t o = blob.size();
    blob.resize(o + ((bytes + 7) / 8) * 8);
    if (bytes) memcpy(blob.data() + o, src, bytes);
    return o;
  };
  P.o_it = put(P.items.data(), sizeof(RealItem) * P.items.size());
  P.o_pm = put(P.perm.data(), sizeof(int) * P.ne);
  P.o_ec = put(P.ecol.data(), sizeof(int) * P.ne);
  P.o_ps = put(P.pstart.data(), sizeof(int) * (ntq + 1));
  P.o_pl = put(P.plist.data(), sizeof(int) * P.np);
  P.o_pk = put(P.pk.data(), sizeof(int) * P.np);
  P.o_pt = put(P.pit.data(), sizeof(int) * P.np);
  P.o_pc = put(P.pcol.data(), sizeof(int) * P.np);
  P.o_ew = put(P.ew.data(), sizeof(double) * P.ne);
  ECALJ_TRY(upload_blob(h, ecalj_ctx::S_RTAB, blob, &P.d));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_RSLOT, sizeof(cd) * (size_t)n * P.ne, (void**)&P.slot));
  return ECALJ_OK;
}

// entry columns of the k-points [k0, k1) into the slots (shared W: the
// entries of every k, on the first call)
template <class T>
static ecalj_status real_entries(ecalj_ctx* h, const RealPlan& P, const RealSpec& rs, const T* zmel,
                                 int64_t zstride, int n, int k0, int k1) {
  int i0 = P.kitem[k0], i1 = P.kitem[k1];
  if (P.shared) {
    if (P.entries_done) return ECALJ_OK;
    i0 = 0;
    i1 = (int)P.items.size();
    P.entries_done = true;
  }
  if (i1 <= i0) return ECALJ_OK;
  hipLaunchKernelGGL((sig_real_entries_kernel<T>), dim3((unsigned)(i1 - i0)), dim3(256), 0, h->stream,
                     (const RealItem*)(P.d + P.o_it), i0, (const int*)(P.d + P.o_pm), (const int*)(P.d + P.o_ec),
                     (const double*)(P.d + P.o_ew), P.shared ? (const int*)(P.d + P.o_pk) : nullptr, zmel, zstride,
                     rs.wr, rs.ldw, rs.wstride, n, P.slot);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

template <class T>
static ecalj_status real_into_cz(ecalj_ctx* h, const RealPlan& P, const RealSpec& rs, const T* zmel,
                                 int64_t zstride, int n, int k0, int k1, T* cz, int64_t czstride) {
  if (P.np == 0) return ECALJ_OK;
  ECALJ_TRY(real_entries<T>(h, P, rs, zmel, zstride, n, k0, k1));
  const int p0 = P.kpole[k0], p1 = P.kpole[k1];
  if (p1 <= p0) return ECALJ_OK;
  const int64_t tot = (int64_t)(p1 - p0) * n;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL((sig_real_add_kernel<T>), dim3(blocks), dim3(256), 0, h->stream, (const int*)(P.d + P.o_pk),
                     (const int*)(P.d + P.o_pc), p0, p1, k0, P.slot, n, cz, czstride);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

template <class T>
static ecalj_status sig_real_t(ecalj_ctx* h, int nb, const T* zmel, int n, int nm, int ntq, const RealSpec& rs,
                               cd* zsec, int64_t ldz) {
  RealPlan P;
  ECALJ_TRY(real_plan(h, rs, nb, n, nm, ntq, P));
  const int64_t zstride = (int64_t)n * nm * ntq;
  if (P.np > 0) ECALJ_TRY(real_entries<T>(h, P, rs, zmel, zstride, n, 0, nb));
  // the final kernel also clips k by k when there are no poles
  if (P.np == 0) {
    std::vector<unsigned char> blob(sizeof(int) * (ntq + 1) + 8, 0);
    ECALJ_TRY(upload_blob(h, ecalj_ctx::S_RTAB, blob, &P.d));
    P.o_ps = P.o_pl = P.o_pk = P.o_pt = 0;
  }
  hipLaunchKernelGGL((sig_real_final_kernel<T>), dim3(ntq * ntq), dim3(256), 0, h->stream,
                     (const int*)(P.d + P.o_ps), (const int*)(P.d + P.o_pl), (const int*)(P.d + P.o_pk),
                     (const int*)(P.d + P.o_pt), P.slot, zmel, zstride, n, nm, ntq, nb, zsec, ldz);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status sig_real(ecalj_ctx* h, int mode, int nb, const void* zmel, int n, int nm, int ntq, const RealSpec& rs,
                      cd* zsec, int64_t ldz) {
  if (mode == M64) return sig_real_t<cd>(h, nb, (const cd*)zmel, n, nm, ntq, rs, zsec, ldz);
  return sig_real_t<cf>(h, nb, (const cf*)zmel, n, nm, ntq, rs, zsec, ldz);
}

// A pole plan kept across the scratch groups of one Sigma_c call: the
// real-axis columns of the k-points [k0, k1) are added into czmelwc (cz of k
// at cz + (k - k0) * n*nm*ntq) before the final product.
struct RealPlanHandle {
  RealPlan P;
};
ecalj_status real_plan_make(ecalj_ctx* h, const RealSpec& rs, int nb, int n, int nm, int ntq, RealPlanHandle** out) {
  auto* r = new RealPlanHandle;
  const ecalj_status st = real_plan(h, rs, nb, n, nm, ntq, r->P);
  if (st != ECALJ_OK) {
    delete r;
    return st;
  }
  *out = r;
  return ECALJ_OK;
}
void real_plan_free(RealPlanHandle* r) { delete r; }
ecalj_status real_plan_into_cz(ecalj_ctx* h, const RealPlanHandle* r, const RealSpec& rs, int mode,
                               const void* zmel, int n, int nm, int ntq, int k0, int k1, void* cz) {
  const int64_t zstride = (int64_t)n * nm * ntq;
  if (mode == M64)
    return real_into_cz<cd>(h, r->P, rs, (const cd*)zmel, zstride, n, k0, k1, (cd*)cz, zstride);
  return real_into_cz<cf>(h, r->P, rs, (const cf*)zmel, zstride, n, k0, k1, (cf*)cz, zstride);
}

}  // namespace ecalj
