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

template <class T>
__global__ __launch_bounds__(256) void sig_real_entries_kernel(const RealItem* __restrict__ items,
                                                               const int* __restrict__ perm,
                                                               const int* __restrict__ ecol,
                                                               const double* __restrict__ ew,
                                                               const T* __restrict__ zmel, int64_t zstride,
                                                               const cd* __restrict__ wr, int64_t ldw,
                                                               int64_t wstride, int n, cd* __restrict__ slot) {
  __shared__ cd wt[RT][RJ + 1];  // Wt(I0+r, J0+jj)
  __shared__ cd zb[RJ][RE];      // w_e * zmel(J0+jj, col_e)
  const RealItem it = items[blockIdx.x];
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
        double xr, xi;
        ld_c2(Z + (int64_t)n * ecol[id] + J, xr, xi);
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

// zsec(m, m') += sum over poles p of external state m (fixed order) of
// sum_I conj(v_p(I)) zmel_k(I, it_p, m'), v_p = slot[3p] + slot[3p+1] + slot[3p+2];
// then Im(zsec_mm) <= 0.  One block per (m, m').
template <class T>
__global__ __launch_bounds__(256) void sig_real_final_kernel(const int* __restrict__ pstart,
                                                             const int* __restrict__ plist,
                                                             const int* __restrict__ pk,
                                                             const int* __restrict__ pit, const cd* __restrict__ slot,
                                                             const T* __restrict__ zmel, int64_t zstride, int n,
                                                             int nm, int ntq, cd* __restrict__ zsec, int64_t ldz) {
  __shared__ double sr[256], si[256];
  const int m = blockIdx.x % ntq, mp = blockIdx.x / ntq;
  double tr = 0.0, ti = 0.0;
  for (int q = pstart[m]; q < pstart[m + 1]; ++q) {
    const int p = plist[q];
    const cd* s0 = slot + (int64_t)n * 3 * p;
    const T* Z = zmel + zstride * pk[p] + (int64_t)n * (pit[p] + (int64_t)nm * mp);
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
  if (threadIdx.x == 0) {
    cd* z = zsec + m + ldz * mp;
    const cd o = *z;
    double zr = o.re + tr, zi = o.im + ti;
    if (m == mp && zi > 0.0) zi = 0.0;
    *z = cd{zr, zi};
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

template <class T>
static ecalj_status sig_real_t(ecalj_ctx* h, int nb, const T* zmel, int n, int nm, int ntq, const cd* wr,
                               int64_t ldw, int64_t wstride, int nw, const double* freq_r, const double* omega,
                               const double* ekc, double ef, double esmr, int nctot, int ns2r, const double* wkkr,
                               cd* zsec, int64_t ldz) {
  const double ddw = 10.0, wfaccut = 1e-8;  // S/m_sxcf_gemm.f90:264,408
  // poles p (k, it, itp) with three entries 3p+j at iw = ixs-1+j
  std::vector<int> pk, pit, pitp, eiw, ecol;
  std::vector<double> ew;
  for (int k = 0; k < nb; ++k) {
    const double* e = ekc + (int64_t)nm * k;
    int nt0p = 0, nt0m = 0;
    for (int i = 0; i < nm; ++i) {
      nt0p += e[i] < ef + ddw * esmr;
      nt0m += e[i] < ef - ddw * esmr;
    }
    for (int itp = 0; itp < ntq; ++itp) {
      const double omg = omega[itp];
      const int itini = (omg >= ef) ? std::max(1, nt0m + 1) : 1;
      const int itend = (omg >= ef) ? ns2r : std::min(nt0p, ns2r);
      for (int it = itini; it <= itend; ++it) {
        double wf = wfacx2_h(omg, ef, e[it - 1], it <= nctot ? 0.0 : esmr);
        if (wf < wfaccut) continue;
        wf *= wkkr[k] * ((omg - ef) >= 0.0 ? 1.0 : -1.0);
        const double we = 0.5 * std::fabs(omg - weavx2_h(omg, ef, e[it - 1], esmr));
        int ixs = 0;
        for (int i = 1; i <= nw; ++i)
          if (freq_r[i] > we) {
            ixs = i;
            break;
          }
        if (ixs < 1 || ixs + 1 > nw) return ECALJ_ERR_DIM;  // pole outside the real mesh
        double a[3][3], ai[3][3];
        for (int i = 0; i < 3; ++i) {
          const double x = freq_r[ixs - 1 + i];
          a[i][0] = 1.0;
          a[i][1] = x * x;
          a[i][2] = x * x * x * x;
        }
        inv3(a, ai);
        const double xv[3] = {1.0, we * we, we * we * we * we};
        pk.push_back(k);
        pit.push_back(it - 1);
        pitp.push_back(itp);
        for (int j = 0; j < 3; ++j) {
          eiw.push_back(ixs - 1 + j);
          ecol.push_back((it - 1) + nm * itp);
          ew.push_back(wf * (xv[0] * ai[0][j] + xv[1] * ai[1][j] + xv[2] * ai[2][j]));
        }
      }
    }
  }
  const int np = (int)pk.size();
  if (np == 0) return ECALJ_OK;
  const int ne = 3 * np;
  // entries ordered by (k, iw); items of <= RE entries x row tiles
  std::vector<int> perm(ne);
  for (int e = 0; e < ne; ++e) perm[e] = e;
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) {
    const int kx = pk[x / 3], ky = pk[y / 3];
    return kx != ky ? kx < ky : eiw[x] < eiw[y];
  });
  std::vector<RealItem> items;
  for (int s = 0; s < ne;) {
    const int k = pk[perm[s] / 3], iw = eiw[perm[s]];
    int e = s;
    while (e < ne && pk[perm[e] / 3] == k && eiw[perm[e]] == iw && e - s < RE) ++e;
    for (int i0 = 0; i0 < n; i0 += RT) items.push_back({k, iw, s, e - s, i0, 0});
    s = e;
  }
  // poles grouped by external state, in (k, it) order
  std::vector<int> pstart(ntq + 1, 0), plist(np);
  for (int p = 0; p < np; ++p) pstart[pitp[p] + 1]++;
  for (int m = 0; m < ntq; ++m) pstart[m + 1] += pstart[m];
  {
    std::vector<int> fill(pstart.begin(), pstart.end() - 1);
    for (int p = 0; p < np; ++p) plist[fill[pitp[p]]++] = p;
  }
  // one table upload: items | perm | ecol | pstart | plist | pk | pit | ew (8-byte aligned)
  std::vector<unsigned char> blob;
  auto put = [&](const void* src, size_t bytes) {
    const size_t o = blob.size();
    blob.resize(o + ((bytes + 7) / 8) * 8);
    if (bytes) memcpy(blob.data() + o, src, bytes);
    return o;
  };
  const size_t o_it = put(items.data(), sizeof(RealItem) * items.size());
  const size_t o_pm = put(perm.data(), sizeof(int) * ne);
  const size_t o_ec = put(ecol.data(), sizeof(int) * ne);
  const size_t o_ps = put(pstart.data(), sizeof(int) * (ntq + 1));
  const size_t o_pl = put(plist.data(), sizeof(int) * np);
  const size_t o_pk = put(pk.data(), sizeof(int) * np);
  const size_t o_pt = put(pit.data(), sizeof(int) * np);
  const size_t o_ew = put(ew.data(), sizeof(double) * ne);
  unsigned char* d;
  ECALJ_TRY(upload_blob(h, ecalj_ctx::S_RTAB, blob, &d));
  cd* slot;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_RSLOT, sizeof(cd) * (size_t)n * ne, (void**)&slot));
  const int64_t zstride = (int64_t)n * nm * ntq;
  if (items.size() > 0x7fffffff) return ECALJ_ERR_DIM;
  hipLaunchKernelGGL((sig_real_entries_kernel<T>), dim3((unsigned)items.size()), dim3(256), 0, h->stream,
                     (const RealItem*)(d + o_it), (const int*)(d + o_pm), (const int*)(d + o_ec),
                     (const double*)(d + o_ew), zmel, zstride, wr, ldw, wstride, n, slot);
  ECALJ_HIP(hipGetLastError());
  hipLaunchKernelGGL((sig_real_final_kernel<T>), dim3(ntq * ntq), dim3(256), 0, h->stream, (const int*)(d + o_ps),
                     (const int*)(d + o_pl), (const int*)(d + o_pk), (const int*)(d + o_pt), slot, zmel, zstride, n,
                     nm, ntq, zsec, ldz);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status sig_real(ecalj_ctx* h, int mode, int nb, const void* zmel, int n, int nm, int ntq, const cd* wr,
                      int64_t ldw, int64_t wstride, int nw, const double* freq_r, const double* omega,
                      const double* ekc, double ef, double esmr, int nctot, int ns2r, const double* wkkr, cd* zsec,
                      int64_t ldz) {
  if (mode == M64)
    return sig_real_t<cd>(h, nb, (const cd*)zmel, n, nm, ntq, wr, ldw, wstride, nw, freq_r, omega, ekc, ef, esmr,
                          nctot, ns2r, wkkr, zsec, ldz);
  return sig_real_t<cf>(h, nb, (const cf*)zmel, n, nm, ntq, wr, ldw, wstride, nw, freq_r, omega, ekc, ef, esmr,
                        nctot, ns2r, wkkr, zsec, ldz);
}

}  // namespace ecalj
