// tetwt.hip -- tetrahedron histogram weights of chi0 (gettetwt +
// tetwt5x_dtet4, S/m_tetwt.f90:36-256, S/tetwt5.f90:9-486) for npm = 1 or 2
// and nctot core states, tetrahedron_matrix_linear = F.
//
// Band indices: the valence bands 0..nband-1, then the core states
// nband..nband+nctot-1, flat in k (S/tetwt5.f90:179-185).  jpm = 0 pairs an
// occupied state at k (valence below E_F or core) with an unoccupied one at
// k+q; jpm = 1 (npm = 2, time reversal off) pairs a state at k with an
// occupied one at k+q (valence or core).
//
// Device work, in the reference's two passes:
//   job 0  tw_job0_kernel: one thread per (tetrahedron, band pair) marks the
//          pairs of each corner k and widens their excitation window
//          [demin, demax] (atomic min / max: exact, so the index maps are
//          bit-identical to a sequential loop); tw_sym0_kernel merges the
//          windows of degenerate bands; tw_count_kernel / tw_lists_kernel
//          build the per-k pair lists (ib-major, rsvwwk00_4's order), the
//          bin windows (hisrange) and the whw pointers (block scans).
//   job 1  tw_job1_kernel: one thread per (tetrahedron, pair) integrates
//          the excitation energy over every bin of the region where the
//          occupied band is below and the unoccupied band above E_F
//          (tetcut.hpp: two plane clips into at most nine tetrahedra, each
//          binned through its cumulative volume), and adds the bins to the
//          pair's histogram at the four corners (FP64 atomics);
//          tw_sym1_kernel averages degenerate bands.
// The host keeps what the reference computes once per call outside the
// loops: per-tetrahedron band counts and the degenerate groups (chkdgn).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "ecalj_gpu.h"
#include "tetcut.hpp"

namespace ecalj {
namespace {

struct TwBuf {
  void* p = nullptr;
  ~TwBuf() {
    if (p) (void)hipFree(p);
  }
  ecalj_status alloc(size_t bytes) {
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return ECALJ_ERR_NOMEM;
    return ECALJ_OK;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// flag / window / pair-number arrays: ib over n1 = nband + nctot, jb over
// n2 = nband + ncc (ncc = nctot for npm = 2, else 0), then k, then jpm
struct PairDims {
  int n1, n2, nqbz;
  __device__ size_t operator()(int ib, int jb, int k, int jpm) const {
    return (size_t)ib + (size_t)n1 * ((size_t)jb + (size_t)n2 * ((size_t)k + (size_t)nqbz * jpm));
  }
};

__device__ inline void atomic_min_d(double* a, double v) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(a);
  unsigned long long old = *p;
  while (v < __longlong_as_double((long long)old)) {
    const unsigned long long prev = atomicCAS(p, old, (unsigned long long)__double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}

__device__ inline void atomic_max_d(double* a, double v) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(a);
  unsigned long long old = *p;
  while (v > __longlong_as_double((long long)old)) {
    const unsigned long long prev = atomicCAS(p, old, (unsigned long long)__double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}

// The four corners' histogram windows of one pair: a bin of the
// tetrahedron's histogram goes to every corner whose window holds it, in
// full (times 4 wtet = 1), or with matrix_linear the corner's share lam[c]
// of 4 W (S/tetwt5.f90:329-340).
struct CornerBins {
  int ihw[4], nhw[4], jhw[4];
  double scale;  // pi / (4 vol_BZ)
  double* whw;
  __device__ void operator()(int ihis, double w) const {
    const double v = w * scale;
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (ihis >= ihw[c] && ihis < ihw[c] + nhw[c]) atomicAdd(whw + (jhw[c] - 1) + (ihis - ihw[c]), v);
  }
  __device__ void operator()(int ihis, double w4, const double* lam) const {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (ihis >= ihw[c] && ihis < ihw[c] + nhw[c])
        atomicAdd(whw + (jhw[c] - 1) + (ihis - ihw[c]), w4 * lam[c] * scale);
  }
};

// ---------------------------------------------------------------- kernels
struct TetRow {
  int k[4];            // corner k-points, 0-based
  int w[4];            // corner indices into qbzw, 0-based
  int ni[2], nj[2];    // pair rectangle of each jpm: ibx < ni, jbx < nj
  int nocc1, nocc2;    // valence bands below E_F at k and at k+q (max over corners)
};

// One (tetrahedron, jpm, pair) per thread; the band map and the checks of
// S/tetwt5.f90:252-288.  e1, e2: energies at k and k+q with the core states
// appended (leading dimension n1).
__device__ inline bool tw_pair(const TetRow& t, int jpm, int p, const double* e1, const double* e2, int n1, int nband,
                               double ef, int& ib, int& jb, double eo[4], double eu[4]) {
  const int nj = t.nj[jpm];
  if (p >= t.ni[jpm] * nj) return false;
  const int ibx = p / nj, jbx = p % nj;
  if (jpm == 0) {
    ib = ibx < t.nocc1 ? ibx : ibx - t.nocc1 + nband;
    jb = jbx;
  } else {
    ib = ibx;
    jb = jbx < t.nocc2 ? jbx : jbx - t.nocc2 + nband;
  }
  double mino = 1e300, maxu = -1e300, maxd = -1e300;
  for (int c = 0; c < 4; ++c) {
    const double a = e1[ib + (size_t)n1 * t.k[c]], b = e2[jb + (size_t)n1 * t.k[c]];
    eo[c] = jpm == 0 ? a : b;
    eu[c] = jpm == 0 ? b : a;
    mino = fmin(mino, eo[c]);
    maxu = fmax(maxu, eu[c]);
    maxd = fmax(maxd, eu[c] - eo[c]);
  }
  if (!(mino <= ef && maxu >= ef)) return false;
  return !(maxd < 0.0);
}

// grid (pairs, tetrahedra, npm)
__global__ __launch_bounds__(256) void tw_job0_kernel(const TetRow* __restrict__ tets, const double* __restrict__ e1,
                                                      const double* __restrict__ e2, int nband, PairDims pd,
                                                      double ef, unsigned char* __restrict__ iwgt,
                                                      double* __restrict__ demin, double* __restrict__ demax) {
  const TetRow t = tets[blockIdx.y];
  const int jpm = blockIdx.z;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  int ib, jb;
  double eo[4], eu[4];
  if (!tw_pair(t, jpm, p, e1, e2, pd.n1, nband, ef, ib, jb, eo, eu)) return;
  double dmx = -0.5 * (eo[0] - eu[0]), dmn = dmx;
  for (int c = 1; c < 4; ++c) {
    const double v = -0.5 * (eo[c] - eu[c]);
    dmx = fmax(dmx, v);
    dmn = fmin(dmn, v);
  }
  for (int c = 0; c < 4; ++c) {
    const size_t e = pd(ib, jb, t.k[c], jpm);
    iwgt[e] = 1;
    atomic_max_d(demax + e, dmx);
    atomic_min_d(demin + e, dmn);
  }
}

struct DgBlock {
  int k, jpm, a0, a1, b0, b1;  // 0-based inclusive band ranges at k
};

__global__ void tw_sym0_kernel(const DgBlock* __restrict__ blk, int nblk, PairDims pd, unsigned char* __restrict__ iwgt,
                               double* __restrict__ demin, double* __restrict__ demax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nblk) return;
  const DgBlock b = blk[i];
  unsigned char w = 0;
  double mx = -1e300, mn = 1e300;
  for (int a = b.a0; a <= b.a1; ++a)
    for (int c = b.b0; c <= b.b1; ++c) {
      const size_t e = pd(a, c, b.k, b.jpm);
      w |= iwgt[e];
      mx = fmax(mx, demax[e]);
      mn = fmin(mn, demin[e]);
    }
  for (int a = b.a0; a <= b.a1; ++a)
    for (int c = b.b0; c <= b.b1; ++c) {
      const size_t e = pd(a, c, b.k, b.jpm);
      iwgt[e] = w;
      demax[e] = mx;
      demin[e] = mn;
    }
}

__global__ void tw_fill_kernel(double* __restrict__ a, double v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// block-wide exclusive scan of ints (blockDim.x = 1024 = 16 waves)
__device__ inline int block_exscan(int v, int* total, int* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      const int t = sh[w];
      sh[w] = s;
      s += t;
    }
    sh[16] = s;
  }
  __syncthreads();
  const int r = sh[wave] + x - v;
  *total = sh[16];
  __syncthreads();
  return r;
}

// grid (nqbz, npm): pairs of (k, jpm)
__global__ __launch_bounds__(1024) void tw_count_kernel(const unsigned char* __restrict__ iwgt, PairDims pd,
                                                        int* __restrict__ nbnb) {
  __shared__ int sh[17];
  const int kj = blockIdx.x + pd.nqbz * blockIdx.y;
  int c = 0;
  const size_t nn = (size_t)pd.n1 * pd.n2;
  for (size_t q = threadIdx.x; q < nn; q += blockDim.x) c += iwgt[q + nn * kj];
  int tot;
  block_exscan(c, &tot, sh);
  if (threadIdx.x == 0) nbnb[kj] = tot;
}

// Pair lists of (k, jpm) in ib-major order (rsvwwk00_4), bin windows
// (hisrange), the pair's offset within (k, jpm)'s part of whw, ibjb, and the
// total bin count of (k, jpm).  grid (nqbz, npm).
__global__ __launch_bounds__(1024) void tw_lists_kernel(const unsigned char* __restrict__ iwgt,
                                                        const double* __restrict__ demin,
                                                        const double* __restrict__ demax, PairDims pd, int ld,
                                                        const double* __restrict__ frhis, int nwhis,
                                                        int* __restrict__ n1b, int* __restrict__ n2b,
                                                        int* __restrict__ ihw, int* __restrict__ nhw,
                                                        int* __restrict__ jloc, int* __restrict__ ibjb,
                                                        int* __restrict__ ktot, int* __restrict__ err) {
  __shared__ int sh[17];
  const int k = blockIdx.x, jpm = blockIdx.y, kj = k + pd.nqbz * jpm;
  int base = 0, hbase = 0;
  const int nn = pd.n1 * pd.n2;
  for (int q0 = 0; q0 < nn; q0 += blockDim.x) {
    const int q = q0 + threadIdx.x;
    int f = 0, ih = 0, nh = 0, ib = 0, jb = 0;
    if (q < nn) {
      ib = q / pd.n2;
      jb = q % pd.n2;
      const size_t e = pd(ib, jb, k, jpm);
      f = iwgt[e];
      if (f) {  // the bins [lo, hi] touches (hisrange, S/tetwt5.f90:487-513)
        const int i0 = tet::first_edge_above(frhis, nwhis, demin[e]);
        const int i1 = max(i0, tet::first_edge_above(frhis, nwhis, demax[e]));
        if (i1 > nwhis) {
          *err = 2;
        } else {
          ih = i0;
          nh = i1 - i0 + 1;
        }
      }
    }
    int tf, th;
    const int pos = block_exscan(f, &tf, sh);
    const int hpos = block_exscan(f ? nh : 0, &th, sh);
    if (f) {
      const size_t o = (size_t)(base + pos) + (size_t)ld * kj;
      n1b[o] = ib + 1;
      n2b[o] = jb + 1;
      ihw[o] = ih;
      nhw[o] = nh;
      jloc[o] = hbase + hpos;
      ibjb[pd(ib, jb, k, jpm)] = base + pos + 1;
    }
    base += tf;
    hbase += th;
  }
  if (threadIdx.x == 0) ktot[kj] = hbase;
}

// nqbz here counts the (k, jpm) slots: nqbz * npm
__global__ void tw_jhw_kernel(const int* __restrict__ nbnb, const long long* __restrict__ kbase, int ld, int nqbz,
                              const int* __restrict__ jloc, int* __restrict__ jhw) {
  const size_t tot = (size_t)ld * nqbz;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(e / ld), i = (int)(e % ld);
    if (i < nbnb[k]) jhw[e] = (int)(kbase[k] + jloc[e] + 1);
  }
}

// grid (pairs, tetrahedra, npm)
template <bool Linear>
__global__ __launch_bounds__(256) void tw_job1_kernel(const TetRow* __restrict__ tets, const double* __restrict__ qbzw,
                                                      const double* __restrict__ e1, const double* __restrict__ e2,
                                                      int nband, PairDims pd, double ef,
                                                      const int* __restrict__ ibjb, int ld,
                                                      const int* __restrict__ ihw, const int* __restrict__ nhw,
                                                      const int* __restrict__ jhw, const double* __restrict__ frhis,
                                                      int nwhis, double scale, double* __restrict__ whw,
                                                      int* __restrict__ err) {
  const TetRow t = tets[blockIdx.y];
  const int jpm = blockIdx.z;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  int ib, jb;
  double eo[4], eu[4];
  if (!tw_pair(t, jpm, p, e1, e2, pd.n1, nband, ef, ib, jb, eo, eu)) return;
  CornerBins s;
  s.scale = scale;
  s.whw = whw;
  double kv[4][3];
  for (int c = 0; c < 4; ++c) {
    const size_t o = (size_t)(ibjb[pd(ib, jb, t.k[c], jpm)] - 1) + (size_t)ld * (t.k[c] + (size_t)pd.nqbz * jpm);
    s.ihw[c] = ihw[o];
    s.nhw[c] = nhw[o];
    s.jhw[c] = jhw[o];
    for (int d = 0; d < 3; ++d) kv[c][d] = qbzw[d + 3 * (size_t)t.w[c]];
  }
  if (!tet::pair_histogram<Linear>(kv, eo, eu, ef, frhis, nwhis, s)) *err = 3;  // outside the histogram mesh
}

// degenerate bands: every pair of a block takes the block's average histogram
__global__ void tw_sym1_kernel(const DgBlock* __restrict__ blk, int nblk, PairDims pd, const int* __restrict__ ibjb,
                               int ld, const int* __restrict__ nhw, const int* __restrict__ jhw,
                               double* __restrict__ whw, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nblk) return;
  const DgBlock b = blk[i];
  const size_t kj = (size_t)b.k + (size_t)pd.nqbz * b.jpm;
  const int bs = (b.a1 - b.a0 + 1) * (b.b1 - b.b0 + 1);
  int isum = 0;
  for (int a = b.a0; a <= b.a1; ++a)
    for (int c = b.b0; c <= b.b1; ++c) isum += ibjb[pd(a, c, b.k, b.jpm)] != 0;
  if (isum == 0) return;
  if (isum != bs) {
    *err = 5;
    return;
  }
  const size_t o0 = (size_t)(ibjb[pd(b.a0, b.b0, b.k, b.jpm)] - 1) + (size_t)ld * kj;
  const int nh = nhw[o0];
  for (int a = b.a0; a <= b.a1; ++a)  // one window for the block (its windows were merged in job 0)
    for (int c = b.b0; c <= b.b1; ++c)
      if (nhw[(size_t)(ibjb[pd(a, c, b.k, b.jpm)] - 1) + (size_t)ld * kj] != nh) {
        *err = 5;
        return;
      }
  for (int h = 0; h < nh; ++h) {
    double sum = 0.0;
    for (int a = b.a0; a <= b.a1; ++a)
      for (int c = b.b0; c <= b.b1; ++c) {
        const size_t o = (size_t)(ibjb[pd(a, c, b.k, b.jpm)] - 1) + (size_t)ld * kj;
        sum += whw[jhw[o] - 1 + h];
      }
    const double avg = sum / bs;
    for (int a = b.a0; a <= b.a1; ++a)
      for (int c = b.b0; c <= b.b1; ++c) {
        const size_t o = (size_t)(ibjb[pd(a, c, b.k, b.jpm)] - 1) + (size_t)ld * kj;
        whw[jhw[o] - 1 + h] = avg;
      }
  }
}

// chkdgn (S/tetwt5.f90:868-912): 0-based inclusive ranges
int chkdgn(const double* e, int n, std::vector<int>& lo, std::vector<int>& hi) {
  lo.assign(n, 0);
  hi.assign(n, 0);
  if (n < 1) return 0;
  lo[0] = 0;
  if (n == 1) {
    hi[0] = 0;
    return 1;
  }
  int i = 0;
  for (int ix = 1; ix < n; ++ix) {
    if (std::fabs(e[ix] - e[ix - 1]) > 1e-4) {
      hi[i] = ix - 1;
      ++i;
      lo[i] = hi[i - 1] + 1;
      if (ix == n - 1) hi[i] = ix;
    } else if (ix == n - 1) {
      hi[i] = n - 1;
    }
  }
  return i + 1;
}

}  // namespace
}  // namespace ecalj

using namespace ecalj;

extern "C" ecalj_status ecalj_tetwt_full(ecalj_handle h, int nband, int nqbz, int ntetf, const int* idtetf,
                                         int nqbzw, const double* qbzw, const int* ib1bz, const double* ekxx1,
                                         const double* ekxx2, double ef, double ebmx, const double* qlat,
                                         const double* frhis, int nwhis, int npm, int nctot, const double* ecore,
                                         int matrix_linear, int nbnbx_cap, int64_t nhw_cap, int* nbnb, int* n1b, int* n2b, int* ihw,
                                         int* nhw, int* jhw, double* whw, int* nbnbx_out, int64_t* nhwtot_out) {
  if (!h || !idtetf || !qbzw || !ib1bz || !ekxx1 || !ekxx2 || !qlat || !frhis || !nbnbx_out || !nhwtot_out)
    return ECALJ_ERR_ARG;
  if (nband < 1 || nqbz < 1 || ntetf < 1 || nqbzw < 1 || nwhis < 1) return ECALJ_ERR_DIM;
  if (npm < 1 || npm > 2 || nctot < 0) return ECALJ_ERR_DIM;
  if (nctot > 0 && !ecore) return ECALJ_ERR_ARG;
  if (std::fabs(frhis[0]) > 1e-12) return ECALJ_ERR_ARG;  // S/m_tetwt.f90:176
  if (h->capturing) return ECALJ_ERR_STATE;  // synchronises with the host between its passes
  ECALJ_HIP(hipSetDevice(h->device));
  hipStream_t st = h->stream;
  const int ncc = npm == 2 ? nctot : 0;
  const PairDims pd{nband + nctot, nband + ncc, nqbz};
  // per-tetrahedron rows: corners and band counts (S/tetwt5.f90:223-264)
  std::vector<TetRow> rows(ntetf);
  int maxp = 0;
  const double ecut = std::min(1e10, ebmx);
  for (int t = 0; t < ntetf; ++t) {
    TetRow& r = rows[t];
    int nocc1 = 0, nocc2 = 0, nbmx1 = nband, nbmx2 = nband;
    for (int c = 0; c < 4; ++c) {
      const int w = idtetf[c + 4 * (size_t)t] - 1;
      if (w < 0 || w >= nqbzw) return ECALJ_ERR_ARG;
      r.w[c] = w;
      r.k[c] = ib1bz[w] - 1;
      if (r.k[c] < 0 || r.k[c] >= nqbz) return ECALJ_ERR_ARG;
      int o1 = 0, o2 = 0, m1 = 0, m2 = 0;
      for (int b = 0; b < nband; ++b) {
        const double a = ekxx1[b + (size_t)nband * r.k[c]], e = ekxx2[b + (size_t)nband * r.k[c]];
        o1 += a < ef;
        o2 += e < ef;
        m1 += a < ecut;
        m2 += e < ecut;
      }
      nocc1 = std::max(nocc1, o1);
      nocc2 = std::max(nocc2, o2);
      nbmx1 = std::min(nbmx1, m1);
      nbmx2 = std::min(nbmx2, m2);
    }
    r.nocc1 = nocc1;
    r.nocc2 = nocc2;
    r.ni[0] = nocc1 + nctot;
    r.nj[0] = nbmx2;
    r.ni[1] = nbmx1;
    r.nj[1] = nocc2 + nctot;
    for (int j = 0; j < npm; ++j) maxp = std::max(maxp, r.ni[j] * r.nj[j]);
  }
  // degenerate blocks with more than one pair (singletons need no merging):
  // valence groups of each side, plus the core groups on the occupied side
  std::vector<DgBlock> blocks;
  {
    std::vector<int> l1, h1, l2, h2, lc, hc;
    const int nrc = nctot > 0 ? chkdgn(ecore, nctot, lc, hc) : 0;
    for (int c = 0; c < nrc; ++c) {
      lc[c] += nband;
      hc[c] += nband;
    }
    for (int k = 0; k < nqbz; ++k) {
      const int r1 = chkdgn(ekxx1 + (size_t)nband * k, nband, l1, h1);
      const int r2 = chkdgn(ekxx2 + (size_t)nband * k, nband, l2, h2);
      for (int jpm = 0; jpm < npm; ++jpm) {
        const int g1 = r1 + (jpm == 0 ? nrc : 0), g2 = r2 + (jpm == 1 ? nrc : 0);
        for (int a = 0; a < g1; ++a)
          for (int b = 0; b < g2; ++b) {
            const int a0 = a < r1 ? l1[a] : lc[a - r1], a1 = a < r1 ? h1[a] : hc[a - r1];
            const int b0 = b < r2 ? l2[b] : lc[b - r2], b1 = b < r2 ? h2[b] : hc[b - r2];
            if (a1 > a0 || b1 > b0) blocks.push_back({k, jpm, a0, a1, b0, b1});
          }
      }
    }
  }
  // energies with the core states appended at every k
  const size_t ne = (size_t)pd.n1 * nqbz;
  std::vector<double> e1(ne), e2(ne);
  for (int k = 0; k < nqbz; ++k) {
    std::memcpy(&e1[(size_t)pd.n1 * k], ekxx1 + (size_t)nband * k, sizeof(double) * nband);
    std::memcpy(&e2[(size_t)pd.n1 * k], ekxx2 + (size_t)nband * k, sizeof(double) * nband);
    for (int c = 0; c < nctot; ++c) e1[(size_t)pd.n1 * k + nband + c] = e2[(size_t)pd.n1 * k + nband + c] = ecore[c];
  }
  const int nkj = nqbz * npm;  // (k, jpm) slots, k fastest
  const size_t nbb = (size_t)pd.n1 * pd.n2 * nkj;
  TwBuf d_rows, d_ek1, d_ek2, d_qw, d_fr, d_iw, d_dmn, d_dmx, d_blk, d_nbnb, d_err, d_ibjb;
  ECALJ_TRY(d_rows.alloc(sizeof(TetRow) * ntetf));
  ECALJ_TRY(d_ek1.alloc(sizeof(double) * ne));
  ECALJ_TRY(d_ek2.alloc(sizeof(double) * ne));
  ECALJ_TRY(d_qw.alloc(sizeof(double) * 3 * (size_t)nqbzw));
  ECALJ_TRY(d_fr.alloc(sizeof(double) * (nwhis + 1)));
  ECALJ_TRY(d_iw.alloc(nbb));
  ECALJ_TRY(d_dmn.alloc(sizeof(double) * nbb));
  ECALJ_TRY(d_dmx.alloc(sizeof(double) * nbb));
  ECALJ_TRY(d_ibjb.alloc(sizeof(int) * nbb));
  ECALJ_TRY(d_blk.alloc(sizeof(DgBlock) * std::max<size_t>(blocks.size(), 1)));
  ECALJ_TRY(d_nbnb.alloc(sizeof(int) * nkj));
  ECALJ_TRY(d_err.alloc(sizeof(int)));
  ECALJ_HIP(hipMemcpyAsync(d_rows.p, rows.data(), sizeof(TetRow) * ntetf, hipMemcpyHostToDevice, st));
  ECALJ_HIP(hipMemcpyAsync(d_ek1.p, e1.data(), sizeof(double) * ne, hipMemcpyHostToDevice, st));
  ECALJ_HIP(hipMemcpyAsync(d_ek2.p, e2.data(), sizeof(double) * ne, hipMemcpyHostToDevice, st));
  ECALJ_HIP(hipMemcpyAsync(d_qw.p, qbzw, sizeof(double) * 3 * (size_t)nqbzw, hipMemcpyHostToDevice, st));
  ECALJ_HIP(hipMemcpyAsync(d_fr.p, frhis, sizeof(double) * (nwhis + 1), hipMemcpyHostToDevice, st));
  if (!blocks.empty())
    ECALJ_HIP(hipMemcpyAsync(d_blk.p, blocks.data(), sizeof(DgBlock) * blocks.size(), hipMemcpyHostToDevice, st));
  ECALJ_HIP(hipMemsetAsync(d_iw.p, 0, nbb, st));
  ECALJ_HIP(hipMemsetAsync(d_ibjb.p, 0, sizeof(int) * nbb, st));
  ECALJ_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), st));
  const unsigned fb = (unsigned)std::min<size_t>((nbb + 255) / 256, 65536);
  hipLaunchKernelGGL(tw_fill_kernel, dim3(fb), dim3(256), 0, st, d_dmn.as<double>(), 1e10, nbb);
  hipLaunchKernelGGL(tw_fill_kernel, dim3(fb), dim3(256), 0, st, d_dmx.as<double>(), -1e10, nbb);
  const int nblk = (int)blocks.size();
  // job 0
  if (maxp > 0) {
    hipLaunchKernelGGL(tw_job0_kernel, dim3((maxp + 255) / 256, ntetf, npm), dim3(256), 0, st, d_rows.as<TetRow>(),
                       d_ek1.as<double>(), d_ek2.as<double>(), nband, pd, ef, d_iw.as<unsigned char>(),
                       d_dmn.as<double>(), d_dmx.as<double>());
    ECALJ_HIP(hipGetLastError());
  }
  if (nblk)
    hipLaunchKernelGGL(tw_sym0_kernel, dim3((nblk + 255) / 256), dim3(256), 0, st, d_blk.as<DgBlock>(), nblk, pd,
                       d_iw.as<unsigned char>(), d_dmn.as<double>(), d_dmx.as<double>());
  hipLaunchKernelGGL(tw_count_kernel, dim3(nqbz, npm), dim3(1024), 0, st, d_iw.as<unsigned char>(), pd,
                     d_nbnb.as<int>());
  ECALJ_HIP(hipGetLastError());
  std::vector<int> hb(nkj);
  ECALJ_HIP(hipMemcpyAsync(hb.data(), d_nbnb.p, sizeof(int) * nkj, hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipStreamSynchronize(st));
  int nbnbx = 1;
  for (int kj = 0; kj < nkj; ++kj) nbnbx = std::max(nbnbx, hb[kj]);
  *nbnbx_out = nbnbx;
  const int ld = std::max(nbnbx, nbnbx_cap);
  TwBuf d_n1b, d_n2b, d_ihw, d_nhw, d_jloc, d_jhw, d_ktot, d_kbase;
  const size_t nl = (size_t)ld * nkj;
  for (TwBuf* b : {&d_n1b, &d_n2b, &d_ihw, &d_nhw, &d_jloc, &d_jhw}) ECALJ_TRY(b->alloc(sizeof(int) * nl));
  ECALJ_TRY(d_ktot.alloc(sizeof(int) * nkj));
  ECALJ_TRY(d_kbase.alloc(sizeof(long long) * nkj));
  for (TwBuf* b : {&d_n1b, &d_n2b, &d_ihw, &d_nhw, &d_jhw}) ECALJ_HIP(hipMemsetAsync(b->p, 0, sizeof(int) * nl, st));
  hipLaunchKernelGGL(tw_lists_kernel, dim3(nqbz, npm), dim3(1024), 0, st, d_iw.as<unsigned char>(),
                     d_dmn.as<double>(), d_dmx.as<double>(), pd, ld, d_fr.as<double>(), nwhis, d_n1b.as<int>(),
                     d_n2b.as<int>(), d_ihw.as<int>(), d_nhw.as<int>(), d_jloc.as<int>(), d_ibjb.as<int>(),
                     d_ktot.as<int>(), d_err.as<int>());
  ECALJ_HIP(hipGetLastError());
  std::vector<int> kt(nkj);
  int herr = 0;
  ECALJ_HIP(hipMemcpyAsync(kt.data(), d_ktot.p, sizeof(int) * nkj, hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipMemcpyAsync(&herr, d_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipStreamSynchronize(st));
  if (herr) return ECALJ_ERR_DIM;  // a window outside the histogram mesh (hisrange)
  // whw pointers run over jpm, then k, then the pairs (S/m_tetwt.f90:199-216)
  std::vector<long long> kb(nkj);
  long long tot = 0;
  for (int kj = 0; kj < nkj; ++kj) {
    kb[kj] = tot;
    tot += kt[kj];
  }
  *nhwtot_out = tot;
  if (nbnbx > nbnbx_cap || tot > nhw_cap) return ECALJ_ERR_DIM;  // sizes reported; call again
  if (!nbnb || !n1b || !n2b || !ihw || !nhw || !jhw || !whw) return ECALJ_ERR_ARG;
  ECALJ_HIP(hipMemcpyAsync(d_kbase.p, kb.data(), sizeof(long long) * nkj, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(tw_jhw_kernel, dim3((unsigned)std::min<size_t>((nl + 255) / 256, 65536)), dim3(256), 0, st,
                     d_nbnb.as<int>(), d_kbase.as<long long>(), ld, nkj, d_jloc.as<int>(), d_jhw.as<int>());
  ECALJ_HIP(hipGetLastError());
  // job 1
  TwBuf d_whw;
  ECALJ_TRY(d_whw.alloc(sizeof(double) * std::max<long long>(tot, 1)));
  ECALJ_HIP(hipMemsetAsync(d_whw.p, 0, sizeof(double) * std::max<long long>(tot, 1), st));
  double qb[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) qb[c][r] = qlat[r + 3 * c];
  const double voltot = std::fabs(qb[0][0] * qb[1][1] * qb[2][2] - qb[0][0] * qb[1][2] * qb[2][1] -
                                  qb[0][1] * qb[1][0] * qb[2][2] + qb[0][1] * qb[1][2] * qb[2][0] +
                                  qb[0][2] * qb[1][0] * qb[2][1] - qb[0][2] * qb[1][1] * qb[2][0]);
  const double pi = 3.14159265358979323846;
  const double scale = pi / 4.0 / voltot;  // piofvoltot (S/tetwt5.f90:335-338)
  if (maxp > 0) {
    hipLaunchKernelGGL(matrix_linear ? tw_job1_kernel<true> : tw_job1_kernel<false>,
                       dim3((maxp + 255) / 256, ntetf, npm), dim3(256), 0, st, d_rows.as<TetRow>(),
                       d_qw.as<double>(), d_ek1.as<double>(), d_ek2.as<double>(), nband, pd, ef, d_ibjb.as<int>(),
                       ld, d_ihw.as<int>(), d_nhw.as<int>(), d_jhw.as<int>(), d_fr.as<double>(), nwhis, scale,
                       d_whw.as<double>(), d_err.as<int>());
    ECALJ_HIP(hipGetLastError());
  }
  if (nblk)
    hipLaunchKernelGGL(tw_sym1_kernel, dim3((nblk + 255) / 256), dim3(256), 0, st, d_blk.as<DgBlock>(), nblk, pd,
                       d_ibjb.as<int>(), ld, d_nhw.as<int>(), d_jhw.as<int>(), d_whw.as<double>(), d_err.as<int>());
  ECALJ_HIP(hipGetLastError());
  ECALJ_HIP(hipMemcpyAsync(nbnb, d_nbnb.p, sizeof(int) * nkj, hipMemcpyDeviceToHost, st));
  // outputs have leading dimension nbnbx_cap (= ld), then k, then jpm
  ECALJ_HIP(hipMemcpyAsync(n1b, d_n1b.p, sizeof(int) * nl, hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipMemcpyAsync(n2b, d_n2b.p, sizeof(int) * nl, hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipMemcpyAsync(ihw, d_ihw.p, sizeof(int) * nl, hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipMemcpyAsync(nhw, d_nhw.p, sizeof(int) * nl, hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipMemcpyAsync(jhw, d_jhw.p, sizeof(int) * nl, hipMemcpyDeviceToHost, st));
  if (tot > 0) ECALJ_HIP(hipMemcpyAsync(whw, d_whw.p, sizeof(double) * tot, hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipMemcpyAsync(&herr, d_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipStreamSynchronize(st));
  if (herr) return ECALJ_ERR_DIM;
  return ECALJ_OK;
}

extern "C" ecalj_status ecalj_tetwt(ecalj_handle h, int nband, int nqbz, int ntetf, const int* idtetf, int nqbzw,
                                    const double* qbzw, const int* ib1bz, const double* ekxx1, const double* ekxx2,
                                    double ef, double ebmx, const double* qlat, const double* frhis, int nwhis,
                                    int nbnbx_cap, int64_t nhw_cap, int* nbnb, int* n1b, int* n2b, int* ihw,
                                    int* nhw, int* jhw, double* whw, int* nbnbx_out, int64_t* nhwtot_out) {
  return ecalj_tetwt_full(h, nband, nqbz, ntetf, idtetf, nqbzw, qbzw, ib1bz, ekxx1, ekxx2, ef, ebmx, qlat, frhis,
                          nwhis, 1, 0, nullptr, 0, nbnbx_cap, nhw_cap, nbnb, n1b, n2b, ihw, nhw, jhw, whw, nbnbx_out,
                          nhwtot_out);
}
