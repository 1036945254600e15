// sigbf3.hip -- Sigma_c frequency contraction in the bf16x3 form.
//
//   czmelwc(I,c) = sum_v wgt(v,c) * sum_J Wt_v(I,J) zmel(J,c),
//   Wt_v(I,J) = (I <= J) ? W_v(I,J) : conj(W_v(J,I))   (keepwv upper-triangle
//   convention, S/m_sxcf_gemm.f90:468-475, after the conjugate transpose of :492)
//
// The weight depends only on (v, c), so it is applied to the per-frequency
// partial sums in FP32 instead of to every B element.  W_v and zmel are split
// once into bf16 hi/lo planes (plane p: 0 = re hi, 1 = re lo, 2 = im hi,
// 3 = im lo) and written pre-tiled: for each 32-wide k slice ks the operand
// rows of one plane are contiguous, 32 bf16 (64 bytes) per row, with the four
// 16-byte chunks of row r stored in the order chunk ^ (((r >> 1) ^ (r >> 2)) & 3).  A block
// tile of one stage is then a few contiguous 4-8 KB runs that the main loop
// moves global -> LDS with direct-to-LDS loads (lane i of a wave writes LDS
// bytes 16i..16i+15 of a 1 KB run), three stages in flight, and the chunk
// order makes the MFMA operand reads (16 rows x one chunk) bank-conflict free
// whether the LDS serves a 16-byte read 8 lanes (128 bytes) or 16 lanes (256
// bytes) at a time: rows of equal parity within 8 and rows equal mod 4 within
// 16 land in distinct 16-byte bank groups.
#include <algorithm>
#include <numeric>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sigtile.hpp"

namespace ecalj {


// ---------------------------------------------------------------- pre-split
// A[b][(((v*NS + ks)*4 + p)*MP + I)*32 + slot] from W (complex128, ldw, per-batch stride w_stride)
constexpr int SPT = 32;
constexpr int SPLIT_KS = 4;  // k slices per block of split_w_kernel / split_w8_kernel
__global__ __launch_bounds__(256) void split_w_kernel(const cd* __restrict__ w, int64_t ldw, int64_t w_stride,
                                                      int n, int nw1, int MP, int NS,
                                                      unsigned short* __restrict__ as, int64_t as_stride) {
  __shared__ cd up[SPT][SPT + 1];  // up[jj][ii] = W(I0+ii, J0+jj)
  __shared__ cd lo[SPT][SPT + 1];  // lo[ii][jj] = W(J0+jj, I0+ii)
  const int I0 = blockIdx.y * SPT;
  const int v = blockIdx.z % nw1, b = blockIdx.z / nw1;
  const cd* W = w + w_stride * b + (int64_t)ldw * n * v;
  const int tx = threadIdx.x % SPT, ty = threadIdx.x / SPT;  // 32 x 8
  typedef unsigned short us4v __attribute__((ext_vector_type(4)));
  // SPLIT_KS consecutive k slices per block
  for (int kq = 0; kq < SPLIT_KS; ++kq) {
    const int ks = blockIdx.x * SPLIT_KS + kq;
    if (ks >= NS) break;
    const int J0 = ks * SPT;
    if (kq) __syncthreads();  // the previous slice's LDS reads are done
    // a tile wholly above the diagonal reads only W(I,J), wholly below only W(J,I)
    const bool need_up = I0 <= J0 + SPT - 1, need_lo = I0 + SPT - 1 > J0;
    for (int r = ty; r < SPT; r += 8) {
      if (need_up) {
        const int I = I0 + tx, J = J0 + r;
        up[r][tx] = (I < n && J < n) ? W[I + ldw * J] : cd{0.0, 0.0};
      }
      if (need_lo) {
        const int J2 = J0 + tx, I2 = I0 + r;
        lo[r][tx] = (I2 < n && J2 < n) ? W[J2 + ldw * I2] : cd{0.0, 0.0};
      }
    }
    __syncthreads();
    // every thread: row r, 4 k values (half hf of chunk ch)
    const int r = threadIdx.x >> 3, ch = (threadIdx.x >> 1) & 3, hf = threadIdx.x & 1;
    const int I = I0 + r;
    us4v rh, rl, ih, il;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int jj = 8 * ch + 4 * hf + e, J = J0 + jj;
      double xr = 0.0, xi = 0.0;
      if (I < n && J < n) {
        if (I <= J) {
          xr = up[jj][r].re;
          xi = up[jj][r].im;
        } else {
          xr = lo[r][jj].re;
          xi = -lo[r][jj].im;
        }
      }
      unsigned short h, l;
      split_bf16((float)xr, h, l);
      rh[e] = h;
      rl[e] = l;
      split_bf16((float)xi, h, l);
      ih[e] = h;
      il[e] = l;
    }
    const int64_t pst = (int64_t)MP * QBK;
    unsigned short* A = as + as_stride * b + ((int64_t)v * NS + ks) * 4 * pst + (int64_t)I * QBK +
                        8 * chunk_slot(I, ch) + 4 * hf;
    *reinterpret_cast<us4v*>(A) = rh;
    *reinterpret_cast<us4v*>(A + pst) = rl;
    *reinterpret_cast<us4v*>(A + 2 * pst) = ih;
    *reinterpret_cast<us4v*>(A + 3 * pst) = il;
  }
}

// B[b][((ks*4 + p)*NCP + c)*32 + slot] from zmel (complex64, n x ncol per batch);
// one thread per (b, ks, c, chunk)
__global__ __launch_bounds__(256) void split_z_kernel(const cf* __restrict__ z, int64_t z_stride, int n,
                                                      int ncol, int NCP, int NS, unsigned short* __restrict__ bs,
                                                      int64_t bs_stride, int nb) {
  typedef unsigned short us8v __attribute__((ext_vector_type(8)));
  const int64_t per = (int64_t)NS * NCP * 4, tot = per * nb;
  const int64_t pst = (int64_t)NCP * QBK;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e % per;
    const int ks = (int)(r / (4 * NCP));
    const int c = (int)((r / 4) % NCP), ch = (int)(r % 4);
    us8v rh, rl, ih, il;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int J = ks * QBK + 8 * ch + q;
      float xr = 0.f, xi = 0.f;
      if (c < ncol && J < n) {
        const cf x = z[z_stride * b + J + (int64_t)n * c];
        xr = x.re;
        xi = x.im;
      }
      unsigned short h, l;
      split_bf16(xr, h, l);
      rh[q] = h;
      rl[q] = l;
      split_bf16(xi, h, l);
      ih[q] = h;
      il[q] = l;
    }
    unsigned short* B = bs + bs_stride * b + (int64_t)ks * 4 * pst + (int64_t)c * QBK + 8 * chunk_slot(c, ch);
    *reinterpret_cast<us8v*>(B) = rh;
    *reinterpret_cast<us8v*>(B + pst) = rl;
    *reinterpret_cast<us8v*>(B + 2 * pst) = ih;
    *reinterpret_cast<us8v*>(B + 3 * pst) = il;
  }
}

// Uniform batch.  G == 0: block (x, y, z) = (row tile, column tile, internal k).
// G > 0, XCD-aware: a 1-D grid whose blocks the dispatcher deals round-robin
// to the 8 XCDs (block p on XCD p % 8, in increasing p).  XCD x's j-th block
// belongs to super-tile (j / S)*8 + x of S = R*G blocks: all R row tiles times
// G consecutive column tiles of one k.  The S blocks of a super-tile run
// together on one XCD, read each W stage of their k together (L2 hits) and
// keep their G zmel panels in that XCD's L2 across the frequency loop; the 8
// XCDs work on consecutive super-tiles, mostly of the same k.  Blocks past
// the last tile exit at once.
template <int BN, int NSTG, int NW>
__global__ __launch_bounds__(NW * 64, 1) void sig_bf3_kernel(const unsigned short* __restrict__ as,
                                                              int64_t as_stride,
                                                              const unsigned short* __restrict__ bs,
                                                              int64_t bs_stride, const double* __restrict__ wgt,
                                                              int64_t g_stride, int n, int nw1, int v0, int nwg,
                                                              int ncol, int MP, int NS, int NCP, int accumulate,
                                                              cf* __restrict__ cz, int64_t cz_stride, int G,
                                                              int nk) {
  __shared__ SigSmem<BN, NSTG> s;
  int b, r, c;
  if (G == 0) {
    r = blockIdx.x;
    c = blockIdx.y;
    b = blockIdx.z;
  } else {
    const int R = MP / QBM, Ct = NCP / BN, NG = (Ct + G - 1) / G, S = R * G;
    const int p = blockIdx.x, x = p & 7, j = p >> 3;
    const int sigma = (j / S) * 8 + x, loc = j % S;
    b = sigma / NG;
    r = loc % R;
    c = (sigma % NG) * G + loc / R;
    if (b >= nk || c >= Ct) return;
  }
  sig_tile<BN, NSTG, NW>(s, as + as_stride * b, bs + bs_stride * b, wgt + g_stride * b, n, nw1, v0, nwg, ncol, MP,
                         NS, NCP, r * QBM, c * BN, accumulate, cz + cz_stride * b);
}

// Grouped, persistent: internal k-points with different column counts.  The
// tile list holds (k, row tile, column tile) for every k; each block walks it
// with a grid stride (all tiles cost the same) until the list is exhausted.
struct SigGroup {
  int64_t boff;   // B planes of k (elements)
  int64_t goff;   // weights of k (elements)
  int64_t czoff;  // output of k (elements)
  int ncol, ncp;  // columns, padded columns
};
template <int BN, int NSTG, int NW>
__global__ __launch_bounds__(NW * 64, 1) void sig_bf3_grouped_kernel(const int* __restrict__ tiles, int ntiles,
                                                                      const SigGroup* __restrict__ grp,
                                                                      const unsigned short* __restrict__ as,
                                                                      int64_t as_stride,
                                                                      const unsigned short* __restrict__ bs,
                                                                      const double* __restrict__ wgt, int n, int nw1,
                                                                      int v0, int nwg, int MP, int NS, int accumulate,
                                                                      cf* __restrict__ cz) {
  __shared__ SigSmem<BN, NSTG> s;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int k = tiles[3 * t], it = tiles[3 * t + 1], jt = tiles[3 * t + 2];
    const SigGroup gk = grp[k];
    __syncthreads();  // the previous tile's LDS reads are done before this tile's loads land
    sig_tile<BN, NSTG, NW>(s, as + as_stride * k, bs + gk.boff, wgt + gk.goff, n, nw1, v0, nwg, gk.ncol, MP, NS,
                           gk.ncp, it * QBM, jt * BN, accumulate, cz + gk.czoff);
  }
}

// ---------------------------------------------------------------- launcher

// Kernel variants of the tuner: {column tile, stages, waves, XCD-aware order}
struct SigVariant { int bn, nstg, nw, xcd; const char* name; };
// The last two are the frequency-inner three-product form (sig3m.hip), whose
// operands are prepared differently and whose sums round differently.
static const SigVariant kSigVariants[] = {{64, 3, 4, 0, "128x64x3"},
                                          {128, 2, 4, 0, "128x128x2"},
                                          {128, 2, 8, 0, "128x128x2/8w"},
                                          {128, 2, 8, 1, "128x128x2/8w/xcd"},
                                          {128, 3, 8, 0, "3m/128x128"},
                                          {128, 3, 8, 1, "3m/128x128/xcd"},
                                          {128, 4, 8, 0, "vin4/128x128"},
                                          {128, 4, 8, 1, "vin4/128x128/xcd"},
                                          {128, 3, 8, 0, "3m/128x128/w8"},
                                          {128, 3, 8, 1, "3m/128x128/w8/xcd"},
                                          {128, 3, 4, 0, "3m32/128x128"},
                                          {128, 3, 4, 1, "3m32/128x128/xcd"}};
constexpr int kNumSigVariants = 12;
constexpr int kFirst3m = 4;
constexpr int kFirstVin4 = 6;
constexpr int kFirst3mW8 = 8;   // three-product form, 8 waves side by side
constexpr int kFirst3m32 = 10;  // three-product form on the 32x32x16 MFMA, 4 waves
// operand preparation of a variant: 0 = W and zmel split into bf16 planes;
// 1 = six W planes + padded complex64 zmel (3m); 2 = four W planes + padded zmel
static int sig_form(int var) {
  return var >= kFirst3mW8 ? 1 : var >= kFirstVin4 ? 2 : var >= kFirst3m ? 1 : 0;
}
// Grouped kernel: the first three tile shapes, then the three-product form
// in its two 8-wave layouts (each with its own tile order)
constexpr int kNumSigGroupedVariants = 5;
static const char* const kSigGroupedNames[kNumSigGroupedVariants] = {"128x64x3", "128x128x2", "128x128x2/8w",
                                                                      "3m/128x128/xcd", "3m/128x128/w8/xcd"};

// Column tiles per super-tile of the XCD-aware order: the super-tile fills
// about 32 CUs (one XCD) and its zmel panels (G x BN columns x NP x 8 bytes)
// stay within ~3.5 MB of the XCD's 4 MB L2.
static int xcd_group(int R, int Ct, int BN, int NP) {
  int G = std::max(1, 32 / std::max(R, 1));
  const int64_t panel = (int64_t)BN * NP * 8;
  while (G > 1 && G * panel > (int64_t)3584 * 1024) --G;
  return std::min(G, Ct);
}

template <int BN, int NSTG, int NW>
static ecalj_status sig_launch_t(ecalj_ctx* h, int nb, int n, int ncol, int nw1, int MP, int NS, int NCP,
                                 const unsigned short* as, int64_t as_stride, const unsigned short* bs,
                                 int64_t bs_stride, const double* wgt, int64_t g_stride, cf* cz, int64_t cz_stride,
                                 bool xcd) {
  dim3 grid(MP / QBM, NCP / BN, nb);
  int G = 0;
  if (xcd) {
    const int R = MP / QBM, Ct = NCP / BN;
    G = xcd_group(R, Ct, BN, NS * QBK);
    const int64_t NG = (Ct + G - 1) / G, S = (int64_t)R * G;
    const int64_t blocks = 8 * S * ((nb * NG + 7) / 8);
    if (blocks > 0x7fffffff) return ECALJ_ERR_DIM;
    grid = dim3((unsigned)blocks, 1, 1);
  } else if (grid.y > 65535 || grid.z > 65535) {
    return ECALJ_ERR_DIM;
  }
  for (int v0 = 0; v0 < nw1; v0 += QNW) {
    const int nwg = std::min(QNW, nw1 - v0);
    hipLaunchKernelGGL((sig_bf3_kernel<BN, NSTG, NW>), grid, dim3(NW * 64), 0, h->stream, as, as_stride, bs,
                       bs_stride, wgt, g_stride, n, nw1, v0, nwg, ncol, MP, NS, NCP, v0 > 0 ? 1 : 0, cz, cz_stride, G,
                       nb);
    ECALJ_HIP(hipGetLastError());
  }
  return ECALJ_OK;
}

// Computes czmelwc for nb internal k-points; returns ECALJ_OK or a status.
ecalj_status sig_bf3(ecalj_ctx* h, int nb, const cf* zmel, int64_t z_stride, int n, int ncol, int nw1,
                     const cd* wc, int64_t ldw, int64_t w_stride, const double* wgt, int64_t g_stride, cf* cz,
                     int64_t cz_stride, unsigned short* as, unsigned short* bs) {
  if (n < 1 || ncol < 1 || nw1 < 1 || nb < 1) return ECALJ_ERR_DIM;
  const int MP = pad_to(n, QBM), NP = pad_to(n, QBK), NCP = pad_to(ncol, QBN_MAX), NS = NP / QBK;
  const int64_t as_stride = 4LL * nw1 * MP * NP;
  const int64_t bs_stride = 4LL * NCP * NP;
  // operands of a variant's form: four W planes and four zmel planes, or
  // (three-product form) six W planes and the padded complex64 zmel
  auto prep = [&](int form) -> ecalj_status {
    if (form == 1)
      return sig3m_prep(h, nb, zmel, z_stride, n, ncol, nw1, wc, ldw, w_stride, as, reinterpret_cast<cf*>(bs));
    // W planes: the split grid covers all MP x NP, so padding comes out zero
    dim3 gw((NS + SPLIT_KS - 1) / SPLIT_KS, MP / SPT, nw1 * nb);
    if (gw.z > 65535 || gw.y > 65535) return ECALJ_ERR_DIM;
    hipLaunchKernelGGL(split_w_kernel, gw, dim3(256), 0, h->stream, wc, ldw, w_stride, n, nw1, MP, NS, as,
                       as_stride);
    ECALJ_HIP(hipGetLastError());
    if (form == 2) return sig_pad_z(h, nb, zmel, z_stride, n, ncol, reinterpret_cast<cf*>(bs));
    const int64_t totz = (int64_t)NS * NCP * 4 * nb;
    const int bz = (int)std::min<int64_t>((totz + 255) / 256, 65536);
    hipLaunchKernelGGL(split_z_kernel, dim3(bz), dim3(256), 0, h->stream, zmel, z_stride, n, ncol, NCP, NS, bs,
                       bs_stride, nb);
    ECALJ_HIP(hipGetLastError());
    return ECALJ_OK;
  };
  auto run = [&](int var) -> ecalj_status {
    const bool xcd = kSigVariants[var].xcd != 0;
    if (var >= kFirst3m32)
      return sig3m_run(h, xcd, 0, nb, n, ncol, nw1, as, reinterpret_cast<const cf*>(bs), wgt, g_stride, cz,
                       cz_stride);
    if (var >= kFirst3mW8)
      return sig3m_run(h, xcd, 1, nb, n, ncol, nw1, as, reinterpret_cast<const cf*>(bs), wgt, g_stride, cz,
                       cz_stride);
    if (var >= kFirstVin4)
      return sigv4_run(h, xcd, nb, n, ncol, nw1, as, as_stride, reinterpret_cast<const cf*>(bs), wgt, g_stride, cz,
                       cz_stride);
    if (var >= kFirst3m)
      return sig3m_run(h, xcd, 2, nb, n, ncol, nw1, as, reinterpret_cast<const cf*>(bs), wgt, g_stride, cz,
                       cz_stride);
    if (kSigVariants[var].bn == 64)
      return sig_launch_t<64, 3, 4>(h, nb, n, ncol, nw1, MP, NS, NCP, as, as_stride, bs, bs_stride, wgt, g_stride, cz,
                                    cz_stride, xcd);
    if (kSigVariants[var].nw == 4)
      return sig_launch_t<128, 2, 4>(h, nb, n, ncol, nw1, MP, NS, NCP, as, as_stride, bs, bs_stride, wgt, g_stride,
                                     cz, cz_stride, xcd);
    return sig_launch_t<128, 2, 8>(h, nb, n, ncol, nw1, MP, NS, NCP, as, as_stride, bs, bs_stride, wgt, g_stride, cz,
                                   cz_stride, xcd);
  };
  // variant: forced by the handle, cached for this shape, or measured now
  // (the variants of one form compute the same sums in the same order; after
  // measuring, the chosen variant runs once more, so the result is always its own)
  int var = -1;
  const TuneKey key{TUNE_SIGC_BF3, n, ncol, nw1, nb};
  if (h->tune_mode > 0) {
    var = std::min(h->tune_mode, kNumSigVariants) - 1;
  } else if (tune_lookup(h, key, &var)) {
  } else if (h->capturing) {  // no timing under capture: first variant
    var = 0;
  } else {
    // each variant is charged with its own operand preparation
    std::vector<double> ms(kNumSigVariants);
    for (int f = 0; f < 3; ++f) {
      double pm = 0.0;
      ECALJ_TRY(tune_time(h, [&] { return prep(f); }, &pm));
      for (int v = 0; v < kNumSigVariants; ++v)
        if (sig_form(v) == f) {
          ECALJ_TRY(tune_time(h, [&] { return run(v); }, &ms[v]));
          ms[v] += pm;
        }
    }
    var = (int)(std::min_element(ms.begin(), ms.end()) - ms.begin());
    // One launch per variant shows the first, cool launch.  Over consecutive
    // launches the XCD-aware order, which fetches ~3x fewer bytes through the
    // fabric, runs faster than a variant that beat it by a few percent once
    // (C2: 22.6 against 23.2 ms per launch over 8 q-point steps), so it is
    // kept when within 3% of the fastest single launch.
    // (measured against the fastest launch itself, and the fastest XCD-aware
    // variant of those within 3%, so that the choice cannot drift along a
    // chain of variants each within 3% of the previous one)
    if (!kSigVariants[var].xcd) {
      const double lim = 1.03 * ms[var];
      int pick = -1;
      for (int v = 0; v < kNumSigVariants; ++v)
        if (kSigVariants[v].xcd && ms[v] <= lim && (pick < 0 || ms[v] < ms[pick])) pick = v;
      if (pick >= 0) var = pick;
    }
    std::vector<const char*> names;
    for (int v = 0; v < kNumSigVariants; ++v) names.push_back(kSigVariants[v].name);
    tune_store(h, key, var, "sigc_bf3", names, ms);
    ECALJ_TRY(prep(sig_form(var)));
    return run(var);
  }
  ECALJ_TRY(prep(sig_form(var)));
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 0, &tev));
  ECALJ_TRY(run(var));
  ECALJ_TRY(timing_end(h, 0, tev));
  return ECALJ_OK;
}

// Grouped form: nk internal k-points with column counts ncol[k] (host); zmel,
// wgt and cz packed per k (n x ncol[k], nw1 x ncol[k], n x ncol[k]).
ecalj_status sig_bf3_grouped(ecalj_ctx* h, int nk, const int* ncol, const cf* zmel, int n, int nw1, const cd* wc,
                             int64_t ldw, int64_t w_stride, const double* wgt, cf* cz, unsigned short* as,
                             unsigned short* bs) {
  if (n < 1 || nw1 < 1 || nk < 1) return ECALJ_ERR_DIM;
  const int MP = pad_to(n, QBM), NP = pad_to(n, QBK), NS = NP / QBK;
  const int64_t as_stride = 4LL * nw1 * MP * NP;
  std::vector<SigGroup> grp(nk);
  int64_t bo = 0, go = 0, co = 0, cols = 0;
  for (int k = 0; k < nk; ++k) {
    if (ncol[k] < 0) return ECALJ_ERR_DIM;
    grp[k].boff = bo;
    grp[k].goff = go;
    grp[k].czoff = co;
    grp[k].ncol = ncol[k];
    grp[k].ncp = pad_to(std::max(ncol[k], 1), QBN_MAX);
    bo += 4LL * grp[k].ncp * NP;
    go += (int64_t)nw1 * ncol[k];
    co += (int64_t)n * ncol[k];
    cols += ncol[k];
  }
  if (cols == 0) return ECALJ_OK;
  constexpr int k3m = 3;  // and k3m + 1: the three-product form, wave layouts 2 x 4 and 1 x 8
  auto prep = [&](bool is3m) -> ecalj_status {
    if (is3m)
      return sig3m_grouped_prep(h, nk, ncol, zmel, n, nw1, wc, ldw, w_stride, as, reinterpret_cast<cf*>(bs));
    dim3 gw((NS + SPLIT_KS - 1) / SPLIT_KS, MP / SPT, nw1 * nk);
    if (gw.z > 65535 || gw.y > 65535) return ECALJ_ERR_DIM;
    hipLaunchKernelGGL(split_w_kernel, gw, dim3(256), 0, h->stream, wc, ldw, w_stride, n, nw1, MP, NS, as,
                       as_stride);
    ECALJ_HIP(hipGetLastError());
    for (int k = 0; k < nk; ++k) {
      if (ncol[k] == 0) continue;
      const int64_t totz = (int64_t)NS * grp[k].ncp * 4;
      const int bz = (int)std::min<int64_t>((totz + 255) / 256, 65536);
      hipLaunchKernelGGL(split_z_kernel, dim3(bz), dim3(256), 0, h->stream, zmel + grp[k].czoff, (int64_t)0, n,
                         ncol[k], grp[k].ncp, NS, bs + grp[k].boff, (int64_t)0, 1);
      ECALJ_HIP(hipGetLastError());
    }
    return ECALJ_OK;
  };
  // table: [SigGroup x nk][tiles: (k, row tile, column tile) x ntiles]
  auto run = [&](int var) -> ecalj_status {
    if (var >= k3m)
      return sig3m_grouped_run(h, var == k3m ? 2 : 1, nk, ncol, n, nw1, as, reinterpret_cast<const cf*>(bs), wgt,
                               cz);
    const int bn = kSigVariants[var].bn;
    std::vector<int> tiles;
    for (int k = 0; k < nk; ++k)
      for (int it = 0; it < MP / QBM; ++it)
        for (int jt = 0; jt < (ncol[k] + bn - 1) / bn; ++jt) {
          tiles.push_back(k);
          tiles.push_back(it);
          tiles.push_back(jt);
        }
    const int ntiles = (int)(tiles.size() / 3);
    std::vector<unsigned char> blob(sizeof(SigGroup) * nk + sizeof(int) * tiles.size());
    memcpy(blob.data(), grp.data(), sizeof(SigGroup) * nk);
    memcpy(blob.data() + sizeof(SigGroup) * nk, tiles.data(), sizeof(int) * tiles.size());
    unsigned char* d;
    ECALJ_TRY(upload_blob(h, ecalj_ctx::S_GTAB, blob, &d));
    const SigGroup* dg = reinterpret_cast<const SigGroup*>(d);
    const int* dt = reinterpret_cast<const int*>(d + sizeof(SigGroup) * nk);
    const int grid = std::max(1, std::min(ntiles, h->ncu));
    for (int v0 = 0; v0 < nw1; v0 += QNW) {
      const int nwg = std::min(QNW, nw1 - v0);
      if (var == 0)
        hipLaunchKernelGGL((sig_bf3_grouped_kernel<64, 3, 4>), dim3(grid), dim3(256), 0, h->stream, dt, ntiles, dg,
                           as, as_stride, bs, wgt, n, nw1, v0, nwg, MP, NS, v0 > 0 ? 1 : 0, cz);
      else if (var == 1)
        hipLaunchKernelGGL((sig_bf3_grouped_kernel<128, 2, 4>), dim3(grid), dim3(256), 0, h->stream, dt, ntiles,
                           dg, as, as_stride, bs, wgt, n, nw1, v0, nwg, MP, NS, v0 > 0 ? 1 : 0, cz);
      else
        hipLaunchKernelGGL((sig_bf3_grouped_kernel<128, 2, 8>), dim3(grid), dim3(512), 0, h->stream, dt, ntiles,
                           dg, as, as_stride, bs, wgt, n, nw1, v0, nwg, MP, NS, v0 > 0 ? 1 : 0, cz);
      ECALJ_HIP(hipGetLastError());
    }
    return ECALJ_OK;
  };
  int var = -1;
  const TuneKey key{TUNE_SIGC_BF3_GROUPED, n, cols, nw1, nk};
  if (h->tune_mode > 0) {
    var = (h->tune_mode - 1) % kNumSigGroupedVariants;
  } else if (tune_lookup(h, key, &var)) {
  } else if (h->capturing) {
    var = 0;
  } else {
    std::vector<double> ms(kNumSigGroupedVariants);
    for (int f = 0; f < 2; ++f) {
      double pm = 0.0;
      ECALJ_TRY(tune_time(h, [&] { return prep(f == 1); }, &pm));
      for (int v = 0; v < kNumSigGroupedVariants; ++v)
        if ((v >= k3m) == (f == 1)) {
          ECALJ_TRY(tune_time(h, [&] { return run(v); }, &ms[v]));
          ms[v] += pm;
        }
    }
    var = (int)(std::min_element(ms.begin(), ms.end()) - ms.begin());
    std::vector<const char*> names(kSigGroupedNames, kSigGroupedNames + kNumSigGroupedVariants);
    tune_store(h, key, var, "sigc_bf3_grouped", names, ms);
    ECALJ_TRY(prep(var >= k3m));
    return run(var);
  }
  ECALJ_TRY(prep(var >= k3m));
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 0, &tev));
  ECALJ_TRY(run(var));
  ECALJ_TRY(timing_end(h, 0, tev));
  return ECALJ_OK;
}

// W plane elements per k of the largest form (the three-product form's six planes)
size_t sig_bf3_a_elems_per_k(int n, int nw1) {
  return std::max((size_t)F8_ROWB / 2 * nw1 * pad_to(n, QBM) * (pad_to(n, QBK) / QBK),
                  sig3m_a_elems_per_k(n, nw1));
}

// ---------------------------------------------------------------- plain batched GEMM on the same core
// C_b (m x ncol, leading dimension m, complex64) = alpha * op(A_b) op(B_b),
// K = k: the operands are split into the pre-tiled bf16 hi/lo planes of the
// Sigma_c kernel (one frequency, weights alpha) and the Sigma_c kernel runs.
// op(X) = X, or X^H when the flag is set.  A stride / B stride 0: one operand
// shared by every batch (split once).

// A planes [b][((ks*4 + p)*MP + I)*32 + slot]; one thread per (b, ks, chunk, I), I fastest
template <class TIN>
__global__ __launch_bounds__(256) void split_a_kernel(const TIN* __restrict__ a, int64_t lda, int64_t a_stride,
                                                      int m, int k, int herm, int MP, int NS,
                                                      unsigned short* __restrict__ as, int64_t as_stride, int nb) {
  typedef unsigned short us8v __attribute__((ext_vector_type(8)));
  const int64_t per = (int64_t)NS * 4 * MP, tot = per * nb;
  const int64_t pst = (int64_t)MP * QBK;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e % per;
    const int I = (int)(r % MP);
    const int ch = (int)((r / MP) % 4), ks = (int)(r / (4 * (int64_t)MP));
    us8v rh, rl, ih, il;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int J = ks * QBK + 8 * ch + q;
      float xr = 0.f, xi = 0.f;
      if (I < m && J < k) {
        const TIN x = herm ? a[a_stride * b + J + lda * I] : a[a_stride * b + I + lda * J];
        xr = (float)x.re;
        xi = herm ? -(float)x.im : (float)x.im;
      }
      unsigned short h, l;
      split_bf16(xr, h, l);
      rh[q] = h;
      rl[q] = l;
      split_bf16(xi, h, l);
      ih[q] = h;
      il[q] = l;
    }
    unsigned short* A = as + as_stride * b + (int64_t)ks * 4 * pst + (int64_t)I * QBK + 8 * chunk_slot(I, ch);
    *reinterpret_cast<us8v*>(A) = rh;
    *reinterpret_cast<us8v*>(A + pst) = rl;
    *reinterpret_cast<us8v*>(A + 2 * pst) = ih;
    *reinterpret_cast<us8v*>(A + 3 * pst) = il;
  }
}

// B planes [b][((ks*4 + p)*NCP + c)*32 + slot]; one thread per (b, ks, chunk, c), c fastest
template <class TIN>
__global__ __launch_bounds__(256) void split_b_kernel(const TIN* __restrict__ bm, int64_t ldb, int64_t b_stride,
                                                      int k, int ncol, int herm, int NCP, int NS,
                                                      unsigned short* __restrict__ bs, int64_t bs_stride, int nb) {
  typedef unsigned short us8v __attribute__((ext_vector_type(8)));
  const int64_t per = (int64_t)NS * 4 * NCP, tot = per * nb;
  const int64_t pst = (int64_t)NCP * QBK;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e % per;
    const int c = (int)(r % NCP);
    const int ch = (int)((r / NCP) % 4), ks = (int)(r / (4 * (int64_t)NCP));
    us8v rh, rl, ih, il;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int J = ks * QBK + 8 * ch + q;
      float xr = 0.f, xi = 0.f;
      if (c < ncol && J < k) {
        const TIN x = herm ? bm[b_stride * b + c + ldb * J] : bm[b_stride * b + J + ldb * c];
        xr = (float)x.re;
        xi = herm ? -(float)x.im : (float)x.im;
      }
      unsigned short h, l;
      split_bf16(xr, h, l);
      rh[q] = h;
      rl[q] = l;
      split_bf16(xi, h, l);
      ih[q] = h;
      il[q] = l;
    }
    unsigned short* B = bs + bs_stride * b + (int64_t)ks * 4 * pst + (int64_t)c * QBK + 8 * chunk_slot(c, ch);
    *reinterpret_cast<us8v*>(B) = rh;
    *reinterpret_cast<us8v*>(B + pst) = rl;
    *reinterpret_cast<us8v*>(B + 2 * pst) = ih;
    *reinterpret_cast<us8v*>(B + 3 * pst) = il;
  }
}

template <class TIN>
static ecalj_status split_a(ecalj_ctx* h, const void* a, int64_t lda, int64_t a_stride, int m, int k, bool herm,
                            int MP, int NS, unsigned short* as, int64_t as_stride, int nb) {
  const int64_t tot = (int64_t)NS * 4 * MP * nb;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(split_a_kernel<TIN>, dim3(blocks), dim3(256), 0, h->stream, (const TIN*)a, lda, a_stride, m, k,
                     herm ? 1 : 0, MP, NS, as, as_stride, nb);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

template <class TIN>
static ecalj_status split_b(ecalj_ctx* h, const void* b, int64_t ldb, int64_t b_stride, int k, int ncol, bool herm,
                            int NCP, int NS, unsigned short* bs, int64_t bs_stride, int nb) {
  const int64_t tot = (int64_t)NS * 4 * NCP * nb;
  const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(split_b_kernel<TIN>, dim3(blocks), dim3(256), 0, h->stream, (const TIN*)b, ldb, b_stride, k,
                     ncol, herm ? 1 : 0, NCP, NS, bs, bs_stride, nb);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status qgemm_bf3(ecalj_ctx* h, int nb, int m, int ncol, int k, const QOperand& A, const QOperand& B,
                       double alpha, cf* C, int64_t c_stride) {
  if (nb < 1 || m < 1 || ncol < 1 || k < 1 || nb > 65535) return ECALJ_ERR_DIM;
  if (!A.p || !B.p || !C) return ECALJ_ERR_ARG;
  const int MP = pad_to(m, QBM), NP = pad_to(k, QBK), NCP = pad_to(ncol, QBN_MAX), NS = NP / QBK;
  if (MP / QBM > 65535 || NCP / QBN_MAX > 65535) return ECALJ_ERR_DIM;
  const int na = A.stride ? nb : 1, nbb = B.stride ? nb : 1;
  const int64_t ape = 4LL * MP * NP, bpe = 4LL * NCP * NP;
  unsigned short *as, *bs;
  double* w;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_QA, sizeof(unsigned short) * (size_t)ape * na, (void**)&as));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_QB, sizeof(unsigned short) * (size_t)bpe * nbb, (void**)&bs));
  std::vector<unsigned char> wb(sizeof(double) * (size_t)ncol);
  for (int c = 0; c < ncol; ++c) std::memcpy(wb.data() + sizeof(double) * c, &alpha, sizeof(double));
  unsigned char* dw = nullptr;
  ECALJ_TRY(upload_blob(h, ecalj_ctx::S_QW, wb, &dw));
  w = (double*)dw;
  if (A.wide) ECALJ_TRY(split_a<cd>(h, A.p, A.ld, A.stride, m, k, A.herm, MP, NS, as, ape, na));
  else ECALJ_TRY(split_a<cf>(h, A.p, A.ld, A.stride, m, k, A.herm, MP, NS, as, ape, na));
  if (B.wide) ECALJ_TRY(split_b<cd>(h, B.p, B.ld, B.stride, k, ncol, B.herm, NCP, NS, bs, bpe, nbb));
  else ECALJ_TRY(split_b<cf>(h, B.p, B.ld, B.stride, k, ncol, B.herm, NCP, NS, bs, bpe, nbb));
  // the 8-wave 128 x 128 variant (the tuner's choice at the Sigma_c shapes); the output has leading dimension m
  return sig_launch_t<128, 2, 8>(h, nb, m, ncol, 1, MP, NS, NCP, as, A.stride ? ape : 0, bs, B.stride ? bpe : 0, w,
                                 0, C, c_stride, false);
}

// Scratch per internal k, sized for the larger of the two operand layouts:
// bf16x3 keeps 4 bf16 planes (128 bytes per row and 32-wide k slice), the fp8
// cross-term form 2 bf16 planes + 4 fp8 planes + 4 scale bytes (260 bytes).
size_t sig_bf3_bytes_per_k(int n, int ncol, int nw1) {
  return 2 * (sig_bf3_a_elems_per_k(n, nw1) + (size_t)F8_ROWB / 2 * pad_to(ncol, QBN_MAX) * (pad_to(n, QBK) / QBK));
}


// ================================================================ fp8 cross-term form
// The same contraction with each real product a*b (a = ah + al, ah = bf16(a))
// taken as ah*bh on the bf16 MFMA plus ah*bl + al*bh on the scaled fp8 MFMA
// (16x16x128, E4M3 operands, one E8M0 scale per 32 values of a row); al*bl
// (|al*bl| <= 2^-18 |ab|) is dropped.  Rounding ah and al to E4M3 (3 mantissa
// bits) costs ~2^-4 of each cross term, ~2^-12 of the product, against
// ~2^-17 for bf16x3; the MFMA time per 32-wide k slice and 16x16 tile drops
// from 12 bf16 instructions (192 cycles) to 4 bf16 + 2 fp8 (128 cycles).
//
// The four cross products of one complex 32-wide k slice fill the 128 k of
// one fp8 instruction, one operand plane per 32-value scale block
// (planes 0..3 = re hi, re lo, im hi, im lo):
//   cr += [arh arl aih ail] . [brl brh -bil -bih]
//   ci += [arh arl aih ail] . [bil bih brl brh]
// Operand layout of the scaled 16x16x128 fp8 MFMA (measured on gfx950): lane
// (r, g) (r = lane & 15, g = lane >> 4) holds k = 16g..16g+15 in bytes 0-15
// and k = 64+16g.. in bytes 16-31; the scale of block s (k = 32s..32s+31)
// comes from lane (r, s).  Block s = plane s therefore needs, in lane g,
// plane g>>1 (bytes 0-15) and plane 2+(g>>1) (bytes 16-31), each its half
// g&1 of the 32-wide slice.
//
// Pre-tiled operand section per 32-wide k slice (rows = W rows I or zmel
// columns c, R of them): [2 bf16 planes (re hi, im hi) x R x 64 bytes, chunk
// order of chunk_slot][fp8 block R x 128 bytes][scales R x 4 bytes].  An fp8
// row holds 8 slots of 16 bytes: logical slot 2g + h = plane (g>>1) + 2h,
// half g&1, stored at slot (2g + h) ^ f8_code(row), which keeps the 16-byte
// reads of 8 (or 16) consecutive rows on distinct banks.  Scale byte p of a
// row = E8M0 exponent of plane p.

__device__ __host__ inline int f8_code(int row) { return (row ^ ((row >> 3) & 1)) & 7; }

// E8M0 exponent e for a block with max |x| = m: x * 2^-e puts the largest
// value in [128, 256), inside E4M3's range (448); 0 for an all-zero block.
__device__ inline int f8_block_exp(float m) {
  if (!(m > 1e-30f)) return 0;
  int E;
  frexpf(m, &E);
  return E - 8;
}

// 8 floats -> 8 E4M3 bytes (round to nearest even) after scaling by 2^-e
__device__ inline uint2 f8_pack8(const float* x, int e) {
  unsigned int w0 = 0, w1 = 0;
  w0 = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(x[0], -e), ldexpf(x[1], -e), w0, false);
  w0 = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(x[2], -e), ldexpf(x[3], -e), w0, true);
  w1 = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(x[4], -e), ldexpf(x[5], -e), w1, false);
  w1 = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(x[6], -e), ldexpf(x[7], -e), w1, true);
  return make_uint2(w0, w1);
}

// One row's 8 k values (k = 8ch..8ch+7 of the slice) of a complex operand
// into a section of R rows: bf16 hi planes, fp8 planes, scales.  The four
// threads ch = 0..3 of the row are consecutive lanes of one wave (the block
// maxima are taken across them) and must all call this.
__device__ inline void f8_put_row(const float* xr, const float* xi, int row, int ch, int R,
                                  unsigned char* __restrict__ sec) {
  typedef unsigned short us8v __attribute__((ext_vector_type(8)));
  us8v rh, ih;
  float pl[4][8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    rh[e] = bf16_bits(xr[e]);
    ih[e] = bf16_bits(xi[e]);
    pl[0][e] = bf16_float(rh[e]);
    pl[1][e] = xr[e] - pl[0][e];
    pl[2][e] = bf16_float(ih[e]);
    pl[3][e] = xi[e] - pl[2][e];
  }
  *reinterpret_cast<us8v*>(sec + (int64_t)row * 64 + 16 * chunk_slot(row, ch)) = rh;
  *reinterpret_cast<us8v*>(sec + (int64_t)R * 64 + (int64_t)row * 64 + 16 * chunk_slot(row, ch)) = ih;
  unsigned char* f8 = sec + (int64_t)R * 128 + (int64_t)row * 128;
  unsigned int sw = 0;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    float m = 0.f;
#pragma unroll
    for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(pl[p][e]));
    m = fmaxf(m, __shfl_xor(m, 1));
    m = fmaxf(m, __shfl_xor(m, 2));
    const int ex = f8_block_exp(m);
    const int g = 2 * (p & 1) + (ch >> 1), t = 2 * g + (p >> 1);
    *reinterpret_cast<uint2*>(f8 + 16 * (t ^ f8_code(row)) + 8 * (ch & 1)) = f8_pack8(pl[p], ex);
    sw |= (unsigned int)(127 + ex) << (8 * p);
  }
  if (ch == 0) *reinterpret_cast<unsigned int*>(sec + (int64_t)R * 256 + (int64_t)row * 4) = sw;
}

// W planes: section (b, v, ks) at as + as_stride*b + (v*NS + ks)*F8_ROWB*MP
__global__ __launch_bounds__(256) void split_w8_kernel(const cd* __restrict__ w, int64_t ldw, int64_t w_stride,
                                                       int n, int nw1, int MP, int NS,
                                                       unsigned char* __restrict__ as, int64_t as_stride) {
  __shared__ cd up[SPT][SPT + 1];  // up[jj][ii] = W(I0+ii, J0+jj)
  __shared__ cd lo[SPT][SPT + 1];  // lo[ii][jj] = W(J0+jj, I0+ii)
  const int I0 = blockIdx.y * SPT;
  const int v = blockIdx.z % nw1, b = blockIdx.z / nw1;
  const cd* W = w + w_stride * b + (int64_t)ldw * n * v;
  const int tx = threadIdx.x % SPT, ty = threadIdx.x / SPT;
  for (int kq = 0; kq < SPLIT_KS; ++kq) {  // SPLIT_KS consecutive k slices per block
    const int ks = blockIdx.x * SPLIT_KS + kq;
    if (ks >= NS) break;
    const int J0 = ks * SPT;
    if (kq) __syncthreads();  // the previous slice's LDS reads are done
    const bool need_up = I0 <= J0 + SPT - 1, need_lo = I0 + SPT - 1 > J0;
    for (int r = ty; r < SPT; r += 8) {
      if (need_up) {
        const int I = I0 + tx, J = J0 + r;
        up[r][tx] = (I < n && J < n) ? W[I + ldw * J] : cd{0.0, 0.0};
      }
      if (need_lo) {
        const int J2 = J0 + tx, I2 = I0 + r;
        lo[r][tx] = (I2 < n && J2 < n) ? W[J2 + ldw * I2] : cd{0.0, 0.0};
      }
    }
    __syncthreads();
    if (threadIdx.x < SPT * 4) {  // waves 0 and 1, whole: 4 consecutive lanes per row
      const int r = threadIdx.x >> 2, ch = threadIdx.x & 3;
      const int I = I0 + r;
      float xr[8], xi[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int jj = 8 * ch + e, J = J0 + jj;
        double a = 0.0, c = 0.0;
        if (I < n && J < n) {
          if (I <= J) {
            a = up[jj][r].re;
            c = up[jj][r].im;
          } else {
            a = lo[r][jj].re;
            c = -lo[r][jj].im;
          }
        }
        xr[e] = (float)a;
        xi[e] = (float)c;
      }
      f8_put_row(xr, xi, I, ch, MP, as + as_stride * b + ((int64_t)v * NS + ks) * F8_ROWB * MP);
    }
  }
}

// zmel planes: section (b, ks) at bs + bs_stride*b + ks*F8_ROWB*NCP; one thread per (b, ks, c, ch)
__global__ __launch_bounds__(256) void split_z8_kernel(const cf* __restrict__ z, int64_t z_stride, int n, int ncol,
                                                       int NCP, int NS, unsigned char* __restrict__ bs,
                                                       int64_t bs_stride, int nb) {
  const int64_t per = (int64_t)NS * NCP * 4, tot = per * nb;
  // tot is a multiple of 4 and the stride keeps groups of 4 lanes together
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e % per;
    const int ks = (int)(r / (4 * NCP));
    const int c = (int)((r / 4) % NCP), ch = (int)(r % 4);
    float xr[8], xi[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int J = ks * QBK + 8 * ch + q;
      xr[q] = xi[q] = 0.f;
      if (c < ncol && J < n) {
        const cf x = z[z_stride * b + J + (int64_t)n * c];
        xr[q] = x.re;
        xi[q] = x.im;
      }
    }
    f8_put_row(xr, xi, c, ch, NCP, bs + bs_stride * b + (int64_t)ks * F8_ROWB * NCP);
  }
}

constexpr int F8_BN = 128;
struct alignas(16) F8Stage {
  unsigned char abf[2][QBM * 64];  // re hi, im hi: 8 KB each
  unsigned char af8[QBM * 128];    // 16 KB
  unsigned int asc[QBM];
  unsigned char bbf[2][F8_BN * 64];
  unsigned char bf8[F8_BN * 128];
  unsigned int bsc[F8_BN];
};
struct alignas(16) F8Smem {
  F8Stage st[2];
  float wt[QNW][F8_BN];
};

typedef int i4q __attribute__((ext_vector_type(4)));
typedef int i8q __attribute__((ext_vector_type(8)));
typedef int i12q __attribute__((ext_vector_type(12)));

// One 128 x 128 block tile, two stages; 8 waves in 2 x 4 (two per SIMD), wave
// tile 64 x 32 (~235 VGPRs).  A 16-wave form (64 x 16 wave tiles within 128
// VGPRs) measured the same time: per 32-wide k slice the block's MFMA work
// and its 65 KB of L2 -> LDS loads each take ~1.8 us at C2 and overlap, so
// neither more waves nor fewer loads alone shortens a stage.
constexpr int F8_NW = 8;
__device__ __forceinline__ void sig8_tile(F8Smem& s, const unsigned char* __restrict__ Ak,
                                          const unsigned char* __restrict__ Bk, const double* __restrict__ g,
                                          int n, int nw1, int v0, int nwg, int ncol, int MP, int NS, int NCP, int I0,
                                          int c0, int accumulate, cf* __restrict__ C) {
  constexpr int NW = F8_NW, WN = NW / 2, TM = 4, TN = F8_BN / (16 * WN), NT = NW * 64, RPW = 64 / NW;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6), wm = wave / WN, wn = wave % WN;
  const int64_t asec = (int64_t)F8_ROWB * MP, bsec = (int64_t)F8_ROWB * NCP;
  const int nst = nwg * NS;

  for (int e = t; e < nwg * F8_BN; e += NT) {
    const int v = e / F8_BN, c = c0 + e % F8_BN;
    s.wt[v][e % F8_BN] = (c < ncol) ? (float)g[(v0 + v) + (int64_t)nw1 * c] : 0.f;
  }
  __syncthreads();

  // stage loads: 64 runs of 1 KB, RPW consecutive ones per wave (runs 0-15:
  // A bf16 planes, 16-31: A fp8 block, 32-47: B bf16, 48-63: B fp8; a wave's
  // runs never straddle two of these), and 4 runs of 256 bytes of scales
  // (waves 0-3)
  auto issue = [&](int st, int buf) {
    const int v = v0 + st / NS, ks = st % NS;
    const int run0 = RPW * wave, cat = run0 / 16, sub0 = run0 % 16;
    const bool isb = cat >= 2;
    const unsigned char* base = isb ? Bk + (int64_t)ks * bsec : Ak + ((int64_t)v * NS + ks) * asec;
    const int64_t R = isb ? NCP : MP;
    const int r0 = isb ? c0 : I0;
    F8Stage& S = s.st[buf];
    if ((cat & 1) == 0) {  // bf16 planes: 8 runs each
      const unsigned char* src = base + (int64_t)r0 * 64 + lane * 16;
      unsigned char* dst = isb ? S.bbf[0] : S.abf[0];
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        const int sub = sub0 + q, pl = sub / 8, part = sub % 8;
        __builtin_amdgcn_global_load_lds((gbl_ptr_t)(src + (int64_t)pl * 64 * R + part * 1024),
                                         (lds_ptr_t)(dst + pl * (QBM * 64) + part * 1024), 16, 0, 0);
      }
    } else {  // fp8 block: 16 runs
      const unsigned char* src = base + 128 * R + (int64_t)r0 * 128 + lane * 16;
      unsigned char* dst = isb ? S.bf8 : S.af8;
#pragma unroll
      for (int q = 0; q < RPW; ++q)
        __builtin_amdgcn_global_load_lds((gbl_ptr_t)(src + (sub0 + q) * 1024), (lds_ptr_t)(dst + (sub0 + q) * 1024),
                                         16, 0, 0);
    }
    if (wave < 4) {  // scales: A halves (waves 0, 1), B halves (waves 2, 3)
      const bool sb = wave >= 2;
      const int hf = wave & 1;
      const unsigned char* sbase = sb ? Bk + (int64_t)ks * bsec : Ak + ((int64_t)v * NS + ks) * asec;
      const int64_t SR = sb ? NCP : MP;
      const int sr0 = sb ? c0 : I0;
      const unsigned char* src = sbase + 256 * SR + (int64_t)(sr0 + 64 * hf) * 4 + lane * 4;
      unsigned int* dst = (sb ? S.bsc : S.asc) + 64 * hf;
      __builtin_amdgcn_global_load_lds((gbl_ptr_t)src, (lds_ptr_t)dst, 4, 0, 0);
    }
  };

  f4 cr[TM][TN], ci[TM][TN], tr[TM][TN], ti[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) cr[i][j] = ci[i][j] = tr[i][j] = ti[i][j] = f4{0, 0, 0, 0};
  const int gl = lane >> 4, code = f8_code(lane & 15);
  const int slot = chunk_slot(lane & 15, gl);
  const int arow = wm * TM * 16 + (lane & 15), brow = wn * TN * 16 + (lane & 15);
  // fp8 16-byte reads: A' = stored lane group gl, B' / B'' from stored group gl ^ 2
  const int ao0 = 16 * ((2 * gl) ^ code), ao1 = 16 * ((2 * gl + 1) ^ code);
  const int bo0 = 16 * ((2 * (gl ^ 2)) ^ code), bo1 = 16 * ((2 * (gl ^ 2) + 1) ^ code);

  if (nst > 0) issue(0, 0);
  int st = 0;
  for (int vl = 0; vl < nwg; ++vl) {
    for (int ks = 0; ks < NS; ++ks, ++st) {
      const int buf = st & 1;
      ECALJ_WAIT_VMCNT(0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_barrier();
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      // the other buffer was last read in iteration st-1, which every wave has left
      if (st + 1 < nst) issue(st + 1, buf ^ 1);
      const F8Stage& S = s.st[buf];

      u4q brh[TN], bih[TN], nbh[TN];
      // per column tile one 12-dword run T = (y, x, -y): B'' = T[0:8] = (y, x),
      // B' = T[4:12] = (x, -y), overlapping windows of one register range
      i12q T[TN];
      int sbp[TN], sbq[TN];
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int r = brow + j * 16;
        brh[j] = *reinterpret_cast<const u4q*>(&S.bbf[0][r * 64 + 16 * slot]);
        bih[j] = *reinterpret_cast<const u4q*>(&S.bbf[1][r * 64 + 16 * slot]);
        nbh[j] = bih[j] ^ 0x80008000u;
        const i4q x = *reinterpret_cast<const i4q*>(&S.bf8[r * 128 + bo0]);
        const i4q y = *reinterpret_cast<const i4q*>(&S.bf8[r * 128 + bo1]);
        const i4q ny = y ^ (int)0x80808080u;
        const i8q yx = __builtin_shufflevector(y, x, 0, 1, 2, 3, 4, 5, 6, 7);
        const i8q n8 = __builtin_shufflevector(ny, ny, 0, 1, 2, 3, 0, 1, 2, 3);
        T[j] = __builtin_shufflevector(yx, n8, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11);
        const unsigned int w = S.bsc[r];
        sbp[j] = (w >> (8 * (gl ^ 1))) & 0xff;
        sbq[j] = (w >> (8 * (gl ^ 3))) & 0xff;
      }
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int r = arow + i * 16;
        const u4q arh = *reinterpret_cast<const u4q*>(&S.abf[0][r * 64 + 16 * slot]);
        const u4q aih = *reinterpret_cast<const u4q*>(&S.abf[1][r * 64 + 16 * slot]);
        const i4q a0 = *reinterpret_cast<const i4q*>(&S.af8[r * 128 + ao0]);
        const i4q a1 = *reinterpret_cast<const i4q*>(&S.af8[r * 128 + ao1]);
        const i8q a8 = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
        const int sa = (S.asc[r] >> (8 * gl)) & 0xff;
        // term outer, column tile inner: an accumulator's next MFMA is 2 TN
        // instructions behind its previous one (the MFMA result latency)
#define ECALJ_F8_BF(ACC, AF, BF)                                                                      \
  _Pragma("unroll") for (int j = 0; j < TN; ++j) ACC[i][j] =                                          \
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, AF), __builtin_bit_cast(bf8, BF[j]), \
                                              ACC[i][j], 0, 0, 0);
        ECALJ_F8_BF(cr, arh, brh)
        ECALJ_F8_BF(ci, arh, bih)
        ECALJ_F8_BF(cr, aih, nbh)
        ECALJ_F8_BF(ci, aih, brh)
#undef ECALJ_F8_BF
#pragma unroll
        for (int j = 0; j < TN; ++j)
          cr[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
              a8, __builtin_shufflevector(T[j], T[j], 4, 5, 6, 7, 8, 9, 10, 11), cr[i][j], 0, 0, 0, sa, 0, sbp[j]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          ci[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
              a8, __builtin_shufflevector(T[j], T[j], 0, 1, 2, 3, 4, 5, 6, 7), ci[i][j], 0, 0, 0, sa, 0, sbq[j]);
        // keep each row tile's 12 MFMAs in this order (the scheduler otherwise
        // chains an accumulator's MFMAs back to back at the ends of the stage)
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const float wv = s.wt[vl][wn * TN * 16 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        tr[i][j] += wv * cr[i][j];
        ti[i][j] += wv * ci[i][j];
        cr[i][j] = f4{0, 0, 0, 0};
        ci[i][j] = f4{0, 0, 0, 0};
      }
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int I = I0 + wm * TM * 16 + i * 16 + acc_row_f32(lane, e);
        const int c = c0 + wn * TN * 16 + j * 16 + (lane & 15);
        if (I < n && c < ncol) {
          cf* p = C + I + (int64_t)n * c;
          if (accumulate) {
            const cf o = *p;
            *p = cf{o.re + tr[i][j][e], o.im + ti[i][j][e]};
          } else {
            *p = cf{tr[i][j][e], ti[i][j][e]};
          }
        }
      }
}

// Uniform batch in the XCD-aware 1-D order of sig_bf3_kernel (G > 0)
__global__ __launch_bounds__(F8_NW * 64, 1) void sig_f8_kernel(const unsigned char* __restrict__ as,
                                                               int64_t as_stride,
                                                               const unsigned char* __restrict__ bs,
                                                               int64_t bs_stride, const double* __restrict__ wgt,
                                                               int64_t g_stride, int n, int nw1, int v0, int nwg,
                                                               int ncol, int MP, int NS, int NCP, int accumulate,
                                                               cf* __restrict__ cz, int64_t cz_stride, int G, int nk) {
  __shared__ F8Smem s;
  const int R = MP / QBM, Ct = NCP / F8_BN, NG = (Ct + G - 1) / G, S = R * G;
  const int p = blockIdx.x, x = p & 7, j = p >> 3;
  const int sigma = (j / S) * 8 + x, loc = j % S;
  const int b = sigma / NG, r = loc % R, c = (sigma % NG) * G + loc / R;
  if (b >= nk || c >= Ct) return;
  sig8_tile(s, as + as_stride * b, bs + bs_stride * b, wgt + g_stride * b, n, nw1, v0, nwg, ncol, MP, NS, NCP,
            r * QBM, c * F8_BN, accumulate, cz + cz_stride * b);
}

// Grouped, persistent (tile list of (k, row tile, column tile)), as sig_bf3_grouped_kernel
__global__ __launch_bounds__(F8_NW * 64, 1) void sig_f8_grouped_kernel(const int* __restrict__ tiles, int ntiles,
                                                                       const SigGroup* __restrict__ grp,
                                                                       const unsigned char* __restrict__ as,
                                                                       int64_t as_stride,
                                                                       const unsigned char* __restrict__ bs,
                                                                       const double* __restrict__ wgt, int n,
                                                                       int nw1, int v0, int nwg, int MP, int NS,
                                                                       int accumulate, cf* __restrict__ cz) {
  __shared__ F8Smem s;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int k = tiles[3 * t], it = tiles[3 * t + 1], jt = tiles[3 * t + 2];
    const SigGroup gk = grp[k];
    __syncthreads();
    sig8_tile(s, as + as_stride * k, bs + gk.boff, wgt + gk.goff, n, nw1, v0, nwg, gk.ncol, MP, NS, gk.ncp,
              it * QBM, jt * F8_BN, accumulate, cz + gk.czoff);
  }
}

ecalj_status sig_f8(ecalj_ctx* h, int nb, const cf* zmel, int64_t z_stride, int n, int ncol, int nw1, const cd* wc,
                    int64_t ldw, int64_t w_stride, const double* wgt, int64_t g_stride, cf* cz, int64_t cz_stride,
                    unsigned short* as16, unsigned short* bs16) {
  if (n < 1 || ncol < 1 || nw1 < 1 || nb < 1) return ECALJ_ERR_DIM;
  unsigned char* as = reinterpret_cast<unsigned char*>(as16);
  unsigned char* bs = reinterpret_cast<unsigned char*>(bs16);
  const int MP = pad_to(n, QBM), NP = pad_to(n, QBK), NCP = pad_to(ncol, QBN_MAX), NS = NP / QBK;
  const int64_t as_stride = (int64_t)F8_ROWB * MP * NS * nw1;  // bytes
  const int64_t bs_stride = (int64_t)F8_ROWB * NCP * NS;
  dim3 gw((NS + SPLIT_KS - 1) / SPLIT_KS, MP / SPT, nw1 * nb);
  if (gw.z > 65535 || gw.y > 65535) return ECALJ_ERR_DIM;
  hipLaunchKernelGGL(split_w8_kernel, gw, dim3(256), 0, h->stream, wc, ldw, w_stride, n, nw1, MP, NS, as, as_stride);
  ECALJ_HIP(hipGetLastError());
  const int64_t totz = (int64_t)NS * NCP * 4 * nb;
  const int bz = (int)std::min<int64_t>((totz + 255) / 256, 65536);
  hipLaunchKernelGGL(split_z8_kernel, dim3(bz), dim3(256), 0, h->stream, zmel, z_stride, n, ncol, NCP, NS, bs,
                     bs_stride, nb);
  ECALJ_HIP(hipGetLastError());
  const int R = MP / QBM, Ct = NCP / F8_BN;
  const int G = xcd_group(R, Ct, F8_BN, NP);
  const int64_t NG = (Ct + G - 1) / G, S = (int64_t)R * G;
  const int64_t blocks = 8 * S * ((nb * NG + 7) / 8);
  if (blocks > 0x7fffffff) return ECALJ_ERR_DIM;
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 0, &tev));
  for (int v0 = 0; v0 < nw1; v0 += QNW) {
    const int nwg = std::min(QNW, nw1 - v0);
    hipLaunchKernelGGL(sig_f8_kernel, dim3((unsigned)blocks), dim3(F8_NW * 64), 0, h->stream, as, as_stride, bs,
                       bs_stride, wgt, g_stride, n, nw1, v0, nwg, ncol, MP, NS, NCP, v0 > 0 ? 1 : 0, cz, cz_stride,
                       G, nb);
    ECALJ_HIP(hipGetLastError());
  }
  ECALJ_TRY(timing_end(h, 0, tev));
  return ECALJ_OK;
}

ecalj_status sig_f8_grouped(ecalj_ctx* h, int nk, const int* ncol, const cf* zmel, int n, int nw1, const cd* wc,
                            int64_t ldw, int64_t w_stride, const double* wgt, cf* cz, unsigned short* as16,
                            unsigned short* bs16) {
  if (n < 1 || nw1 < 1 || nk < 1) return ECALJ_ERR_DIM;
  unsigned char* as = reinterpret_cast<unsigned char*>(as16);
  unsigned char* bs = reinterpret_cast<unsigned char*>(bs16);
  const int MP = pad_to(n, QBM), NP = pad_to(n, QBK), NS = NP / QBK;
  const int64_t as_stride = (int64_t)F8_ROWB * MP * NS * nw1;
  std::vector<SigGroup> grp(nk);
  int64_t bo = 0, go = 0, co = 0, cols = 0;
  for (int k = 0; k < nk; ++k) {
    if (ncol[k] < 0) return ECALJ_ERR_DIM;
    grp[k].boff = bo;  // bytes
    grp[k].goff = go;
    grp[k].czoff = co;
    grp[k].ncol = ncol[k];
    grp[k].ncp = pad_to(std::max(ncol[k], 1), QBN_MAX);
    bo += (int64_t)F8_ROWB * grp[k].ncp * NS;
    go += (int64_t)nw1 * ncol[k];
    co += (int64_t)n * ncol[k];
    cols += ncol[k];
  }
  dim3 gw((NS + SPLIT_KS - 1) / SPLIT_KS, MP / SPT, nw1 * nk);
  if (gw.z > 65535 || gw.y > 65535) return ECALJ_ERR_DIM;
  hipLaunchKernelGGL(split_w8_kernel, gw, dim3(256), 0, h->stream, wc, ldw, w_stride, n, nw1, MP, NS, as, as_stride);
  ECALJ_HIP(hipGetLastError());
  for (int k = 0; k < nk; ++k) {
    if (ncol[k] == 0) continue;
    const int64_t totz = (int64_t)NS * grp[k].ncp * 4;
    const int bz = (int)std::min<int64_t>((totz + 255) / 256, 65536);
    hipLaunchKernelGGL(split_z8_kernel, dim3(bz), dim3(256), 0, h->stream, zmel + grp[k].czoff, (int64_t)0, n,
                       ncol[k], grp[k].ncp, NS, bs + grp[k].boff, (int64_t)0, 1);
    ECALJ_HIP(hipGetLastError());
  }
  if (cols == 0) return ECALJ_OK;
  std::vector<int> tiles;
  for (int k = 0; k < nk; ++k)
    for (int it = 0; it < MP / QBM; ++it)
      for (int jt = 0; jt < (ncol[k] + F8_BN - 1) / F8_BN; ++jt) {
        tiles.push_back(k);
        tiles.push_back(it);
        tiles.push_back(jt);
      }
  const int ntiles = (int)(tiles.size() / 3);
  std::vector<unsigned char> blob(sizeof(SigGroup) * nk + sizeof(int) * tiles.size());
  memcpy(blob.data(), grp.data(), sizeof(SigGroup) * nk);
  memcpy(blob.data() + sizeof(SigGroup) * nk, tiles.data(), sizeof(int) * tiles.size());
  unsigned char* d;
  ECALJ_TRY(upload_blob(h, ecalj_ctx::S_GTAB, blob, &d));
  const SigGroup* dg = reinterpret_cast<const SigGroup*>(d);
  const int* dt = reinterpret_cast<const int*>(d + sizeof(SigGroup) * nk);
  const int grid = std::max(1, std::min(ntiles, h->ncu));
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 0, &tev));
  for (int v0 = 0; v0 < nw1; v0 += QNW) {
    const int nwg = std::min(QNW, nw1 - v0);
    hipLaunchKernelGGL(sig_f8_grouped_kernel, dim3(grid), dim3(F8_NW * 64), 0, h->stream, dt, ntiles, dg, as,
                       as_stride, bs, wgt, n, nw1, v0, nwg, MP, NS, v0 > 0 ? 1 : 0, cz);
    ECALJ_HIP(hipGetLastError());
  }
  ECALJ_TRY(timing_end(h, 0, tev));
  return ECALJ_OK;
}


}  // namespace ecalj
