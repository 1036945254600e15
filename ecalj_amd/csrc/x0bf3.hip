// x0bf3.hip -- x0gemm (S/x0gemm.f90:2-97) in the bf16x3 form on the block tile of sigtile.hpp.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "sigtile.hpp"

namespace ecalj {

// ================================================================ x0gemm on the bf16x3 core
// rcxq(:,:,slice(bin)) += sum_{t in bin} conj(z_t(i)) w_t z_t(ipr_col-1+j)
// (S/x0gemm.f90:51-97) as one grouped GEMM per histogram bin: the bin's
// columns are gathered once into pre-split bf16 planes (A = conj(z_t), B =
// w_t z_t, K = the bin's entries padded to 32 with zeros), and every (bin,
// row tile, column tile) runs the Sigma_c tile code (one frequency, unit
// weights, accumulating).  A square output (ipr_col = 1, npr_col = npr) is
// Hermitian: only tiles ti <= tj run, off-diagonal ones also add the
// conjugate transpose.

// one thread per (global K slice s, row r, chunk ch), ch fastest: the four
// chunks of a row are adjacent lanes, so a wave writes 16 whole 64-byte rows
// of each plane
__global__ __launch_bounds__(256) void x0_split_kernel(const cf* __restrict__ z, const int64_t* __restrict__ colp,
                                                       const double* __restrict__ wp, int64_t nslice, int npr,
                                                       int npr_col, int64_t boff, int MP, int NCP, int RP,
                                                       unsigned short* __restrict__ as,
                                                       unsigned short* __restrict__ bs) {
  typedef unsigned short us8v __attribute__((ext_vector_type(8)));
  const int64_t tot = nslice * 4 * RP;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int ch = (int)(e % 4);
    const int r = (int)((e / 4) % RP);
    const int64_t sl = e / (4 * (int64_t)RP);
    us8v arh, arl, aih, ail, brh, brl, bih, bil;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int64_t t = sl * QBK + 8 * ch + q;
      const int64_t c = colp[t];
      float xr = 0.f, xi = 0.f, yr = 0.f, yi = 0.f;
      if (c >= 0) {
        if (r < npr) {
          const cf x = z[c + r];
          xr = x.re;
          xi = -x.im;
        }
        if (r < npr_col) {
          const float w = (float)wp[t];
          const cf y = z[c + boff + r];
          yr = w * y.re;
          yi = w * y.im;
        }
      }
      unsigned short h, l;
      split_bf16(xr, h, l);
      arh[q] = h;
      arl[q] = l;
      split_bf16(xi, h, l);
      aih[q] = h;
      ail[q] = l;
      split_bf16(yr, h, l);
      brh[q] = h;
      brl[q] = l;
      split_bf16(yi, h, l);
      bih[q] = h;
      bil[q] = l;
    }
    if (r < MP) {
      const int64_t pst = (int64_t)MP * QBK;
      unsigned short* A = as + sl * 4 * pst + (int64_t)r * QBK + 8 * chunk_slot(r, ch);
      *reinterpret_cast<us8v*>(A) = arh;
      *reinterpret_cast<us8v*>(A + pst) = arl;
      *reinterpret_cast<us8v*>(A + 2 * pst) = aih;
      *reinterpret_cast<us8v*>(A + 3 * pst) = ail;
    }
    if (r < NCP) {
      const int64_t pst = (int64_t)NCP * QBK;
      unsigned short* B = bs + sl * 4 * pst + (int64_t)r * QBK + 8 * chunk_slot(r, ch);
      *reinterpret_cast<us8v*>(B) = brh;
      *reinterpret_cast<us8v*>(B + pst) = brl;
      *reinterpret_cast<us8v*>(B + 2 * pst) = bih;
      *reinterpret_cast<us8v*>(B + 3 * pst) = bil;
    }
  }
}

struct X0Bin {
  int64_t aoff, boff;  // plane offsets (elements) of the bin's first K slice
  int ns, slice;       // K slices, output slice
};

__global__ __launch_bounds__(512, 1) void x0_bf3_kernel(const int* __restrict__ tiles, int ntiles,
                                                        const X0Bin* __restrict__ bins,
                                                        const unsigned short* __restrict__ as,
                                                        const unsigned short* __restrict__ bs,
                                                        const double* __restrict__ ones, int npr, int npr_col,
                                                        int MP, int NCP, int mirror, cf* __restrict__ rcxq,
                                                        int64_t slice_elems) {
  __shared__ SigSmem<128, 2> s;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int b = tiles[3 * t], ti = tiles[3 * t + 1], tj = tiles[3 * t + 2];
    const X0Bin bb = bins[b];
    __syncthreads();  // the previous tile's LDS reads are done before this tile's loads land
    sig_tile<128, 2, 8>(s, as + bb.aoff, bs + bb.boff, ones, npr, 1, 0, 1, npr_col, MP, bb.ns, NCP, ti * QBM,
                        tj * 128, 1, rcxq + slice_elems * bb.slice, mirror && ti != tj);
  }
}

// Scratch for the gathered planes of one group of bins (A and B planes):
// larger entry lists run as several groups, each accumulating into rcxq.
constexpr size_t X0_SCRATCH = (size_t)4 << 30;

ecalj_status x0_bf3(ecalj_ctx* h, const cf* z, int npr, int npr_col, int ipr_col, int nbin, const int64_t* cnt,
                    const int64_t* col, const double* wt, const int* slice, cf* rcxq, int64_t slice_elems,
                    int mirror) {
  const int MP = pad_to(std::max(npr, 1), QBM), NCP = pad_to(std::max(npr_col, 1), 128);
  const int RP = std::max(MP, NCP);
  const int R = MP / QBM, Ct = NCP / 128;
  // entries per group (multiple of 32): the group's planes fit X0_SCRATCH;
  // ECALJ_X0_GROUP_ENTRIES sets a smaller group (tests of the grouping)
  int64_t kcap = std::max<int64_t>(QBK, (int64_t)(X0_SCRATCH / (8 * (size_t)(MP + NCP))) / QBK * QBK);
  if (const char* e = std::getenv("ECALJ_X0_GROUP_ENTRIES")) {
    const long long v = std::atoll(e);
    if (v > 0) kcap = std::min<int64_t>(kcap, std::max<int64_t>(QBK, (int64_t)v / QBK * QBK));
  }
  // pieces: (bin, first entry, entries) with at most kcap entries; groups of consecutive pieces
  struct Piece { int bin; int64_t t0, k; };
  std::vector<std::vector<Piece>> groups(1);
  int64_t gk = 0;
  for (int b = 0; b < nbin; ++b) {
    const int64_t K = cnt[b + 1] - cnt[b];
    for (int64_t t0 = 0; t0 < K; t0 += kcap) {
      const int64_t k = std::min(kcap, K - t0), k32 = (k + QBK - 1) / QBK * QBK;
      if (gk + k32 > kcap) {
        groups.emplace_back();
        gk = 0;
      }
      groups.back().push_back({b, cnt[b] + t0, k});
      gk += k32;
    }
  }
  if (groups.back().empty()) groups.pop_back();
  if (groups.empty()) return ECALJ_OK;
  // blobs of every group first: the list slot and the planes are sized once (no growth between groups)
  struct Blob { std::vector<unsigned char> v; size_t o_w, o_one, o_bin, o_til; int64_t kp; int ntiles; };
  std::vector<Blob> blobs(groups.size());
  size_t maxblob = 0;
  int64_t maxkp = 0;
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const std::vector<Piece>& g = groups[gi];
    std::vector<X0Bin> bins;
    int64_t kp = 0;
    for (const Piece& pc : g) {
      const int64_t k32 = (pc.k + QBK - 1) / QBK * QBK;
      if (k32 / QBK > 0x7fffffff) return ECALJ_ERR_DIM;
      bins.push_back({kp / QBK * 4 * (int64_t)MP * QBK, kp / QBK * 4 * (int64_t)NCP * QBK, (int)(k32 / QBK),
                      slice[pc.bin]});
      kp += k32;
    }
    std::vector<int64_t> colp(kp, -1);
    std::vector<double> wp(kp, 0.0);
    int64_t o = 0;
    for (const Piece& pc : g) {
      for (int64_t t = 0; t < pc.k; ++t) {
        colp[o + t] = col[pc.t0 + t];
        wp[o + t] = wt[pc.t0 + t];
      }
      o += (pc.k + QBK - 1) / QBK * QBK;
    }
    // tiles: longest K first (persistent grid)
    std::vector<int> order(bins.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return bins[x].ns > bins[y].ns; });
    std::vector<int> tiles;
    for (int b : order)
      for (int ti = 0; ti < R; ++ti)
        for (int tj = mirror ? ti : 0; tj < Ct; ++tj) {
          tiles.push_back(b);
          tiles.push_back(ti);
          tiles.push_back(tj);
        }
    // blob: [colp | wp | ones (NCP) | bins | tiles]
    Blob& B = blobs[gi];
    B.kp = kp;
    B.ntiles = (int)(tiles.size() / 3);
    B.o_w = sizeof(int64_t) * kp;
    B.o_one = B.o_w + sizeof(double) * kp;
    B.o_bin = B.o_one + sizeof(double) * NCP;
    B.o_til = B.o_bin + sizeof(X0Bin) * bins.size();
    B.v.resize(B.o_til + sizeof(int) * tiles.size());
    memcpy(B.v.data(), colp.data(), B.o_w);
    memcpy(B.v.data() + B.o_w, wp.data(), B.o_one - B.o_w);
    for (int c = 0; c < NCP; ++c) {
      const double one = 1.0;
      memcpy(B.v.data() + B.o_one + sizeof(double) * c, &one, sizeof(double));
    }
    memcpy(B.v.data() + B.o_bin, bins.data(), B.o_til - B.o_bin);
    memcpy(B.v.data() + B.o_til, tiles.data(), sizeof(int) * tiles.size());
    maxblob = std::max(maxblob, B.v.size());
    maxkp = std::max(maxkp, kp);
  }
  unsigned char* d;
  unsigned short *as, *bs;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_LIST, maxblob, (void**)&d));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_X0A, sizeof(unsigned short) * 4 * (size_t)MP * maxkp, (void**)&as));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_X0B, sizeof(unsigned short) * 4 * (size_t)NCP * maxkp, (void**)&bs));
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 1, &tev));
  for (const Blob& B : blobs) {
    ECALJ_TRY(upload_blob(h, ecalj_ctx::S_LIST, B.v, &d));
    const int64_t nslice = B.kp / QBK;
    const int64_t tot = nslice * 4 * RP;
    const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 65536 * 4);
    hipLaunchKernelGGL(x0_split_kernel, dim3(blocks), dim3(256), 0, h->stream, z, (const int64_t*)d,
                       (const double*)(d + B.o_w), nslice, npr, npr_col, (int64_t)(ipr_col - 1), MP, NCP, RP, as,
                       bs);
    ECALJ_HIP(hipGetLastError());
    const int grid = std::max(1, std::min(B.ntiles, h->ncu));
    hipLaunchKernelGGL(x0_bf3_kernel, dim3(grid), dim3(512), 0, h->stream, (const int*)(d + B.o_til), B.ntiles,
                       (const X0Bin*)(d + B.o_bin), as, bs, (const double*)(d + B.o_one), npr, npr_col, MP, NCP,
                       mirror, rcxq, slice_elems);
    ECALJ_HIP(hipGetLastError());
  }
  ECALJ_TRY(timing_end(h, 1, tev));
  return ECALJ_OK;
}

}  // namespace ecalj
