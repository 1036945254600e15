// mlp.hip -- the learned chi0-tail network: FP32 training kernels (forward,
// backward, Adam) on FP32 MFMA, bf16 MFMA inference, and the feature /
// reconstruction kernels that connect it to chi0.
//
// The reference has no learned component; its handle on the band tail is the
// truncation of the chi0 band sum (nband_chi0 / emax_chi0,
// S/m_readgwinput.f90:34-35, passed to tetwt5x_dtet4 at S/m_tetwt.f90:140-145).
// Here the sum is truncated the same way and the network predicts the part
// the truncation drops.
//
// Network (L layers, widths dims[0..L]), row-major batches:
//   z_l = h_l W_l^T + b_l,  h_0 = x,  h_{l+1} = silu(z_l) (l < L-1),  y = z_{L-1};
// params: per layer W_l (dims[l+1] x dims[l], row-major) then b_l.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace ecalj {

__device__ inline float silu_f(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float dsilu_f(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------- FP32 GEMM
// C (M x N, row-major ldc) = op(A) op(B):
//   A(i,k) = TA ? A[k*lda + i] : A[i*lda + k],  B(k,j) = TB ? B[j*ldb + k] : B[k*ldb + j].
enum MlpEpi { ME_STORE = 0, ME_BIAS = 1, ME_BIAS_SILU = 2, ME_DSILU = 3 };
constexpr int MT = 64, MK = 16;

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                       const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                       int ldc, const float* __restrict__ bias,
                                                       float* __restrict__ Zout, const float* __restrict__ Zin) {
  __shared__ float As[MK][MT + 1];  // As[k][i]
  __shared__ float Bs[MK][MT + 1];  // Bs[k][j]
  const int i0 = blockIdx.y * MT, j0 = blockIdx.x * MT;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, wm = wave >> 1, wn = wave & 1;
  f4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += MK) {
#pragma unroll
    for (int q = 0; q < MT * MK / 256; ++q) {
      const int e = t + 256 * q;
      int i, k;
      if (TA) {
        i = e % MT;
        k = e / MT;
      } else {
        k = e % MK;
        i = e / MK;
      }
      const int gi = i0 + i, gk = k0 + k;
      float v = 0.f;
      if (gi < M && gk < K) v = TA ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk];
      As[k][i] = v;
      int j, kb;
      if (TB) {
        kb = e % MK;
        j = e / MK;
      } else {
        j = e % MT;
        kb = e / MT;
      }
      const int gj = j0 + j, gkb = k0 + kb;
      float w = 0.f;
      if (gj < N && gkb < K) w = TB ? B[(int64_t)gj * ldb + gkb] : B[(int64_t)gkb * ldb + gj];
      Bs[kb][j] = w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < MK; kk += 4) {
      float av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[kk + (lane >> 4)][wm * 32 + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = Bs[kk + (lane >> 4)][wn * 32 + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wm * 32 + a * 16 + acc_row_f32(lane, v);
        const int j = j0 + wn * 32 + b * 16 + (lane & 15);
        if (i >= M || j >= N) continue;
        const int64_t o = (int64_t)i * ldc + j;
        float z = acc[a][b][v];
        if (EPI == ME_BIAS || EPI == ME_BIAS_SILU) z += bias[j];
        if (EPI == ME_BIAS_SILU) {
          Zout[o] = z;
          z = silu_f(z);
        }
        if (EPI == ME_DSILU) z *= dsilu_f(Zin[o]);
        C[o] = z;
      }
}

template <bool TA, bool TB, int EPI>
static ecalj_status mlp_gemm(ecalj_ctx* h, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                             float* C, int ldc, const float* bias = nullptr, float* Zout = nullptr,
                             const float* Zin = nullptr) {
  if (M <= 0 || N <= 0) return ECALJ_OK;
  dim3 grid((N + MT - 1) / MT, (M + MT - 1) / MT);
  if (grid.y > 65535) return ECALJ_ERR_DIM;
  hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, h->stream, M, N, K, A, lda, B, ldb, C, ldc,
                     bias, Zout, Zin);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

// ---------------------------------------------------------------- bf16 inference GEMM
// C (M x N) = X (M x K, fp32 -> bf16 on load) W^T (W: N x K bf16 row-major) + bias,
// silu when SILU.  NTERM 1: one bf16 product; NTERM 3: hi/lo split of both
// operands, hi*hi + hi*lo + lo*hi (W given as hi and lo planes).
constexpr int BQK = 32;
template <int NTERM, bool SILU>
__global__ __launch_bounds__(256) void mlp_infer_kernel(int M, int N, int K, const float* __restrict__ X, int ldx,
                                                        const unsigned short* __restrict__ Wh,
                                                        const unsigned short* __restrict__ Wl, int ldw,
                                                        const float* __restrict__ bias, float* __restrict__ C,
                                                        int ldc) {
  constexpr int LDK = BQK + 8;
  __shared__ alignas(16) unsigned short xh[MT * LDK], xl[MT * LDK], wh[MT * LDK], wl[MT * LDK];
  const int i0 = blockIdx.y * MT, j0 = blockIdx.x * MT;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, wm = wave >> 1, wn = wave & 1;
  f4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += BQK) {
#pragma unroll
    for (int q = 0; q < MT * BQK / 256; ++q) {
      const int e = t + 256 * q, k = e % BQK, r = e / BQK;
      const int gk = k0 + k;
      float x = 0.f;
      if (i0 + r < M && gk < K) x = X[(int64_t)(i0 + r) * ldx + gk];
      unsigned short hb, lb;
      split_bf16(x, hb, lb);
      xh[r * LDK + k] = hb;
      if (NTERM == 3) xl[r * LDK + k] = lb;
      unsigned short w0 = 0, w1 = 0;
      if (j0 + r < N && gk < K) {
        w0 = Wh[(int64_t)(j0 + r) * ldw + gk];
        if (NTERM == 3) w1 = Wl[(int64_t)(j0 + r) * ldw + gk];
      }
      wh[r * LDK + k] = w0;
      if (NTERM == 3) wl[r * LDK + k] = w1;
    }
    __syncthreads();
    const int ko = 8 * (lane >> 4);
    us8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int r = wm * 32 + a * 16 + (lane & 15);
      ah[a] = *reinterpret_cast<const us8*>(&xh[r * LDK + ko]);
      if (NTERM == 3) al[a] = *reinterpret_cast<const us8*>(&xl[r * LDK + ko]);
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int r = wn * 32 + b * 16 + (lane & 15);
      bh[b] = *reinterpret_cast<const us8*>(&wh[r * LDK + ko]);
      if (NTERM == 3) bl[b] = *reinterpret_cast<const us8*>(&wl[r * LDK + ko]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, ah[a]),
                                                            __builtin_bit_cast(bf8, bh[b]), acc[a][b], 0, 0, 0);
        if (NTERM == 3) {
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, ah[a]),
                                                              __builtin_bit_cast(bf8, bl[b]), acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, al[a]),
                                                              __builtin_bit_cast(bf8, bh[b]), acc[a][b], 0, 0, 0);
        }
      }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wm * 32 + a * 16 + acc_row_f32(lane, v);
        const int j = j0 + wn * 32 + b * 16 + (lane & 15);
        if (i >= M || j >= N) continue;
        float z = acc[a][b][v] + bias[j];
        if (SILU) z = silu_f(z);
        C[(int64_t)i * ldc + j] = z;
      }
}

// ---------------------------------------------------------------- small kernels
// dY = 2 (Y - T) / (B d); loss = mean (Y - T)^2, reduced in a fixed order (one block)
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ y, const float* __restrict__ tg,
                                                  int64_t n, float* __restrict__ dy, float* __restrict__ loss) {
  __shared__ double red[256];
  const double nrm = 1.0 / (double)n;
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += 256) {
    const float d = y[e] - tg[e];
    dy[e] = (float)(2.0 * d * nrm);
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) *loss = (float)(red[0] * nrm);
}

// db[j] = sum_i dz[i*ld + j], fixed order
__global__ void colsum_kernel(const float* __restrict__ dz, int rows, int cols, int ld, float* __restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (int i = 0; i < rows; ++i) s += dz[(int64_t)i * ld + j];
  db[j] = s;
}

__global__ void adam_kernel(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

__global__ void to_bf16_kernel(int64_t n, const float* __restrict__ p, unsigned short* __restrict__ hi,
                               unsigned short* __restrict__ lo) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    unsigned short h, l;
    split_bf16(p[e], h, l);
    hi[e] = h;
    if (lo) lo[e] = l;
  }
}

// ---------------------------------------------------------------- tail features / reconstruction
// Per basis function I of a truncated chi0 (ni x ni x nw, complex128, frequencies {0, nu_1..}):
//   d_j = -Re chi(I,I,j),  s = d_0;
//   feat(I, j-1) = d_j / s (j = 1..nw-1),  feat(I, nw-1) = log(s);  standardised (f - mu) / sd.
//   aux(I) = {s, d_ref} (FP64), d_ref = -Re chi(I,I,iref).
// A row with s <= 0 or d_ref <= 0 (no truncated spectral weight) gets zero features and s = 0,
// which turns its tail off.
__global__ void tail_features_kernel(const cd* __restrict__ chi, int ni, int nw, int iref,
                                     const float* __restrict__ mu, const float* __restrict__ sd,
                                     float* __restrict__ feat, double* __restrict__ aux) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= ni) return;
  const int64_t nn = (int64_t)ni * ni, dg = (int64_t)I * (ni + 1);
  const double s = -chi[dg].re, dref = -chi[dg + nn * iref].re;
  const bool ok = s > 0.0 && dref > 0.0;
  for (int j = 1; j <= nw; ++j) {
    double f = 0.0;
    if (ok) f = (j < nw) ? -chi[dg + nn * j].re / s : log(s);
    const int c = j - 1;
    feat[(int64_t)I * nw + c] = ok ? (float)((f - mu[c]) / sd[c]) : 0.f;
  }
  aux[2 * I] = ok ? s : 0.0;
  aux[2 * I + 1] = ok ? dref : 1.0;
}

// tail(I, j) = max(0, s_I (mu_y[j] + sd_y[j] pred(I, j)))     (FP64)
__global__ void tail_scale_kernel(const float* __restrict__ pred, int ni, int nw, const float* __restrict__ mu,
                                  const float* __restrict__ sd, const double* __restrict__ aux,
                                  double* __restrict__ tail) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)ni * nw) return;
  const int I = (int)(e / nw), j = (int)(e % nw);
  const double r = (double)mu[j] + (double)sd[j] * (double)pred[e];
  tail[(int64_t)j * ni + I] = fmax(0.0, aux[2 * I] * r);
}

// chi_out(I,J,j) = chi(I,J,j) + rho(I,J) sqrt(tail(I,j) tail(J,j)),
// rho(I,J) = chi(I,J,iref) / sqrt(d_ref(I) d_ref(J)): a Hermitian negative
// semidefinite correction with the truncated sum's correlation structure at
// the reference frequency and the predicted diagonal.  In place allowed.
__global__ void tail_apply_kernel(const cd* __restrict__ chi, int ni, int nw, int iref,
                                  const double* __restrict__ aux, const double* __restrict__ tail,
                                  cd* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)ni * ni;
  if (e >= nn) return;
  const int I = (int)(e % ni), J = (int)(e / ni);
  const cd r = chi[e + nn * iref];
  const double q = 1.0 / sqrt(aux[2 * I + 1] * aux[2 * J + 1]);
  const double rr = r.re * q, ri = r.im * q;
  for (int j = 0; j < nw; ++j) {
    const double a = sqrt(tail[(int64_t)j * ni + I] * tail[(int64_t)j * ni + J]);
    const cd c = chi[e + nn * j];
    out[e + nn * j] = cd{c.re + rr * a, c.im + ri * a};
  }
}

}  // namespace ecalj

using namespace ecalj;

// ---------------------------------------------------------------- C-ABI
namespace {

struct NetShape {
  int L = 0;
  std::vector<int> d;
  std::vector<int64_t> off;  // parameter offset of layer l
  int64_t nparam = 0;
  int maxw = 0;
};

ecalj_status net_shape(int nlayer, const int* dims, NetShape* s) {
  if (nlayer < 1 || nlayer > 16 || !dims) return ECALJ_ERR_ARG;
  s->L = nlayer;
  s->d.assign(dims, dims + nlayer + 1);
  s->off.resize(nlayer + 1);
  s->off[0] = 0;
  s->maxw = 0;
  for (int l = 0; l <= nlayer; ++l) {
    if (s->d[l] < 1 || s->d[l] > 65536) return ECALJ_ERR_DIM;
    s->maxw = std::max(s->maxw, s->d[l]);
  }
  for (int l = 0; l < nlayer; ++l) s->off[l + 1] = s->off[l] + (int64_t)s->d[l + 1] * s->d[l] + s->d[l + 1];
  s->nparam = s->off[nlayer];
  return ECALJ_OK;
}

}  // namespace

extern "C" {

ecalj_status ecalj_mlp_nparam(int nlayer, const int* dims, int64_t* nparam) {
  NetShape s;
  ECALJ_TRY(net_shape(nlayer, dims, &s));
  if (!nparam) return ECALJ_ERR_ARG;
  *nparam = s.nparam;
  return ECALJ_OK;
}

// workspace: Z_l and H_{l+1} per layer (B x d[l+1] floats each), then dZ ping-pong (2 x B x maxw)
static ecalj_status mlp_work(ecalj_ctx* h, const NetShape& s, int batch, float** w) {
  int64_t n = 0;
  for (int l = 0; l < s.L; ++l) n += 2LL * batch * s.d[l + 1];
  n += 2LL * batch * s.maxw;
  return slot_ensure(h, ecalj_ctx::S_MLP, sizeof(float) * (size_t)n, (void**)w);
}

static ecalj_status mlp_forward_ws(ecalj_ctx* h, const NetShape& s, const float* p, const float* x, int B,
                                   float* ws, float* y, std::vector<const float*>* Hs, std::vector<float*>* Zs) {
  const float* hin = x;
  float* cur = ws;
  for (int l = 0; l < s.L; ++l) {
    const int din = s.d[l], dout = s.d[l + 1];
    const float* W = p + s.off[l];
    const float* b = W + (int64_t)dout * din;
    float* Z = cur;
    float* H = cur + (int64_t)B * dout;
    cur += 2LL * B * dout;
    if (Hs) Hs->push_back(hin);
    if (Zs) Zs->push_back(Z);
    if (l < s.L - 1) {
      ECALJ_TRY((mlp_gemm<false, true, ME_BIAS_SILU>(h, B, dout, din, hin, din, W, din, H, dout, b, Z)));
      hin = H;
    } else {
      ECALJ_TRY((mlp_gemm<false, true, ME_BIAS>(h, B, dout, din, hin, din, W, din, y ? y : Z, dout, b)));
    }
  }
  return ECALJ_OK;
}

ecalj_status ecalj_mlp_forward_dev(ecalj_handle h, int nlayer, const int* dims, const void* params, const void* x,
                                   int batch, void* y) {
  if (!h || !params || !x || !y) return ECALJ_ERR_ARG;
  if (batch < 0) return ECALJ_ERR_DIM;
  if (batch == 0) return ECALJ_OK;
  NetShape s;
  ECALJ_TRY(net_shape(nlayer, dims, &s));
  float* ws;
  ECALJ_TRY(mlp_work(h, s, batch, &ws));
  return mlp_forward_ws(h, s, (const float*)params, (const float*)x, batch, ws, (float*)y, nullptr, nullptr);
}

ecalj_status ecalj_mlp_grad_dev(ecalj_handle h, int nlayer, const int* dims, const void* params, const void* x,
                                const void* target, int batch, void* grad, void* loss) {
  if (!h || !params || !x || !target || !grad) return ECALJ_ERR_ARG;
  if (batch < 1) return ECALJ_ERR_DIM;
  NetShape s;
  ECALJ_TRY(net_shape(nlayer, dims, &s));
  float* ws;
  ECALJ_TRY(mlp_work(h, s, batch, &ws));
  const int B = batch;
  const float* p = (const float*)params;
  float* g = (float*)grad;
  std::vector<const float*> Hs;
  std::vector<float*> Zs;
  ECALJ_TRY(mlp_forward_ws(h, s, p, (const float*)x, B, ws, nullptr, &Hs, &Zs));
  int64_t used = 0;
  for (int l = 0; l < s.L; ++l) used += 2LL * B * s.d[l + 1];
  float* dz = ws + used;
  float* dz2 = dz + (int64_t)B * s.maxw;
  const int dL = s.d[s.L];
  hipLaunchKernelGGL(mse_kernel, dim3(1), dim3(256), 0, h->stream, Zs[s.L - 1], (const float*)target,
                     (int64_t)B * dL, dz, (float*)loss);
  ECALJ_HIP(hipGetLastError());
  for (int l = s.L - 1; l >= 0; --l) {
    const int din = s.d[l], dout = s.d[l + 1];
    const float* W = p + s.off[l];
    float* gW = g + s.off[l];
    float* gb = gW + (int64_t)dout * din;
    // dW = dZ^T H  (dout x din)
    ECALJ_TRY((mlp_gemm<true, false, ME_STORE>(h, dout, din, B, dz, dout, Hs[l], din, gW, din)));
    hipLaunchKernelGGL(colsum_kernel, dim3((dout + 127) / 128), dim3(128), 0, h->stream, dz, B, dout, dout, gb);
    ECALJ_HIP(hipGetLastError());
    if (l > 0) {  // dZ_{l-1} = (dZ W) * silu'(Z_{l-1})
      ECALJ_TRY((mlp_gemm<false, false, ME_DSILU>(h, B, din, dout, dz, dout, W, din, dz2, din, nullptr, nullptr,
                                                  Zs[l - 1])));
      std::swap(dz, dz2);
    }
  }
  return ECALJ_OK;
}

ecalj_status ecalj_mlp_adam_dev(ecalj_handle h, int64_t nparam, void* params, const void* grad, void* m, void* v,
                                int step, double lr, double beta1, double beta2, double eps) {
  if (!h || !params || !grad || !m || !v) return ECALJ_ERR_ARG;
  if (nparam < 0 || step < 1) return ECALJ_ERR_DIM;
  const double c1 = 1.0 / (1.0 - std::pow(beta1, step)), c2 = 1.0 / (1.0 - std::pow(beta2, step));
  const int blocks = (int)std::min<int64_t>((nparam + 255) / 256, 4096);
  if (nparam == 0) return ECALJ_OK;
  hipLaunchKernelGGL(adam_kernel, dim3(blocks), dim3(256), 0, h->stream, nparam, (float*)params,
                     (const float*)grad, (float*)m, (float*)v, (float)lr, (float)beta1, (float)beta2, (float)eps,
                     (float)c1, (float)c2);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status ecalj_mlp_to_bf16_dev(ecalj_handle h, int64_t nparam, const void* params, void* hi, void* lo) {
  if (!h || !params || !hi) return ECALJ_ERR_ARG;
  if (nparam < 0) return ECALJ_ERR_DIM;
  if (nparam == 0) return ECALJ_OK;
  const int blocks = (int)std::min<int64_t>((nparam + 255) / 256, 4096);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks), dim3(256), 0, h->stream, nparam, (const float*)params,
                     (unsigned short*)hi, (unsigned short*)lo);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status ecalj_mlp_infer_bf16_dev(ecalj_handle h, int nlayer, const int* dims, const void* params,
                                      const void* w_hi, const void* w_lo, const void* x, int batch, void* y) {
  if (!h || !params || !w_hi || !x || !y) return ECALJ_ERR_ARG;
  if (batch < 0) return ECALJ_ERR_DIM;
  if (batch == 0) return ECALJ_OK;
  NetShape s;
  ECALJ_TRY(net_shape(nlayer, dims, &s));
  float* ws;
  ECALJ_TRY(mlp_work(h, s, batch, &ws));
  const float* p = (const float*)params;
  const unsigned short* wh = (const unsigned short*)w_hi;
  const unsigned short* wl = (const unsigned short*)w_lo;
  const float* hin = (const float*)x;
  float* bufs[2] = {ws, ws + (int64_t)batch * s.maxw};
  for (int l = 0; l < s.L; ++l) {
    const int din = s.d[l], dout = s.d[l + 1];
    const float* b = p + s.off[l] + (int64_t)dout * din;
    float* out = (l == s.L - 1) ? (float*)y : bufs[l & 1];
    dim3 grid((dout + MT - 1) / MT, (batch + MT - 1) / MT);
    if (grid.y > 65535) return ECALJ_ERR_DIM;
    const bool silu = l < s.L - 1;
    if (wl) {
      if (silu)
        hipLaunchKernelGGL((mlp_infer_kernel<3, true>), grid, dim3(256), 0, h->stream, batch, dout, din, hin, din,
                           wh + s.off[l], wl + s.off[l], din, b, out, dout);
      else
        hipLaunchKernelGGL((mlp_infer_kernel<3, false>), grid, dim3(256), 0, h->stream, batch, dout, din, hin, din,
                           wh + s.off[l], wl + s.off[l], din, b, out, dout);
    } else {
      if (silu)
        hipLaunchKernelGGL((mlp_infer_kernel<1, true>), grid, dim3(256), 0, h->stream, batch, dout, din, hin, din,
                           wh + s.off[l], nullptr, din, b, out, dout);
      else
        hipLaunchKernelGGL((mlp_infer_kernel<1, false>), grid, dim3(256), 0, h->stream, batch, dout, din, hin, din,
                           wh + s.off[l], nullptr, din, b, out, dout);
    }
    ECALJ_HIP(hipGetLastError());
    hin = out;
  }
  return ECALJ_OK;
}

ecalj_status ecalj_tail_features_dev(ecalj_handle h, const void* chi, int ni, int nw, int iref, const void* mu,
                                     const void* sd, void* feat, void* aux) {
  if (!h || !chi || !mu || !sd || !feat || !aux) return ECALJ_ERR_ARG;
  if (ni < 1 || nw < 2 || iref < 0 || iref >= nw) return ECALJ_ERR_DIM;
  hipLaunchKernelGGL(tail_features_kernel, dim3((ni + 127) / 128), dim3(128), 0, h->stream, (const cd*)chi, ni, nw,
                     iref, (const float*)mu, (const float*)sd, (float*)feat, (double*)aux);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

ecalj_status ecalj_tail_apply_dev(ecalj_handle h, const void* chi, int ni, int nw, int iref, const void* pred,
                                  const void* mu_y, const void* sd_y, const void* aux, void* chi_out) {
  if (!h || !chi || !pred || !mu_y || !sd_y || !aux || !chi_out) return ECALJ_ERR_ARG;
  if (ni < 1 || nw < 1 || iref < 0 || iref >= nw) return ECALJ_ERR_DIM;
  double* tail;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_TAIL, sizeof(double) * (size_t)ni * nw, (void**)&tail));
  const int64_t n1 = (int64_t)ni * nw;
  hipLaunchKernelGGL(tail_scale_kernel, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, h->stream,
                     (const float*)pred, ni, nw, (const float*)mu_y, (const float*)sd_y, (const double*)aux, tail);
  ECALJ_HIP(hipGetLastError());
  const int64_t nn = (int64_t)ni * ni;
  hipLaunchKernelGGL(tail_apply_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, h->stream,
                     (const cd*)chi, ni, nw, iref, (const double*)aux, tail, (cd*)chi_out);
  ECALJ_HIP(hipGetLastError());
  return ECALJ_OK;
}

}  // extern "C"
