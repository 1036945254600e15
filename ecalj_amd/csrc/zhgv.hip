// zhgv.hip -- the Hermitian-definite generalized eigenproblem A z = lambda B z
// in FP64 on the GPU: the Coulomb-matrix diagonalisation of hvccfp0
// (S/main_hvccfp0.f90:322-368) through zhgv_h / zhgv_d (S/m_lapack.f90:64-104,
// 203-229; LAPACK zhegvx, cuSOLVER Zhegvdx).
//
//   1. Hermitian copies of A and B from their upper triangles;
//   2. B = L L^H, blocked Cholesky: 32x32 diagonal blocks factored in LDS
//      (with their inverses), panels and trailing updates as FP64-MFMA GEMMs;
//   3. L^-1 by the library's unpivoted blocked Gauss-Jordan inverse (the
//      pivots are L's positive diagonal);
//   4. C = L^-1 A L^-H (two GEMMs), made exactly Hermitian;
//   5. C = V diag(lambda) V^H by cyclic two-sided Jacobi: in each round the
//      indices form m/2 disjoint pairs (round-robin tournament order), every
//      pair's 2x2 block is diagonalised by one complex rotation, the
//      rotations of a round are applied to C's columns and V's (one
//      workgroup per pair, contiguous columns) and then to C's rows (one
//      workgroup per column); a sweep is m-1 rounds, repeated until a sweep
//      applies no rotation;
//   6. z = L^-H V (GEMM), columns ordered by ascending lambda.
// Jacobi converges quadratically and handles the degenerate eigenvalues of
// the Coulomb matrix (shells of equal |q+G|) without special cases; the
// eigenvectors of a degenerate eigenvalue are some orthonormal basis of its
// space, as with any solver.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "ctx.hpp"
#include "kernels.hpp"

namespace ecalj {
namespace {

constexpr int CB = 32;  // Cholesky diagonal block

__device__ inline cd cmul(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline cd cmulc(cd a, cd b) { return cd{a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; }  // a conj(b)

// dst (n x n, ld n) = the Hermitian matrix whose upper triangle is src's (ld lds); real diagonal
__global__ void herm_upper_kernel(const cd* __restrict__ src, int64_t lds, int n, cd* __restrict__ dst) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    cd v;
    if (i < j) {
      v = src[i + lds * j];
    } else if (i > j) {
      const cd u = src[j + lds * i];
      v = cd{u.re, -u.im};
    } else {
      v = cd{src[i + lds * j].re, 0.0};
    }
    dst[e] = v;
  }
}

// c (n x n) <- (c + c^H) / 2
__global__ void hermitize_kernel(cd* __restrict__ c, int n) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i > j) continue;
    const cd a = c[e], b = c[j + (int64_t)n * i];
    if (i == j) {
      c[e] = cd{a.re, 0.0};
    } else {
      const cd m{0.5 * (a.re + b.re), 0.5 * (a.im - b.im)};
      c[e] = m;
      c[j + (int64_t)n * i] = cd{m.re, -m.im};
    }
  }
}

// dst (m x b, ld m) = src rows r0.., columns c0.. (ld n)
__global__ void copy_block_kernel(const cd* __restrict__ src, int n, int r0, int c0, int m, int b, cd* __restrict__ dst) {
  const int64_t tot = (int64_t)m * b;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % m), j = (int)(e / m);
    dst[e] = src[(r0 + i) + (int64_t)n * (c0 + j)];
  }
}

// One workgroup: the b x b diagonal block at k0 of a (ld n, lower triangle
// used) is replaced by its Cholesky factor L11 (strict upper zeroed) and
// inv(L11) is written to linv (ld CB).  A non-positive pivot sets *flag and
// continues with pivot 1 so that the kernel always finishes.
__global__ void __launch_bounds__(256) chol_diag_kernel(cd* __restrict__ a, int n, int k0, int b,
                                                        cd* __restrict__ linv, int* __restrict__ flag) {
  __shared__ cd L[CB][CB + 1];
  __shared__ cd X[CB][CB + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < b * b; e += blockDim.x) {
    const int r = e % b, c = e / b;
    L[r][c] = (r >= c) ? a[(k0 + r) + (int64_t)n * (k0 + c)] : cd{0.0, 0.0};
  }
  __syncthreads();
  for (int j = 0; j < b; ++j) {
    if (tid == 0) {
      double d = L[j][j].re;
      if (!(d > 0.0)) {
        atomicExch(flag, 1);
        d = 1.0;
      }
      L[j][j] = cd{sqrt(d), 0.0};
    }
    __syncthreads();
    const double inv = 1.0 / L[j][j].re;
    for (int i = j + 1 + tid; i < b; i += blockDim.x) L[i][j] = cd{L[i][j].re * inv, L[i][j].im * inv};
    __syncthreads();
    const int m = b - j - 1;
    for (int e = tid; e < m * m; e += blockDim.x) {
      const int i = j + 1 + e % m, k = j + 1 + e / m;
      if (k <= i) {
        const cd p = cmulc(L[i][j], L[k][j]);
        L[i][k] = cd{L[i][k].re - p.re, L[i][k].im - p.im};
      }
    }
    __syncthreads();
  }
  // inv(L11) by forward substitution, one column per thread
  if (tid < b) {
    const int c = tid;
    for (int i = 0; i < c; ++i) X[i][c] = cd{0.0, 0.0};
    X[c][c] = cd{1.0 / L[c][c].re, 0.0};
    for (int i = c + 1; i < b; ++i) {
      double sr = 0.0, si = 0.0;
      for (int k = c; k < i; ++k) {
        const cd p = cmul(L[i][k], X[k][c]);
        sr += p.re;
        si += p.im;
      }
      const double inv = -1.0 / L[i][i].re;
      X[i][c] = cd{sr * inv, si * inv};
    }
  }
  __syncthreads();
  for (int e = tid; e < b * b; e += blockDim.x) {
    const int r = e % b, c = e / b;
    a[(k0 + r) + (int64_t)n * (k0 + c)] = L[r][c];
    linv[r + CB * c] = X[r][c];
  }
}

// zero the strict upper triangle of a (n x n, ld n)
__global__ void zero_upper_kernel(cd* __restrict__ a, int n) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn; e += (int64_t)gridDim.x * blockDim.x)
    if ((int)(e % n) < (int)(e / n)) a[e] = cd{0.0, 0.0};
}

// Round-robin tournament order over m players (m even): in round r, pair i
// joins player(i) and player(m-1-i), where player 0 stays and players
// 1..m-1 rotate by r.  Every pair of indices meets once per m-1 rounds.
__device__ inline int rr_player(int j, int r, int m) { return j == 0 ? 0 : 1 + (j - 1 + r) % (m - 1); }

struct Rot {
  double c, sr, si;   // rotation [[c, s u], [-s conj(u), c]] with s u = sr + i si
  double app, aqq;    // diagonal after the rotation
  int p, q, on;
};

// One workgroup per pair of round r: the rotation that diagonalises the
// pair's 2x2 block of c, applied to c's and v's columns p and q.
__global__ void __launch_bounds__(256) jac_cols_kernel(cd* __restrict__ c, cd* __restrict__ v, int n, int m, int r,
                                                       double rel, double absv, Rot* __restrict__ rot,
                                                       int* __restrict__ count) {
  const int i = blockIdx.x;
  int p = rr_player(i, r, m), q = rr_player(m - 1 - i, r, m);
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
  __shared__ Rot R;
  if (threadIdx.x == 0) {
    Rot x;
    x.p = p;
    x.q = q;
    x.on = 0;
    x.c = 1.0;
    x.sr = x.si = 0.0;
    x.app = x.aqq = 0.0;
    if (q < n) {
      const double a = c[p + (int64_t)n * p].re, d = c[q + (int64_t)n * q].re;
      const cd b = c[p + (int64_t)n * q];
      const double ab = hypot(b.re, b.im);
      if (ab > absv && ab > rel * sqrt(fabs(a) * fabs(d))) {
        const double tau = (d - a) / (2.0 * ab);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double cc = 1.0 / sqrt(1.0 + t * t), s = t * cc;
        x.c = cc;
        x.sr = s * b.re / ab;
        x.si = s * b.im / ab;
        x.app = a - t * ab;
        x.aqq = d + t * ab;
        x.on = 1;
        atomicAdd(count, 1);
      }
    }
    rot[i] = x;
    R = x;
  }
  __syncthreads();
  if (!R.on) return;
  const double cc = R.c, sr = R.sr, si = R.si;
  // new_p = c x_p - conj(s u) x_q ; new_q = (s u) x_p + c x_q
  for (int row = threadIdx.x; row < n; row += blockDim.x) {
    cd* mats[2] = {c, v};
    for (int t = 0; t < 2; ++t) {
      cd* M = mats[t];
      const cd xp = M[row + (int64_t)n * p], xq = M[row + (int64_t)n * q];
      M[row + (int64_t)n * p] = cd{cc * xp.re - (sr * xq.re + si * xq.im), cc * xp.im - (sr * xq.im - si * xq.re)};
      M[row + (int64_t)n * q] = cd{(sr * xp.re - si * xp.im) + cc * xq.re, (sr * xp.im + si * xp.re) + cc * xq.im};
    }
  }
}

// The rotations of the round applied to c's rows: one workgroup per column
// j, one thread per pair.  The pair's own 2x2 block gets its exact values.
__global__ void __launch_bounds__(256) jac_rows_kernel(cd* __restrict__ c, int n, int npair,
                                                       const Rot* __restrict__ rot) {
  const int j = blockIdx.x;
  cd* col = c + (int64_t)n * j;
  for (int i = threadIdx.x; i < npair; i += blockDim.x) {
    const Rot R = rot[i];
    if (!R.on) continue;
    const int p = R.p, q = R.q;
    if (j == p) {
      col[p] = cd{R.app, 0.0};
      col[q] = cd{0.0, 0.0};
      continue;
    }
    if (j == q) {
      col[p] = cd{0.0, 0.0};
      col[q] = cd{R.aqq, 0.0};
      continue;
    }
    // new_p = c y_p - (s u) y_q ; new_q = conj(s u) y_p + c y_q
    const cd yp = col[p], yq = col[q];
    col[p] = cd{R.c * yp.re - (R.sr * yq.re - R.si * yq.im), R.c * yp.im - (R.sr * yq.im + R.si * yq.re)};
    col[q] = cd{(R.sr * yp.re + R.si * yp.im) + R.c * yq.re, (R.sr * yp.im - R.si * yp.re) + R.c * yq.im};
  }
}

__global__ void eye_kernel(cd* __restrict__ v, int n) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn; e += (int64_t)gridDim.x * blockDim.x)
    v[e] = cd{(e % n) == (e / n) ? 1.0 : 0.0, 0.0};
}

__global__ void diag_kernel(const cd* __restrict__ c, int n, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = c[i + (int64_t)n * i].re;
}

// a(:, j) (ld lda) = z(:, perm[j]) (ld n); evl[j] = d[perm[j]]
__global__ void sort_out_kernel(const cd* __restrict__ z, const double* __restrict__ d, const int* __restrict__ perm,
                                int n, cd* __restrict__ a, int64_t lda, double* __restrict__ evl) {
  const int j = blockIdx.y;
  const int pj = perm[j];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    a[i + lda * j] = z[i + (int64_t)n * pj];
  if (blockIdx.x == 0 && threadIdx.x == 0) evl[j] = d[pj];
}

__global__ void set_int_kernel(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

inline unsigned grid_for(int64_t tot) { return (unsigned)std::min<int64_t>((tot + 255) / 256, 65536); }

}  // namespace

// z (ld ldz) receives the B-orthonormal eigenvectors of all n eigenvalues,
// ascending, evl_dev the eigenvalues; a and b are read through their upper
// triangles and left unchanged.
ecalj_status zhgv_impl(ecalj_ctx* h, cd* z, int64_t ldz, const cd* a, int64_t lda, const cd* b, int64_t ldb, int n,
                       double* evl_dev) {
  if (n < 1 || lda < n || ldb < n || ldz < n) return ECALJ_ERR_DIM;
  if (n > 32768) return ECALJ_ERR_DIM;
  const int64_t nn = (int64_t)n * n;
  const int m = n + (n & 1);  // tournament players (one idle index when n is odd)
  const int npair = m / 2;
  cd *L, *C, *T, *V;
  Rot* rot;
  int* ctr;
  double* dg;
  int* perm;
  cd* linv11;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_EIG1, sizeof(cd) * nn, (void**)&L));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_EIG2, sizeof(cd) * nn, (void**)&C));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_EIG3, sizeof(cd) * nn, (void**)&T));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_EIG4, sizeof(cd) * nn, (void**)&V));
  const size_t small = sizeof(Rot) * npair + 256 + sizeof(double) * n + 256 + sizeof(int) * n + 256 +
                       sizeof(cd) * CB * CB + 256;
  char* sp;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_EIG5, small, (void**)&sp));
  auto carve = [&](size_t bytes) {
    char* r = sp;
    sp += (bytes + 255) / 256 * 256;
    return r;
  };
  rot = (Rot*)carve(sizeof(Rot) * npair);
  ctr = (int*)carve(sizeof(int));
  dg = (double*)carve(sizeof(double) * n);
  perm = (int*)carve(sizeof(int) * n);
  linv11 = (cd*)carve(sizeof(cd) * CB * CB);
  hipStream_t st = h->stream;
  // 1. Hermitian B and A from their upper triangles
  hipLaunchKernelGGL(herm_upper_kernel, dim3(grid_for(nn)), dim3(256), 0, st, b, ldb, n, L);
  hipLaunchKernelGGL(herm_upper_kernel, dim3(grid_for(nn)), dim3(256), 0, st, a, lda, n, C);
  ECALJ_HIP(hipGetLastError());
  // 2. B = L L^H
  for (int k0 = 0; k0 < n; k0 += CB) {
    const int bs = std::min(CB, n - k0);
    hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(256), 0, st, L, n, k0, bs, linv11, h->d_flag);
    ECALJ_HIP(hipGetLastError());
    const int k1 = k0 + bs, mrow = n - k1;
    if (mrow <= 0) break;
    hipLaunchKernelGGL(copy_block_kernel, dim3(grid_for((int64_t)mrow * bs)), dim3(256), 0, st, L, n, k1, k0, mrow,
                       bs, T);
    ECALJ_HIP(hipGetLastError());
    GemmCall pnl;  // L21 = A21 inv(L11)^H
    pnl.mode = ECALJ_PREC_F64;
    pnl.tb = true; pnl.cb = true;
    pnl.M = mrow; pnl.N = bs; pnl.K = bs;
    pnl.A = T; pnl.lda = mrow;
    pnl.B = linv11; pnl.ldb = CB;
    pnl.C = L + k1 + (int64_t)n * k0; pnl.ldc = n;
    ECALJ_TRY(gemm(h, pnl));
    GemmCall tr;  // A22 -= L21 L21^H
    tr.mode = ECALJ_PREC_F64;
    tr.tb = true; tr.cb = true;
    tr.M = mrow; tr.N = mrow; tr.K = bs;
    tr.A = L + k1 + (int64_t)n * k0; tr.lda = n;
    tr.B = L + k1 + (int64_t)n * k0; tr.ldb = n;
    tr.C = L + k1 + (int64_t)n * k1; tr.ldc = n;
    tr.alpha = cd{-1.0, 0.0}; tr.beta = cd{1.0, 0.0};
    ECALJ_TRY(gemm(h, tr));
  }
  {  // B must be positive definite (zhegvx: info > n)
    int f = 0;
    ECALJ_HIP(hipMemcpyAsync(&f, h->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
    ECALJ_HIP(hipStreamSynchronize(st));
    if (f) {
      hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(64), 0, st, h->d_flag, 0);
      ECALJ_HIP(hipStreamSynchronize(st));
      return ECALJ_ERR_NOT_PD;
    }
  }
  // 3. L^-1 (lower triangular; pivots = L's positive diagonal)
  hipLaunchKernelGGL(zero_upper_kernel, dim3(grid_for(nn)), dim3(256), 0, st, L, n);
  ECALJ_HIP(hipGetLastError());
  ECALJ_TRY(gj_inverse(h, L, n, n, nn, 1, nullptr));
  ECALJ_TRY(check_flag(h));
  hipLaunchKernelGGL(zero_upper_kernel, dim3(grid_for(nn)), dim3(256), 0, st, L, n);  // exact zeros above
  ECALJ_HIP(hipGetLastError());
  // 4. C = L^-1 A L^-H
  GemmCall g1;  // T = A L^-H
  g1.mode = ECALJ_PREC_F64;
  g1.tb = true; g1.cb = true;
  g1.M = n; g1.N = n; g1.K = n;
  g1.A = C; g1.lda = n; g1.B = L; g1.ldb = n; g1.C = T; g1.ldc = n;
  ECALJ_TRY(gemm(h, g1));
  GemmCall g2;  // C = L^-1 T
  g2.mode = ECALJ_PREC_F64;
  g2.M = n; g2.N = n; g2.K = n;
  g2.A = L; g2.lda = n; g2.B = T; g2.ldb = n; g2.C = C; g2.ldc = n;
  ECALJ_TRY(gemm(h, g2));
  hipLaunchKernelGGL(hermitize_kernel, dim3(grid_for(nn)), dim3(256), 0, st, C, n);
  hipLaunchKernelGGL(eye_kernel, dim3(grid_for(nn)), dim3(256), 0, st, V, n);
  ECALJ_HIP(hipGetLastError());
  // 5. Jacobi sweeps until one applies no rotation
  const double rel = 1e-16;
  double absv = 0.0;
  {
    hipLaunchKernelGGL(diag_kernel, dim3((n + 255) / 256), dim3(256), 0, st, C, n, dg);
    std::vector<double> d0(n);
    ECALJ_HIP(hipMemcpyAsync(d0.data(), dg, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    ECALJ_HIP(hipStreamSynchronize(st));
    double mx = 0.0;
    for (double x : d0) mx = std::max(mx, std::fabs(x));
    absv = 1e-300 + 1e-30 * mx;  // below this an off-diagonal element is zero for every purpose
  }
  const int max_sweeps = 60;
  bool converged = false;
  for (int sw = 0; sw < max_sweeps && !converged; ++sw) {
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(64), 0, st, ctr, 0);
    for (int r = 0; r < m - 1; ++r) {
      hipLaunchKernelGGL(jac_cols_kernel, dim3(npair), dim3(256), 0, st, C, V, n, m, r, rel, absv, rot, ctr);
      hipLaunchKernelGGL(jac_rows_kernel, dim3(n), dim3(256), 0, st, C, n, npair, (const Rot*)rot);
    }
    ECALJ_HIP(hipGetLastError());
    int cnt = 0;
    ECALJ_HIP(hipMemcpyAsync(&cnt, ctr, sizeof(int), hipMemcpyDeviceToHost, st));
    ECALJ_HIP(hipStreamSynchronize(st));
    converged = (cnt == 0);
  }
  if (!converged) return ECALJ_ERR_NOT_CONVERGED;
  // 6. z = L^-H V, ascending eigenvalues
  GemmCall g3;
  g3.mode = ECALJ_PREC_F64;
  g3.ta = true; g3.ca = true;
  g3.M = n; g3.N = n; g3.K = n;
  g3.A = L; g3.lda = n; g3.B = V; g3.ldb = n; g3.C = T; g3.ldc = n;
  ECALJ_TRY(gemm(h, g3));
  hipLaunchKernelGGL(diag_kernel, dim3((n + 255) / 256), dim3(256), 0, st, C, n, dg);
  std::vector<double> d(n);
  ECALJ_HIP(hipMemcpyAsync(d.data(), dg, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  ECALJ_HIP(hipStreamSynchronize(st));
  std::vector<int> pm(n);
  std::iota(pm.begin(), pm.end(), 0);
  std::stable_sort(pm.begin(), pm.end(), [&](int x, int y) { return d[x] < d[y]; });
  ECALJ_HIP(hipMemcpyAsync(perm, pm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(sort_out_kernel, dim3((n + 255) / 256, n), dim3(256), 0, st, T, dg, perm, n, z, ldz, evl_dev);
  ECALJ_HIP(hipGetLastError());
  ECALJ_HIP(hipStreamSynchronize(st));  // perm's host source lives on this frame
  return ECALJ_OK;
}

}  // namespace ecalj

using namespace ecalj;

extern "C" {

ecalj_status ecalj_zhgv_dev(ecalj_handle h, void* a, const void* b, int n, double* evl, int il, int iu, int lda,
                            int ldb) {
  if (!h || !a || !b || !evl) return ECALJ_ERR_ARG;
  if (n < 1 || il < 1 || iu < il || iu > n) return ECALJ_ERR_DIM;
  if (h->capturing) return ECALJ_ERR_STATE;  // sweeps until convergence: decisions on the host
  ECALJ_HIP(hipSetDevice(h->device));
  double* ev_all;
  cd* z_all;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_EIG6, sizeof(cd) * (size_t)n * n + sizeof(double) * n + 256,
                        (void**)&z_all));
  ev_all = (double*)(z_all + (size_t)n * n);
  ECALJ_TRY(zhgv_impl(h, z_all, n, (const cd*)a, lda, (const cd*)b, ldb, n, ev_all));
  // eigenpairs il..iu (1-based, ascending) into evl(1:m) and a(:, 1:m)
  const int m = iu - il + 1;
  ECALJ_HIP(hipMemcpy2DAsync(a, sizeof(cd) * (size_t)lda, z_all + (size_t)n * (il - 1), sizeof(cd) * (size_t)n,
                             sizeof(cd) * (size_t)n, m, hipMemcpyDeviceToDevice, h->stream));
  ECALJ_HIP(hipMemcpyAsync(evl, ev_all + (il - 1), sizeof(double) * m, hipMemcpyDeviceToDevice, h->stream));
  return ECALJ_OK;
}

ecalj_status ecalj_zhgv(ecalj_handle h, double* a, const double* b, int n, double* evl, int il, int iu, int lda,
                        int ldb) {
  if (!h || !a || !b || !evl) return ECALJ_ERR_ARG;
  if (n < 1 || lda < n || ldb < n || il < 1 || iu < il || iu > n) return ECALJ_ERR_DIM;
  if (h->capturing) return ECALJ_ERR_STATE;
  ECALJ_HIP(hipSetDevice(h->device));
  const size_t ba = sizeof(cd) * (size_t)lda * n, bb = sizeof(cd) * (size_t)ldb * n;
  DevBuf da, db, de;
  ECALJ_TRY(da.put(h, a, ba));
  ECALJ_TRY(db.put(h, b, bb));
  ECALJ_TRY(de.alloc(sizeof(double) * n));
  ECALJ_TRY(ecalj_zhgv_dev(h, da.p, db.p, n, (double*)de.p, il, iu, lda, ldb));
  ECALJ_HIP(hipMemcpyAsync(a, da.p, ba, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipMemcpyAsync(evl, de.p, sizeof(double) * (iu - il + 1), hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

}  // extern "C"
