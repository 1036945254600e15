/* FP64 CPU oracle of the learned chi0-tail MLP (test infrastructure only:
 * the product path is ecalj_amd/csrc/mlp.hip).
 *
 * The reference has no learned component; its only handle on the band tail
 * is the truncation of the chi0 band sum, nband_chi0 / emax_chi0
 * (S/m_readgwinput.f90:34-35, passed to tetwt5x_dtet4 at S/m_tetwt.f90:140-145).
 * The network predicts what that truncation drops.
 *
 * Network (L layers, widths dims[0..L]):
 *   z_l = h_l W_l^T + b_l,  h_0 = x,
 *   h_{l+1} = silu(z_l) = z_l / (1 + exp(-z_l))   for l < L-1,
 *   y = z_{L-1}.
 * params: for l = 0..L-1, W_l (dims[l+1] x dims[l], row-major) then b_l (dims[l+1]).
 * Batches are row-major: x is B x dims[0], y is B x dims[L].
 * Loss: mean squared error over the B x dims[L] outputs.
 * Pins: gradients against central finite differences of the loss
 * (tests/test_mlp_oracle.py), silu and its derivative against closed forms. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double silu(double z) { return z / (1.0 + exp(-z)); }
static double dsilu(double z) {
  const double s = 1.0 / (1.0 + exp(-z));
  return s * (1.0 + z * (1.0 - s));
}

int64_t orc_mlp_nparam(int L, const int* dims) {
  int64_t n = 0;
  for (int l = 0; l < L; ++l) n += (int64_t)dims[l + 1] * dims[l] + dims[l + 1];
  return n;
}

/* forward; keeps z_l and h_l of every layer in zs / hs when non-NULL
 * (each B x maxw, maxw = max(dims)) */
static void forward(int L, const int* dims, const double* p, const double* x, int B, double* y, double* zs,
                    double* hs, int maxw) {
  double* hin = (double*)malloc(sizeof(double) * (size_t)B * maxw);
  double* hout = (double*)malloc(sizeof(double) * (size_t)B * maxw);
  for (int i = 0; i < B; ++i)
    for (int c = 0; c < dims[0]; ++c) hin[(size_t)i * maxw + c] = x[(size_t)i * dims[0] + c];
  const double* q = p;
  for (int l = 0; l < L; ++l) {
    const int din = dims[l], dout = dims[l + 1];
    const double* W = q;
    const double* b = q + (size_t)dout * din;
    q = b + dout;
    if (hs)
      for (int i = 0; i < B; ++i)
        memcpy(hs + ((size_t)l * B + i) * maxw, hin + (size_t)i * maxw, sizeof(double) * din);
    for (int i = 0; i < B; ++i)
      for (int o = 0; o < dout; ++o) {
        double s = b[o];
        for (int c = 0; c < din; ++c) s += hin[(size_t)i * maxw + c] * W[(size_t)o * din + c];
        if (zs) zs[((size_t)l * B + i) * maxw + o] = s;
        hout[(size_t)i * maxw + o] = (l < L - 1) ? silu(s) : s;
      }
    double* t = hin;
    hin = hout;
    hout = t;
  }
  for (int i = 0; i < B; ++i)
    for (int o = 0; o < dims[L]; ++o) y[(size_t)i * dims[L] + o] = hin[(size_t)i * maxw + o];
  free(hin);
  free(hout);
}

static int maxdim(int L, const int* dims) {
  int m = 0;
  for (int l = 0; l <= L; ++l) m = dims[l] > m ? dims[l] : m;
  return m;
}

void orc_mlp_forward(int L, const int* dims, const double* p, const double* x, int B, double* y) {
  forward(L, dims, p, x, B, y, NULL, NULL, maxdim(L, dims));
}

double orc_mlp_loss_grad(int L, const int* dims, const double* p, const double* x, const double* t, int B,
                         double* g) {
  const int maxw = maxdim(L, dims), dL = dims[L];
  double* zs = (double*)malloc(sizeof(double) * (size_t)L * B * maxw);
  double* hs = (double*)malloc(sizeof(double) * (size_t)L * B * maxw);
  double* y = (double*)malloc(sizeof(double) * (size_t)B * dL);
  forward(L, dims, p, x, B, y, zs, hs, maxw);
  double loss = 0.0;
  const double nrm = 1.0 / ((double)B * dL);
  double* dz = (double*)malloc(sizeof(double) * (size_t)B * maxw);
  double* dh = (double*)malloc(sizeof(double) * (size_t)B * maxw);
  for (int i = 0; i < B; ++i)
    for (int o = 0; o < dL; ++o) {
      const double e = y[(size_t)i * dL + o] - t[(size_t)i * dL + o];
      loss += e * e * nrm;
      dz[(size_t)i * maxw + o] = 2.0 * e * nrm;
    }
  if (g) {
    /* parameter offsets */
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (L + 1));
    off[0] = 0;
    for (int l = 0; l < L; ++l) off[l + 1] = off[l] + (int64_t)dims[l + 1] * dims[l] + dims[l + 1];
    for (int l = L - 1; l >= 0; --l) {
      const int din = dims[l], dout = dims[l + 1];
      const double* W = p + off[l];
      double* gW = g + off[l];
      double* gb = gW + (size_t)dout * din;
      for (int o = 0; o < dout; ++o) {
        double sb = 0.0;
        for (int i = 0; i < B; ++i) sb += dz[(size_t)i * maxw + o];
        gb[o] = sb;
        for (int c = 0; c < din; ++c) {
          double s = 0.0;
          for (int i = 0; i < B; ++i) s += dz[(size_t)i * maxw + o] * hs[((size_t)l * B + i) * maxw + c];
          gW[(size_t)o * din + c] = s;
        }
      }
      if (l > 0) {
        for (int i = 0; i < B; ++i)
          for (int c = 0; c < din; ++c) {
            double s = 0.0;
            for (int o = 0; o < dout; ++o) s += dz[(size_t)i * maxw + o] * W[(size_t)o * din + c];
            dh[(size_t)i * maxw + c] = s * dsilu(zs[((size_t)(l - 1) * B + i) * maxw + c]);
          }
        double* tt = dz;
        dz = dh;
        dh = tt;
      }
    }
    free(off);
  }
  free(zs);
  free(hs);
  free(y);
  free(dz);
  free(dh);
  return loss;
}

double orc_silu(double z) { return silu(z); }
double orc_dsilu(double z) { return dsilu(z); }
