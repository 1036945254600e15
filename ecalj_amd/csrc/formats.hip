// formats.hip -- the on-disk records the unmodified Fortran steps exchange
// with the chi0 -> W -> Sigma_c path, read and written by the library so
// that a GPU step can sit between two Fortran steps:
//   __Vcoud.iq   sequential unformatted: ngb; q(3); vcoud(ngb); zcousq(ngb,ngb)
//                (written by hvccfp0, S/main_hvccfp0.f90:396-401; read by
//                Readvcoud, S/m_readVcoud.f90:36-43)
//   __WVI.iq / __WVR.iq  fixed-length records of W_c(nblochpmx, nblochpmx),
//                complex(8), record rec at byte (rec-1)*mrecl with mrecl =
//                16*nblochpmx**2 (MPI-IO, S/m_mpiio.f90:12-74; mrecl
//                S/rdpp.f90:43; records iw (WVI) and iw-nw_i+1 (WVR),
//                S/m_llw.f90:177,373); in the reference's mixed-precision
//                build (__MP, S/m_kind.f90:3-6) the records are complex(4)
//                with mrecl = 8*nblochpmx**2 (kp = 4 in S/rdpp.f90:43)
//   __WV.d       one formatted line: nprecx mrecl nblochpmx nw+1 niw nq nw_i
//                (S/main_hrcxq.f90:134-136)
//   SEC2U        sequential unformatted: nspin nq ntq nqbz nqibz n1 n2 n3; per
//                q: is q(3) zsec(ntq,ntq) (S/main_hsfp0.sc.f90:339-359)
// Sequential unformatted records carry a 4-byte length before and after the
// payload (the record format of gfortran, flang and ifort).  Records of W
// are moved between the file and device memory through pinned staging.
#include <cstdio>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "ecalj_gpu.h"

namespace {

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(fopen(path, mode)) {}
  ~File() {
    if (f) fclose(f);
  }
};

bool put_rec(FILE* f, const std::vector<std::pair<const void*, size_t>>& parts) {
  size_t n = 0;
  for (auto& p : parts) n += p.second;
  if (n > 0x7fffffffu) return false;  // one 4-byte-marked record
  const int32_t m = (int32_t)n;
  if (fwrite(&m, 4, 1, f) != 1) return false;
  for (auto& p : parts)
    if (p.second && fwrite(p.first, 1, p.second, f) != p.second) return false;
  return fwrite(&m, 4, 1, f) == 1;
}

// reads one record; returns its payload length or -1
long get_rec(FILE* f, std::vector<unsigned char>& buf) {
  int32_t m1 = 0, m2 = 0;
  if (fread(&m1, 4, 1, f) != 1 || m1 < 0) return -1;
  buf.resize((size_t)m1);
  if (m1 && fread(buf.data(), 1, (size_t)m1, f) != (size_t)m1) return -1;
  if (fread(&m2, 4, 1, f) != 1 || m2 != m1) return -1;
  return m1;
}

}  // namespace

extern "C" ecalj_status ecalj_vcoud_write(const char* path, int ngb, const double* q, const double* vcoud,
                                          const double* zcousq) {
  if (!path || !q || !vcoud || !zcousq || ngb < 1) return ECALJ_ERR_ARG;
  File fl(path, "wb");
  if (!fl.f) return ECALJ_ERR_ARG;
  const size_t nz = sizeof(double) * 2 * (size_t)ngb * ngb;
  if (!put_rec(fl.f, {{&ngb, 4}}) || !put_rec(fl.f, {{q, 24}}) || !put_rec(fl.f, {{vcoud, sizeof(double) * ngb}}) ||
      !put_rec(fl.f, {{zcousq, nz}}))
    return ECALJ_ERR_ARG;
  return ECALJ_OK;
}

extern "C" ecalj_status ecalj_vcoud_read(const char* path, int* ngb, double* q, double* vcoud, double* zcousq,
                                         int ngb_cap) {
  if (!path || !ngb) return ECALJ_ERR_ARG;
  File fl(path, "rb");
  if (!fl.f) return ECALJ_ERR_ARG;
  std::vector<unsigned char> b;
  if (get_rec(fl.f, b) != 4) return ECALJ_ERR_DIM;
  int n;
  memcpy(&n, b.data(), 4);
  *ngb = n;
  if (n < 1) return ECALJ_ERR_DIM;
  if (n > ngb_cap || !q || !vcoud || !zcousq) return ECALJ_ERR_DIM;  // size reported; call again
  if (get_rec(fl.f, b) != 24) return ECALJ_ERR_DIM;
  memcpy(q, b.data(), 24);
  if (get_rec(fl.f, b) != (long)(sizeof(double) * n)) return ECALJ_ERR_DIM;
  memcpy(vcoud, b.data(), sizeof(double) * n);
  const size_t nz = sizeof(double) * 2 * (size_t)n * n;
  if (get_rec(fl.f, b) != (long)nz) return ECALJ_ERR_DIM;
  memcpy(zcousq, b.data(), nz);
  return ECALJ_OK;
}

extern "C" ecalj_status ecalj_sec2_write(const char* path, const int* header8, int nq, int ntq, const int* is,
                                         const double* q, const double* zsec) {
  if (!path || !header8 || !is || !q || !zsec || nq < 0 || ntq < 1) return ECALJ_ERR_ARG;
  File fl(path, "wb");
  if (!fl.f) return ECALJ_ERR_ARG;
  if (!put_rec(fl.f, {{header8, 32}})) return ECALJ_ERR_ARG;
  const size_t nz = sizeof(double) * 2 * (size_t)ntq * ntq;
  for (int ip = 0; ip < nq; ++ip)
    if (!put_rec(fl.f, {{is + ip, 4}, {q + 3 * (size_t)ip, 24}, {zsec + 2 * (size_t)ntq * ntq * ip, nz}}))
      return ECALJ_ERR_ARG;
  return ECALJ_OK;
}

// header8 = nspin nq ntq nqbz nqibz n1 n2 n3; then up to nq_cap records
extern "C" ecalj_status ecalj_sec2_read(const char* path, int* header8, int nq_cap, int ntq_cap, int* is, double* q,
                                        double* zsec) {
  if (!path || !header8) return ECALJ_ERR_ARG;
  File fl(path, "rb");
  if (!fl.f) return ECALJ_ERR_ARG;
  std::vector<unsigned char> b;
  if (get_rec(fl.f, b) != 32) return ECALJ_ERR_DIM;
  memcpy(header8, b.data(), 32);
  const int nq = header8[1], ntq = header8[2];
  if (nq < 0 || ntq < 1) return ECALJ_ERR_DIM;
  if (nq > nq_cap || ntq > ntq_cap || !is || !q || !zsec) return ECALJ_ERR_DIM;
  const size_t nz = sizeof(double) * 2 * (size_t)ntq * ntq;
  for (int ip = 0; ip < nq; ++ip) {
    if (get_rec(fl.f, b) != (long)(28 + nz)) return ECALJ_ERR_DIM;
    memcpy(is + ip, b.data(), 4);
    memcpy(q + 3 * (size_t)ip, b.data() + 4, 24);
    memcpy(zsec + 2 * (size_t)ntq * ntq * ip, b.data() + 28, nz);
  }
  return ECALJ_OK;
}

extern "C" ecalj_status ecalj_wv_d_write(const char* path, const int64_t* v7) {
  if (!path || !v7) return ECALJ_ERR_ARG;
  File fl(path, "w");
  if (!fl.f) return ECALJ_ERR_ARG;
  fprintf(fl.f, " ");  // format (1x,10i14)
  for (int i = 0; i < 7; ++i) fprintf(fl.f, "%14lld", (long long)v7[i]);
  fprintf(fl.f, "\n");
  return ECALJ_OK;
}

extern "C" ecalj_status ecalj_wv_d_read(const char* path, int64_t* v7) {
  if (!path || !v7) return ECALJ_ERR_ARG;
  File fl(path, "r");
  if (!fl.f) return ECALJ_ERR_ARG;
  for (int i = 0; i < 7; ++i) {
    long long x;
    if (fscanf(fl.f, "%lld", &x) != 1) return ECALJ_ERR_DIM;
    v7[i] = x;
  }
  return ECALJ_OK;
}

// nrec records of W (ld x ld complex128 each, contiguous in device memory)
// at records rec0..rec0+nrec-1 of a fixed-length file (record length 16*ld^2)
extern "C" ecalj_status ecalj_wv_write_dev(ecalj_handle h, const char* path, int rec0, int nrec, int ld,
                                           const void* w) {
  if (!h || !path || !w || rec0 < 1 || nrec < 0 || ld < 1) return ECALJ_ERR_ARG;
  if (h->capturing) return ECALJ_ERR_STATE;
  const size_t rl = 16 * (size_t)ld * ld;
  FILE* f = fopen(path, "r+b");
  if (!f) f = fopen(path, "w+b");
  if (!f) return ECALJ_ERR_ARG;
  void* pin = nullptr;
  ecalj_status st = ECALJ_OK;
  if (hipHostMalloc(&pin, rl) != hipSuccess) st = ECALJ_ERR_NOMEM;
  for (int r = 0; r < nrec && st == ECALJ_OK; ++r) {
    if (hipMemcpyAsync(pin, (const char*)w + rl * r, rl, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
      st = ECALJ_ERR_HIP;
      break;
    }
    if (fseeko(f, (off_t)rl * (rec0 - 1 + r), SEEK_SET) != 0 || fwrite(pin, 1, rl, f) != rl) st = ECALJ_ERR_ARG;
  }
  if (pin) (void)hipHostFree(pin);
  fclose(f);
  return st;
}

extern "C" ecalj_status ecalj_wv_read_dev(ecalj_handle h, const char* path, int rec0, int nrec, int ld, void* w) {
  if (!h || !path || !w || rec0 < 1 || nrec < 0 || ld < 1) return ECALJ_ERR_ARG;
  if (h->capturing) return ECALJ_ERR_STATE;
  const size_t rl = 16 * (size_t)ld * ld;
  File fl(path, "rb");
  if (!fl.f) return ECALJ_ERR_ARG;
  void* pin = nullptr;
  ecalj_status st = ECALJ_OK;
  if (hipHostMalloc(&pin, rl) != hipSuccess) return ECALJ_ERR_NOMEM;
  for (int r = 0; r < nrec && st == ECALJ_OK; ++r) {
    if (fseeko(fl.f, (off_t)rl * (rec0 - 1 + r), SEEK_SET) != 0 || fread(pin, 1, rl, fl.f) != rl) {
      st = ECALJ_ERR_DIM;  // short file
      break;
    }
    if (hipMemcpyAsync((char*)w + rl * r, pin, rl, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      st = ECALJ_ERR_HIP;
  }
  (void)hipHostFree(pin);
  return st;
}

// The same records in the complex(4) form of the mixed-precision build
// (record length 8*ld^2): device W stays complex128; each element is rounded
// to complex(4) on the way out (round to nearest, as cmplx(w, kind=4)) and
// widened exactly on the way in.  The conversion runs on the pinned staging.
extern "C" ecalj_status ecalj_wv4_write_dev(ecalj_handle h, const char* path, int rec0, int nrec, int ld,
                                            const void* w) {
  if (!h || !path || !w || rec0 < 1 || nrec < 0 || ld < 1) return ECALJ_ERR_ARG;
  if (h->capturing) return ECALJ_ERR_STATE;
  const size_t ne = (size_t)ld * ld, rl8 = 16 * ne, rl4 = 8 * ne;
  FILE* f = fopen(path, "r+b");
  if (!f) f = fopen(path, "w+b");
  if (!f) return ECALJ_ERR_ARG;
  void* pin = nullptr;
  ecalj_status st = ECALJ_OK;
  if (hipHostMalloc(&pin, rl8) != hipSuccess) st = ECALJ_ERR_NOMEM;
  std::vector<float> out(2 * ne);
  for (int r = 0; r < nrec && st == ECALJ_OK; ++r) {
    if (hipMemcpyAsync(pin, (const char*)w + rl8 * r, rl8, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
      st = ECALJ_ERR_HIP;
      break;
    }
    const double* d = static_cast<const double*>(pin);
    for (size_t i = 0; i < 2 * ne; ++i) out[i] = (float)d[i];
    if (fseeko(f, (off_t)rl4 * (rec0 - 1 + r), SEEK_SET) != 0 || fwrite(out.data(), 1, rl4, f) != rl4)
      st = ECALJ_ERR_ARG;
  }
  if (pin) (void)hipHostFree(pin);
  fclose(f);
  return st;
}

extern "C" ecalj_status ecalj_wv4_read_dev(ecalj_handle h, const char* path, int rec0, int nrec, int ld, void* w) {
  if (!h || !path || !w || rec0 < 1 || nrec < 0 || ld < 1) return ECALJ_ERR_ARG;
  if (h->capturing) return ECALJ_ERR_STATE;
  const size_t ne = (size_t)ld * ld, rl8 = 16 * ne, rl4 = 8 * ne;
  File fl(path, "rb");
  if (!fl.f) return ECALJ_ERR_ARG;
  void* pin = nullptr;
  if (hipHostMalloc(&pin, rl8) != hipSuccess) return ECALJ_ERR_NOMEM;
  std::vector<float> in(2 * ne);
  ecalj_status st = ECALJ_OK;
  for (int r = 0; r < nrec && st == ECALJ_OK; ++r) {
    if (fseeko(fl.f, (off_t)rl4 * (rec0 - 1 + r), SEEK_SET) != 0 || fread(in.data(), 1, rl4, fl.f) != rl4) {
      st = ECALJ_ERR_DIM;  // short file
      break;
    }
    double* d = static_cast<double*>(pin);
    for (size_t i = 0; i < 2 * ne; ++i) d[i] = in[i];
    if (hipMemcpyAsync((char*)w + rl8 * r, pin, rl8, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      st = ECALJ_ERR_HIP;
  }
  (void)hipHostFree(pin);
  return st;
}
