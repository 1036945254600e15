// api.hip -- the extern "C" entry points of include/ecalj_gpu.h.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "ctx.hpp"
#include "kernels.hpp"

using ecalj::cd;

namespace ecalj {

bool tune_lookup(ecalj_ctx* h, const TuneKey& k, int* var) {
  auto it = h->tuned.find(k);
  if (it == h->tuned.end()) return false;
  *var = it->second;
  return true;
}

void tune_store(ecalj_ctx* h, const TuneKey& k, int var, const char* op, const std::vector<const char*>& names,
                const std::vector<double>& ms) {
  h->tuned[k] = var;
  char line[512];
  int p = snprintf(line, sizeof line, "%s [%lld %lld %lld %lld]:", op, (long long)k[1], (long long)k[2],
                   (long long)k[3], (long long)k[4]);
  for (size_t v = 0; v < names.size() && p < (int)sizeof line; ++v)
    p += snprintf(line + p, sizeof line - p, " %s %.3f ms%s", names[v], ms[v], (int)v == var ? " *" : "");
  h->tune_log += line;
  h->tune_log += "\n";
}

ecalj_status slot_ensure(ecalj_ctx* h, int slot, size_t bytes, void** out) {
  if (bytes == 0) bytes = 256;
  if (h->arena) {  // fixed arena: no allocation, no synchronisation
    if (h->arena_len[slot] < bytes) {  // first use of the slot in this call, or a larger one: a fresh region
      const size_t off = (h->arena_used + 255) & ~(size_t)255;
      if (off > h->arena_cap || bytes > h->arena_cap - off) return ECALJ_ERR_DIM;
      h->arena_off[slot] = off;
      h->arena_len[slot] = bytes;
      h->arena_used = off + bytes;
    }
    *out = h->arena + h->arena_off[slot];
    return ECALJ_OK;
  }
  if (h->slot_bytes[slot] < bytes && h->capturing) return ECALJ_ERR_STATE;  // warm up before capture
  if (h->slot_bytes[slot] < bytes) {
    if (h->slot[slot]) {
      ECALJ_HIP(hipStreamSynchronize(h->stream));
      ECALJ_HIP(hipFree(h->slot[slot]));
      h->slot[slot] = nullptr;
      h->slot_bytes[slot] = 0;
      ++h->state_gen;  // graphs captured before this point referenced the old slot
    }
    size_t want = bytes + bytes / 8;
    if (hipMalloc(&h->slot[slot], want) != hipSuccess) {
      h->slot[slot] = nullptr;
      return ECALJ_ERR_NOMEM;
    }
    h->slot_bytes[slot] = want;
  }
  *out = h->slot[slot];
  return ECALJ_OK;
}

ecalj_status timing_begin(ecalj_ctx* h, int kclass, int* idx) {
  *idx = -1;
  if (!h->timing) return ECALJ_OK;
  auto& pool = h->ev_pool[kclass];
  int i = h->ev_used[kclass];
  if (i >= (int)pool.size()) {
    ecalj_ctx::EvPair p;
    ECALJ_HIP(hipEventCreate(&p.a));
    ECALJ_HIP(hipEventCreate(&p.b));
    pool.push_back(p);
  }
  ECALJ_HIP(hipEventRecord(pool[i].a, h->stream));
  h->ev_used[kclass] = i + 1;
  *idx = i;
  return ECALJ_OK;
}

ecalj_status timing_end(ecalj_ctx* h, int kclass, int idx) {
  if (idx < 0) return ECALJ_OK;
  ECALJ_HIP(hipEventRecord(h->ev_pool[kclass][idx].b, h->stream));
  return ECALJ_OK;
}

// Host restatement of hilbertmat (S/hilbertmat.f90:14-79), FP64.
static void hilbertmat_host(std::complex<double> zz, int nwhis, const double* L, const double* C,
                            const double* R, std::complex<double>* rmat) {
  using Z = std::complex<double>;
  const double eps = 1e-8;
  const Z zze = zz + Z(0.0, 1e-13);
  auto h2 = [&](double a, double b, Z& facm, Z& facn) {
    Z db = zze - b, da = zze - a;
    facm = 0.0;
    facn = 0.0;
    if (std::abs(da) > eps && std::abs(db) > eps) {
      facn = std::log(db / da);
      facm = zze * facn + (b - a);
    }
  };
  const double* l = L + nwhis;
  const double* c = C + nwhis;
  const double* r = R + nwhis;
  Z* rm = rmat + nwhis;
  for (int i = -nwhis; i <= nwhis; ++i) rm[i] = 0.0;
  for (int iw = -nwhis; iw <= nwhis; ++iw) {
    if (iw == 0) continue;
    const double ci = c[iw];
    const double cim = (iw == -nwhis) ? l[iw] : (iw == 1 ? 0.0 : c[iw - 1]);
    Z facm, facn;
    h2(l[iw], c[iw], facm, facn);
    rm[iw] += (facm - facn * cim) / (ci - cim);
    if (iw != -nwhis && iw != 1) rm[iw - 1] += (-facm + facn * ci) / (ci - cim);
    const double cip = (iw == nwhis) ? r[iw] : (iw == -1 ? 0.0 : c[iw + 1]);
    h2(c[iw], r[iw], facm, facn);
    rm[iw] += (-facm + facn * cip) / (cip - ci);
    if (iw != nwhis && iw != -1) rm[iw + 1] += (facm - facn * ci) / (cip - ci);
  }
}

// his_L/his_C/his_R (S/dpsion5.f90:84-86), index iw + nwhis, and the bin
// scale -wfac/(R-L) of dpsion_chiq (S/dpsion5.f90:193-200).
static void his_edges(int nwhis, const double* frhis, double ecut, std::vector<double>& L, std::vector<double>& C,
                      std::vector<double>& R, std::vector<double>& scale) {
  const int n2 = 2 * nwhis + 1;
  L.assign(n2, 0.0);
  C.assign(n2, 0.0);
  R.assign(n2, 0.0);
  for (int iw = -nwhis; iw <= nwhis; ++iw) {
    const int j = iw + nwhis;
    if (iw < 0) {
      L[j] = -frhis[-iw];
      R[j] = -frhis[-iw - 1];
    } else if (iw == 0) {
      L[j] = 0.0;
      R[j] = 0.0;
    } else {
      L[j] = frhis[iw - 1];
      R[j] = frhis[iw];
    }
    C[j] = 0.5 * (L[j] + R[j]);
  }
  scale.assign(nwhis, 0.0);
  for (int iw = 1; iw <= nwhis; ++iw) {
    const double hl = frhis[iw - 1], hr = frhis[iw], hc = 0.5 * (hl + hr);
    const double wfac = (ecut < 1e9) ? std::exp(-(hc / ecut) * (hc / ecut)) : 1.0;
    scale[iw - 1] = -wfac / (hr - hl);
  }
}

// imatt(iv, iw) * (-wfac(iw)/(R-L)) for npm = 1 (S/dpsion5.f90:84-89,121-129,215-220).
static void hilbert_coef(int nwhis, int niw, const double* frhis, const double* freqi, double ecut,
                         std::vector<double>& coef, std::vector<double>& scale) {
  const int n2 = 2 * nwhis + 1;
  std::vector<double> L, C, R;
  his_edges(nwhis, frhis, ecut, L, C, R, scale);
  const int ldc = hilbert_ldc(niw);
  coef.assign((size_t)ldc * nwhis, 0.0);
  std::vector<std::complex<double>> rr(n2);
  const double pi = 3.14159265358979323846;
  for (int it = 0; it < niw; ++it) {
    hilbertmat_host(std::complex<double>(0.0, freqi[it]), nwhis, L.data(), C.data(), R.data(), rr.data());
    for (int iw = 1; iw <= nwhis; ++iw)
      coef[it + (size_t)ldc * (iw - 1)] =
          std::real(rr[nwhis + iw] - rr[nwhis - iw]) / pi * scale[iw - 1];
  }
}

// Real axis, npm = 1 without chipm (S/dpsion5.f90:91-103,244-256):
// coef(it, iw) = rmatt(it, iw) * bscale(iw), rmatt = Re(rmat(iw) - rmat(-iw))/pi
// of hilbertmat at zz = freq_r(it); dscale(it) multiplies the i*rcxq(:,:,it)
// term: 1 for slot 0 (never scaled), bscale(it) for it >= 1.  bscale is the
// bin scale -wfac/(R-L), or 1 when rcxq already holds the scaled values.
static void hilbert_coef_real(int nwhis, int nw, const double* frhis, const double* freqr, double ecut, bool scaled,
                              std::vector<double>& coef, std::vector<double>& dscale) {
  std::vector<double> L, C, R, scale;
  his_edges(nwhis, frhis, ecut, L, C, R, scale);
  if (scaled) scale.assign(nwhis, 1.0);
  const int n2 = 2 * nwhis + 1, nv = nw + 1, ldc = hilbert_ldc(nv);
  coef.assign((size_t)ldc * nwhis, 0.0);
  dscale.assign(nwhis + 1, 0.0);
  dscale[0] = 1.0;
  for (int iw = 1; iw <= nwhis; ++iw) dscale[iw] = scale[iw - 1];
  const double pi = 3.14159265358979323846;
  std::vector<std::complex<double>> rr(n2);
  for (int it = 0; it < nv; ++it) {
    hilbertmat_host(std::complex<double>(freqr[it], 0.0), nwhis, L.data(), C.data(), R.data(), rr.data());
    for (int iw = 1; iw <= nwhis; ++iw)
      coef[it + (size_t)ldc * (iw - 1)] = std::real(rr[nwhis + iw] - rr[nwhis - iw]) / pi * scale[iw - 1];
  }
}

template <class T>
static ecalj_status upload(ecalj_ctx* h, int slot, const std::vector<T>& v, T** out) {
  const size_t bytes = sizeof(T) * v.size();
  ECALJ_TRY(slot_ensure(h, slot, bytes, (void**)out));
  if (v.empty()) return ECALJ_OK;
  if (h->capturing) {  // the graph's copy node reads this buffer at every replay: give it its own
    void* pin = nullptr;
    if (hipHostMalloc(&pin, bytes, hipHostMallocDefault) != hipSuccess) return ECALJ_ERR_NOMEM;
    h->cap_pinned.push_back(pin);
    memcpy(pin, v.data(), bytes);
    ECALJ_HIP(hipMemcpyAsync(*out, pin, bytes, hipMemcpyHostToDevice, h->stream));
    return ECALJ_OK;
  }
  // stage through pinned memory so the copy is asynchronous; the host waits
  // only for the previous copy out of the same staging buffer
  if (h->pin_ev[slot]) ECALJ_HIP(hipEventSynchronize(h->pin_ev[slot]));
  else ECALJ_HIP(hipEventCreateWithFlags(&h->pin_ev[slot], hipEventDisableTiming));
  if (h->pin_bytes[slot] < bytes) {
    if (h->pin[slot]) ECALJ_HIP(hipHostFree(h->pin[slot]));
    h->pin[slot] = nullptr;
    h->pin_bytes[slot] = 0;
    const size_t want = bytes + bytes / 4;
    if (hipHostMalloc(&h->pin[slot], want, hipHostMallocDefault) != hipSuccess) {
      h->pin[slot] = nullptr;
      return ECALJ_ERR_NOMEM;
    }
    h->pin_bytes[slot] = want;
  }
  memcpy(h->pin[slot], v.data(), bytes);
  ECALJ_HIP(hipMemcpyAsync(*out, h->pin[slot], bytes, hipMemcpyHostToDevice, h->stream));
  ECALJ_HIP(hipEventRecord(h->pin_ev[slot], h->stream));
  return ECALJ_OK;
}

ecalj_status upload_blob(ecalj_ctx* h, int slot, const std::vector<unsigned char>& v, unsigned char** out) {
  return upload(h, slot, v, out);
}

ecalj_status check_flag(ecalj_ctx* h) {
  int f = 0;
  ECALJ_HIP(hipMemcpyAsync(&f, h->d_flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  if (f) {
    int z = 0;
    ECALJ_HIP(hipMemcpyAsync(h->d_flag, &z, sizeof(int), hipMemcpyHostToDevice, h->stream));
    ECALJ_HIP(hipStreamSynchronize(h->stream));
    return ECALJ_ERR_SINGULAR;
  }
  return ECALJ_OK;
}

}  // namespace ecalj

using namespace ecalj;

extern "C" {

const char* ecalj_status_string(int s) {
  switch (s) {
    case ECALJ_OK: return "ok";
    case ECALJ_ERR_DIM: return "inconsistent or unsupported dimensions";
    case ECALJ_ERR_SINGULAR: return "epsilon matrix singular";
    case ECALJ_ERR_HIP: return "HIP runtime error";
    case ECALJ_ERR_ARG: return "bad argument";
    case ECALJ_ERR_NOMEM: return "device allocation failed";
    case ECALJ_ERR_STATE: return "required set-up call missing";
    case ECALJ_ERR_NOT_PD: return "matrix not positive definite";
    case ECALJ_ERR_NOT_CONVERGED: return "eigensolver did not converge";
    case ECALJ_ERR_SCRATCH: return "request exceeds the fixed scratch of the entry point that allocates nothing";
    default: return "unknown status";
  }
}

// Scratch of the entry points that allocate nothing (ecalj_zminv_noalloc_dev):
// one fixed array per device, held until the process ends.  In host code the
// name `scratch` is only a host-side stand-in; kernels get the device address
// taken by hipGetSymbolAddress (per device, at the first ecalj_create on it).
static __device__ __align__(256) unsigned char scratch[ECALJ_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_tried[kMaxDevices] = {};

// With `device` current: the device address of its copy of `scratch`
// (nullptr when the lookup failed; ecalj_zminv_noalloc_dev then refuses).
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_tried[device]) return;
  g_scratch_tried[device] = true;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess && p) g_scratch_ptr[device] = p;
  else (void)hipGetLastError();  // leave no sticky error for the next call
}

void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}
}  // namespace

ecalj_status ecalj_create(ecalj_handle* out, int device) {
  if (!out) return ECALJ_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  ECALJ_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return ECALJ_ERR_ARG;
  ECALJ_HIP(hipSetDevice(device));
  scratch_setup(device);
  static std::atomic<uint64_t> next_id{0};
  ecalj_ctx* h = new ecalj_ctx;
  h->device = device;
  h->id = ++next_id;
  if (hipStreamCreate(&h->own_stream) != hipSuccess) {
    delete h;
    return ECALJ_ERR_HIP;
  }
  h->stream = h->own_stream;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    h->ncu = ncu;
  if (hipMalloc(&h->d_flag, sizeof(int)) != hipSuccess) {
    (void)hipStreamDestroy(h->own_stream);
    delete h;
    return ECALJ_ERR_NOMEM;
  }
  int z = 0;
  (void)hipMemcpy(h->d_flag, &z, sizeof(int), hipMemcpyHostToDevice);
  *out = h;
  return ECALJ_OK;
}

ecalj_status ecalj_destroy(ecalj_handle h) {
  if (!h) return ECALJ_ERR_ARG;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  for (int s = 0; s < ecalj_ctx::S_NSLOT; ++s) {
    if (h->slot[s]) (void)hipFree(h->slot[s]);
    if (h->pin[s]) (void)hipHostFree(h->pin[s]);
    if (h->pin_ev[s]) (void)hipEventDestroy(h->pin_ev[s]);
  }
  if (h->d_ppb) (void)hipFree(h->d_ppb);
  if (h->d_ppb2) (void)hipFree(h->d_ppb2);
  if (h->d_ppbc) (void)hipFree(h->d_ppbc);
  if (h->d_vcousq) (void)hipFree(h->d_vcousq);
  if (h->d_zcousq) (void)hipFree(h->d_zcousq);
  if (h->d_zcousq32) (void)hipFree(h->d_zcousq32);
  for (void* p : {(void*)h->d_ggg, (void*)h->d_igggi, (void*)h->d_nvgcgp2, (void*)h->d_igcgp2i})
    if (p) (void)hipFree(p);
  if (h->d_flag) (void)hipFree(h->d_flag);
  rot_free(h);
  for (auto& pool : h->ev_pool)
    for (auto& p : pool) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
  return ECALJ_OK;
}

// ------------------------------------------------------------------ HIP graphs
struct ecalj_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<void*> pinned;   // staging the graph's copy nodes read at every replay
  uint64_t owner = 0;          // id of the capturing handle
  uint64_t gen = 0;            // its state generation at capture
  hipEvent_t done = nullptr;   // recorded after each launch; destroy waits for it
};

// True when the process runs under rocprofv3 (it exports its options as
// ROCPROF_* variables to the program it starts).  A kernel-trace run of HIP
// graph replays here ended in a segmentation fault inside hipGraphLaunch, in
// a host memcpy issued by the tracer's library at a page-aligned address; the
// graphs' own host memory (pinned staging, kept until destroy) was not it.
// Graph capture is refused under the profiler: profile the eager calls.
static bool under_profiler() {
  return getenv("ROCPROF_OUTPUT_PATH") || getenv("ROCPROF_KERNEL_TRACE") || getenv("ROCPROF_STATS");
}

ecalj_status ecalj_graph_begin(ecalj_handle h) {
  if (!h) return ECALJ_ERR_ARG;
  if (h->capturing) return ECALJ_ERR_STATE;
  if (under_profiler()) return ECALJ_ERR_STATE;
  ECALJ_HIP(hipSetDevice(h->device));
  ECALJ_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));  // pinned staging is allocated while capturing
  h->capturing = true;
  h->timing_before_capture = h->timing;
  h->timing = false;
  return ECALJ_OK;
}

ecalj_status ecalj_graph_end(ecalj_handle h, ecalj_graph_t* out) {
  if (!h || !out) return ECALJ_ERR_ARG;
  if (!h->capturing) return ECALJ_ERR_STATE;
  *out = nullptr;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &g);
  h->capturing = false;
  h->timing = h->timing_before_capture;
  auto* gr = new ecalj_graph_s;
  gr->pinned.swap(h->cap_pinned);
  if (e != hipSuccess || !g) {
    for (void* p : gr->pinned) (void)hipHostFree(p);
    delete gr;
    return ECALJ_ERR_HIP;
  }
  gr->graph = g;
  gr->owner = h->id;
  gr->gen = h->state_gen;
  if (hipGraphInstantiate(&gr->exec, g, nullptr, nullptr, 0) != hipSuccess ||
      hipEventCreateWithFlags(&gr->done, hipEventDisableTiming) != hipSuccess) {
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    (void)hipGraphDestroy(g);
    for (void* p : gr->pinned) (void)hipHostFree(p);
    delete gr;
    return ECALJ_ERR_HIP;
  }
  *out = gr;
  return ECALJ_OK;
}

ecalj_status ecalj_graph_launch(ecalj_handle h, ecalj_graph_t g) {
  if (!h || !g) return ECALJ_ERR_ARG;
  if (h->capturing) return ECALJ_ERR_STATE;
  if (g->owner != h->id) return ECALJ_ERR_ARG;          // the nodes reference the capturing handle's memory
  if (g->gen != h->state_gen) return ECALJ_ERR_STATE;   // that memory was freed or moved since the capture
  ECALJ_HIP(hipSetDevice(h->device));
  ECALJ_HIP(hipGraphLaunch(g->exec, h->stream));
  ECALJ_HIP(hipEventRecord(g->done, h->stream));
  return ECALJ_OK;
}

ecalj_status ecalj_graph_destroy(ecalj_graph_t g) {
  if (!g) return ECALJ_ERR_ARG;
  if (g->done) {  // the pinned staging may still be read by a launch in flight
    (void)hipEventSynchronize(g->done);
    (void)hipEventDestroy(g->done);
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (void* p : g->pinned) (void)hipHostFree(p);
  delete g;
  return ECALJ_OK;
}

ecalj_status ecalj_set_tuning(ecalj_handle h, int mode) {
  if (!h || mode < 0) return ECALJ_ERR_ARG;
  h->tune_mode = mode;
  return ECALJ_OK;
}

ecalj_status ecalj_tuning_report(ecalj_handle h, char* buf, int64_t len) {
  if (!h || !buf || len < 1) return ECALJ_ERR_ARG;
  const size_t n = std::min<size_t>(h->tune_log.size(), (size_t)len - 1);
  memcpy(buf, h->tune_log.data(), n);
  buf[n] = 0;
  return ECALJ_OK;
}

ecalj_status ecalj_set_precision(ecalj_handle h, int p) {
  if (!h) return ECALJ_ERR_ARG;
  if (p != ECALJ_PREC_F64 && p != ECALJ_PREC_F32 && p != ECALJ_PREC_BF16X3 && p != ECALJ_PREC_BF16X3_F8)
    return ECALJ_ERR_ARG;
  h->sigc_fp8 = p == ECALJ_PREC_BF16X3_F8;
  h->precision = h->sigc_fp8 ? ECALJ_PREC_BF16X3 : p;
  return ECALJ_OK;
}

ecalj_status ecalj_set_stream(ecalj_handle h, void* s) {
  if (!h) return ECALJ_ERR_ARG;
  h->stream = s ? static_cast<hipStream_t>(s) : h->own_stream;
  return ECALJ_OK;
}

// h's stream waits (on the device) for the work issued so far on other's stream.
ecalj_status ecalj_stream_wait(ecalj_handle h, ecalj_handle other) {
  if (!h || !other) return ECALJ_ERR_ARG;
  if (h->capturing || other->capturing) return ECALJ_ERR_STATE;
  if (h->stream == other->stream) return ECALJ_OK;
  if (!other->ev_join) ECALJ_HIP(hipEventCreateWithFlags(&other->ev_join, hipEventDisableTiming));
  ECALJ_HIP(hipEventRecord(other->ev_join, other->stream));
  ECALJ_HIP(hipStreamWaitEvent(h->stream, other->ev_join, 0));
  return ECALJ_OK;
}

ecalj_status ecalj_sync(ecalj_handle h) {
  if (!h) return ECALJ_ERR_ARG;
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return check_flag(h);
}

ecalj_status ecalj_kernel_timing(ecalj_handle h, int enable) {
  if (!h) return ECALJ_ERR_ARG;
  h->timing = enable != 0;
  for (int c = 0; c < 5; ++c) h->ev_used[c] = 0;
  return ECALJ_OK;
}

ecalj_status ecalj_kernel_time(ecalj_handle h, int kclass, double* ms, int64_t* launches) {
  if (!h || kclass < 0 || kclass >= 5 || !ms || !launches) return ECALJ_ERR_ARG;
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  double tot = 0.0;
  for (int i = 0; i < h->ev_used[kclass]; ++i) {
    float t = 0.f;
    ECALJ_HIP(hipEventElapsedTime(&t, h->ev_pool[kclass][i].a, h->ev_pool[kclass][i].b));
    tot += t;
  }
  *ms = tot;
  *launches = h->ev_used[kclass];
  h->ev_used[kclass] = 0;
  return ECALJ_OK;
}

// ------------------------------------------------------------------ device memory
ecalj_status ecalj_dev_alloc(ecalj_handle h, size_t bytes, void** ptr) {
  if (!h || !ptr) return ECALJ_ERR_ARG;
  *ptr = nullptr;
  ECALJ_HIP(hipSetDevice(h->device));
  if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) {
    *ptr = nullptr;
    return ECALJ_ERR_NOMEM;
  }
  return ECALJ_OK;
}

ecalj_status ecalj_dev_free(ecalj_handle h, void* ptr) {
  if (!h) return ECALJ_ERR_ARG;
  if (!ptr) return ECALJ_OK;
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  ECALJ_HIP(hipFree(ptr));
  return ECALJ_OK;
}

ecalj_status ecalj_memcpy_h2d(ecalj_handle h, void* dst, const void* src, size_t bytes) {
  if (!h || (bytes && (!dst || !src))) return ECALJ_ERR_ARG;
  if (!bytes) return ECALJ_OK;
  ECALJ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

ecalj_status ecalj_memcpy_d2h(ecalj_handle h, void* dst, const void* src, size_t bytes) {
  if (!h || (bytes && (!dst || !src))) return ECALJ_ERR_ARG;
  if (!bytes) return ECALJ_OK;
  ECALJ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

ecalj_status ecalj_memcpy_d2d(ecalj_handle h, void* dst, const void* src, size_t bytes) {
  if (!h || (bytes && (!dst || !src))) return ECALJ_ERR_ARG;
  if (!bytes) return ECALJ_OK;
  ECALJ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->stream));
  return ECALJ_OK;
}

ecalj_status ecalj_dev_zero(ecalj_handle h, void* ptr, size_t bytes) {
  if (!h || (bytes && !ptr)) return ECALJ_ERR_ARG;
  return dev_zero(h, ptr, bytes);
}

ecalj_status ecalj_dev_copy(ecalj_handle h, void* dst, const void* src, size_t bytes) {
  if (!h || (bytes && (!dst || !src))) return ECALJ_ERR_ARG;
  return dev_copy(h, dst, src, bytes);
}

ecalj_status ecalj_dev_phase_similarity(ecalj_handle h, const void* a, int ngb, int nmat, const void* d, int nk,
                                        void* out) {
  if (!h || !a || !d || !out) return ECALJ_ERR_ARG;
  if (ngb < 0 || nmat < 0 || nk < 0) return ECALJ_ERR_DIM;
  return phase_similarity(h, (const cd*)a, ngb, nmat, (const cd*)d, nk, (cd*)out);
}

// ------------------------------------------------------------------ state
ecalj_status ecalj_set_ppb(ecalj_handle h, int natom, const int* nv, const int* mdim, const double* ppb) {
  if (!h || natom < 1 || !nv || !mdim || !ppb) return ECALJ_ERR_ARG;
  int64_t tot = 0;
  h->nv.assign(nv, nv + natom);
  h->mdim.assign(mdim, mdim + natom);
  h->ppb_off.resize(natom);
  h->ndima_used = 0;
  h->ngb_ppb = 0;
  for (int a = 0; a < natom; ++a) {
    if (nv[a] < 0 || mdim[a] < 0) return ECALJ_ERR_DIM;
    h->ppb_off[a] = tot;
    tot += (int64_t)nv[a] * nv[a] * mdim[a];
    h->ndima_used += nv[a];
    h->ngb_ppb += mdim[a];
  }
  h->natom = natom;
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  ++h->state_gen;
  h->ncore.assign(natom, 0);  // core blocks come only with ecalj_set_ppb_full
  h->ppbc_off.assign(natom, 0);
  if (h->d_ppbc) (void)hipFree(h->d_ppbc);
  h->d_ppbc = nullptr;
  if (h->d_ppb) (void)hipFree(h->d_ppb);
  if (h->d_ppb2) (void)hipFree(h->d_ppb2);
  h->d_ppb = h->d_ppb2 = nullptr;
  const size_t bytes = sizeof(double) * std::max<int64_t>(tot, 1);
  if (hipMalloc(&h->d_ppb, bytes) != hipSuccess) return ECALJ_ERR_NOMEM;
  if (hipMalloc(&h->d_ppb2, bytes) != hipSuccess) return ECALJ_ERR_NOMEM;
  ECALJ_HIP(hipMemcpyAsync(h->d_ppb, ppb, sizeof(double) * tot, hipMemcpyHostToDevice, h->stream));
  for (int a = 0; a < natom; ++a)
    ECALJ_TRY(ppb_relayout(h, h->d_ppb + h->ppb_off[a], h->d_ppb2 + h->ppb_off[a], nv[a], mdim[a]));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

ecalj_status ecalj_set_ppb_full(ecalj_handle h, int natom, int nlnmx, int mdimx, const int* nlnmc, const int* nlnmv,
                                const int* nblocha, const double* ppb) {
  if (!h || natom < 1 || !nlnmc || !nlnmv || !nblocha || !ppb) return ECALJ_ERR_ARG;
  for (int a = 0; a < natom; ++a)
    if (nlnmc[a] < 0 || nlnmv[a] < 0 || nblocha[a] < 0 || nlnmc[a] + nlnmv[a] > nlnmx || nblocha[a] > mdimx)
      return ECALJ_ERR_DIM;
  // valence blocks ppb(nc+1:nc+nv, nc+1:nc+nv, 1:mdim, a) for set_ppb, core blocks
  // ppb(nc+1:nc+nv, 1:nc, 1:mdim, a) (valence x core x product basis) for the CoreValence rows
  const int64_t s1 = nlnmx, s2 = (int64_t)nlnmx * nlnmx, s3 = s2 * mdimx;
  std::vector<double> pv, pc;
  std::vector<int64_t> coff(natom);
  for (int a = 0; a < natom; ++a) {
    const int nc = nlnmc[a], nv = nlnmv[a], md = nblocha[a];
    const double* P = ppb + s3 * a;
    for (int i = 0; i < md; ++i)
      for (int b = 0; b < nv; ++b)
        for (int c = 0; c < nv; ++c) pv.push_back(P[(nc + c) + s1 * (nc + b) + s2 * i]);
    coff[a] = (int64_t)pc.size();
    for (int i = 0; i < md; ++i)
      for (int c = 0; c < nc; ++c)
        for (int b = 0; b < nv; ++b) pc.push_back(P[(nc + b) + s1 * c + s2 * i]);
  }
  ECALJ_TRY(ecalj_set_ppb(h, natom, nlnmv, nblocha, pv.data()));
  h->ncore.assign(nlnmc, nlnmc + natom);
  h->ppbc_off = coff;
  if (!pc.empty()) {
    if (hipMalloc(&h->d_ppbc, sizeof(double) * pc.size()) != hipSuccess) {
      h->d_ppbc = nullptr;
      return ECALJ_ERR_NOMEM;
    }
    ECALJ_HIP(hipMemcpyAsync(h->d_ppbc, pc.data(), sizeof(double) * pc.size(), hipMemcpyHostToDevice, h->stream));
    ECALJ_HIP(hipStreamSynchronize(h->stream));
  }
  return ECALJ_OK;
}

ecalj_status ecalj_set_coulomb(ecalj_handle h, int ngb, const double* vcoud, const double* zcousq) {
  if (!h || ngb < 1 || !vcoud || !zcousq) return ECALJ_ERR_ARG;
  std::vector<double> vs(ngb);
  for (int i = 0; i < ngb; ++i) {
    if (!(vcoud[i] >= 0.0)) return ECALJ_ERR_ARG;
    vs[i] = std::sqrt(vcoud[i]);  // S/m_readVcoud.f90:40
  }
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  ++h->state_gen;
  if (h->d_vcousq) (void)hipFree(h->d_vcousq);
  if (h->d_zcousq) (void)hipFree(h->d_zcousq);
  if (h->d_zcousq32) (void)hipFree(h->d_zcousq32);
  h->d_vcousq = nullptr;
  h->d_zcousq = nullptr;
  h->d_zcousq32 = nullptr;
  if (hipMalloc(&h->d_vcousq, sizeof(double) * ngb) != hipSuccess) return ECALJ_ERR_NOMEM;
  if (hipMalloc(&h->d_zcousq, sizeof(cd) * (size_t)ngb * ngb) != hipSuccess) return ECALJ_ERR_NOMEM;
  if (hipMalloc(&h->d_zcousq32, sizeof(ecalj::cf) * (size_t)ngb * ngb) != hipSuccess) return ECALJ_ERR_NOMEM;
  ECALJ_HIP(hipMemcpyAsync(h->d_vcousq, vs.data(), sizeof(double) * ngb, hipMemcpyHostToDevice, h->stream));
  ECALJ_HIP(hipMemcpyAsync(h->d_zcousq, zcousq, sizeof(cd) * (size_t)ngb * ngb, hipMemcpyHostToDevice, h->stream));
  ECALJ_TRY(convert(h, h->d_zcousq, true, h->d_zcousq32, false, (int64_t)ngb * ngb));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  h->ngb = ngb;
  return ECALJ_OK;
}

// ------------------------------------------------------------------ helpers
static inline bool wide(const ecalj_ctx* h) { return h->precision == ECALJ_PREC_F64; }
static inline size_t celt(const ecalj_ctx* h) { return wide(h) ? sizeof(cd) : sizeof(ecalj::cf); }

// Host staging for the host-pointer entry points: a complex128 host array
// lives on the device in the handle's storage type.
struct DevC {
  DevBuf buf;
  ecalj_status put(ecalj_ctx* h, const void* host, int64_t n) {
    ECALJ_TRY(buf.alloc(celt(h) * std::max<int64_t>(n, 1)));
    if (!host || n == 0) return ECALJ_OK;
    if (wide(h)) {
      ECALJ_HIP(hipMemcpyAsync(buf.p, host, sizeof(cd) * n, hipMemcpyHostToDevice, h->stream));
      return ECALJ_OK;
    }
    DevBuf t;
    ECALJ_TRY(t.put(h, host, sizeof(cd) * n));
    ECALJ_TRY(convert(h, t.p, true, buf.p, false, n));
    ECALJ_HIP(hipStreamSynchronize(h->stream));
    return ECALJ_OK;
  }
  ecalj_status get(ecalj_ctx* h, void* host, int64_t n) {
    if (n == 0) return ECALJ_OK;
    if (wide(h)) {
      ECALJ_HIP(hipMemcpyAsync(host, buf.p, sizeof(cd) * n, hipMemcpyDeviceToHost, h->stream));
    } else {
      DevBuf t;
      ECALJ_TRY(t.alloc(sizeof(cd) * n));
      ECALJ_TRY(convert(h, buf.p, false, t.p, true, n));
      ECALJ_HIP(hipMemcpyAsync(host, t.p, sizeof(cd) * n, hipMemcpyDeviceToHost, h->stream));
      ECALJ_HIP(hipStreamSynchronize(h->stream));
    }
    ECALJ_HIP(hipStreamSynchronize(h->stream));
    return ECALJ_OK;
  }
};

// ------------------------------------------------------------------ zmel
// get_zmel_init_gemm (S/m_zmel.f90:146-597), muffin-tin part.  For each atom
// block the smaller side is contracted with ppb first:
//   m-first: T(b,I,n) = sum_a ppb(b,a,I) conj(cphim(a,n));  zmel(I,n,m) = phase sum_b T(b,I,n) cphiq(b,m)
//   q-first: T(a,I,m) = sum_b ppb(b,a,I) cphiq(b,m);        zmel(I,n,m) = phase sum_a conj(cphim(a,n)) T(a,I,m)
// cphim / cphiq are complex128; zmel is in the handle's storage type.
// Optional parts of the middle-state range (ZmelRange): the valence columns
// of cphim land at column voff of nmt middle states; core states (columns
// core_c0.. in the range) come from the core blocks of ecalj_set_ppb_full
// (S/m_zmel.f90:374-390); iatomp maps atom ia to the atom whose product
// basis, ppb block and row offset it uses (S/m_zmel.f90:295-305, 334).
struct ZmelRange {
  int nmt = -1;            // middle states in zmel (-1: the valence count nm)
  int voff = 0;            // column of the first valence state
  const int* iatomp = nullptr;  // host, 0-based; nullptr = identity
  int nm1c = 0, nm2c = -1; // core states in the range (1-based, the reference's nm1c:nm2c)
  int nm1 = 1;             // first middle state of the range (ns1)
  int nctot = 0;
};

static ecalj_status zmel_impl(ecalj_ctx* h, const double* phase, int ndima, int nm, const cd* cphim64,
                              int64_t sm, int nq, const cd* cphiq64, int64_t sq, int is_m_basis, int conjg,
                              int nbatch, void* zmel, const ZmelRange& rg = ZmelRange()) {
  if (!h->d_ppb) return ECALJ_ERR_STATE;
  if (nm < 0 || nq < 0 || nbatch < 1 || ndima < h->ndima_used) return ECALJ_ERR_DIM;
  const int ngb = h->ngb_ppb;
  if (!is_m_basis && (!h->d_zcousq || h->ngb != ngb)) return ECALJ_ERR_STATE;
  const int nmt = rg.nmt < 0 ? nm : rg.nmt;
  const bool any_core = rg.nm2c >= rg.nm1c && rg.nctot > 0;
  if (any_core && !h->d_ppbc) return ECALJ_ERR_STATE;
  if (nmt == 0 || nq == 0) return ECALJ_OK;
  const int mode = h->precision;
  const size_t es = celt(h);
  const int bm = (sm == 0) ? 1 : nbatch;
  const int bq = (sq == 0) ? 1 : nbatch;
  // per-atom offsets: eigenvector rows (own atom), product-basis rows and ppb (mapped atom)
  const int na = h->natom;
  std::vector<int64_t> ias_v(na + 1, 0), ims_v(na + 1, 0), ics_v(na + 1, 0);
  for (int a = 0; a < na; ++a) {
    ias_v[a + 1] = ias_v[a] + h->nv[a];
    ims_v[a + 1] = ims_v[a] + h->mdim[a];
    ics_v[a + 1] = ics_v[a] + (a < (int)h->ncore.size() ? h->ncore[a] : 0);
  }
  if (rg.iatomp)
    for (int a = 0; a < na; ++a) {
      const int ap = rg.iatomp[a];
      if (ap < 0 || ap >= na || h->nv[ap] != h->nv[a]) return ECALJ_ERR_DIM;
      if (any_core && h->ncore[ap] != h->ncore[a]) return ECALJ_ERR_DIM;
    }
  // eigenvectors in the storage type
  const void* cphim = cphim64;
  const void* cphiq = cphiq64;
  if (!wide(h)) {
    void *m32, *q32;
    const int64_t nme = (bm == 1) ? (int64_t)ndima * nm : sm * (nbatch - 1) + (int64_t)ndima * nm;
    const int64_t nqe = (bq == 1) ? (int64_t)ndima * nq : sq * (nbatch - 1) + (int64_t)ndima * nq;
    ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_CVT1, sizeof(ecalj::cf) * std::max<int64_t>(nme, 1), &m32));
    ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_CVT2, sizeof(ecalj::cf) * nqe, &q32));
    if (nme > 0) ECALJ_TRY(convert(h, cphim64, true, m32, false, nme));
    ECALJ_TRY(convert(h, cphiq64, true, q32, false, nqe));
    cphim = m32;
    cphiq = q32;
  }
  const int64_t zsz = (int64_t)ngb * nmt * nq;  // per batch
  char* zM = (char*)zmel;
  if (!is_m_basis) ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_TMP2, es * zsz * nbatch, (void**)&zM));
  if (nmt != nm) ECALJ_TRY(dev_zero(h, zM, es * zsz * nbatch));  // core columns: rows of other atoms stay 0
  const bool conj_now = conjg && is_m_basis;
  const int64_t vcol = (int64_t)ngb * rg.voff;  // element offset of the first valence column
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 2, &tev));
  for (int a = 0; a < na; ++a) {
    const int ap = rg.iatomp ? rg.iatomp[a] : a;
    const int nv = h->nv[a], md = h->mdim[ap];
    const int64_t ias = ias_v[a], ims = ims_v[ap];
    const cd ph{phase ? phase[2 * a] : 1.0, phase ? phase[2 * a + 1] : 0.0};
    const int64_t nvmd = (int64_t)nv * md;
    if (nm > 0) {
      char* T;
      if ((int64_t)nm * bm <= (int64_t)nq * bq) {
        ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_T, es * nvmd * nm * bm, (void**)&T));
        GemmCall g;  // T[(b+nv*I), (n,k)] = P2 * conj(cphim)
        g.mode = mode;
        g.areal = true; g.cb = true;
        g.M = (int)nvmd; g.K = nv;
        g.A = h->d_ppb2 + h->ppb_off[ap]; g.lda = nvmd;
        g.B = (const char*)cphim + es * ias; g.ldb = ndima;
        g.C = T; g.ldc = nvmd;
        if (bm == 1 || sm == (int64_t)ndima * nm) {
          g.N = nm * bm;
        } else {
          g.N = nm;
          g.bB = {sm, 0}; g.bC = {nvmd * nm, 0};
          g.batch = bm; g.div = bm;
        }
        if (mode == ECALJ_PREC_BF16X3)
          ECALJ_TRY(zmel_p_bf3(h, (int)nvmd, g.N, nv, h->d_ppb2 + h->ppb_off[ap], nvmd,
                               (const ecalj::cf*)cphim + ias, ndima, g.batch > 1 ? sm : 0, true, (ecalj::cf*)T, nvmd,
                               g.batch > 1 ? nvmd * nm : 0, g.batch));
        else
          ECALJ_TRY(gemm(h, g));
        GemmCall z;  // per (n,k): zmel(ims+I, voff+n, m) = phase * sum_b T[b + nv*I + nvmd*(n + nm*k)] cphiq_k(ias+b, m)
        z.mode = mode;
        z.ta = true;
        z.M = md; z.N = nq; z.K = nv;
        z.A = T; z.lda = nv; z.bA = {nvmd, (bm == 1) ? 0 : nvmd * nm};
        z.B = (const char*)cphiq + es * ias; z.ldb = ndima; z.bB = {0, sq};
        z.C = zM + es * (ims + vcol); z.ldc = (int64_t)ngb * nmt; z.bC = {ngb, zsz};
        z.batch = nm * nbatch; z.div = nm;
        z.alpha = ph; z.conjC = conj_now;
        if (mode == ECALJ_PREC_BF16X3)
          ECALJ_TRY(zmel_z_bf3(h, md, nq, nv, (const ecalj::cf*)T, nv, nvmd, (bm == 1) ? 0 : nvmd * nm,
                               (const ecalj::cf*)cphiq + ias, ndima, sq, false, ph, conj_now,
                               (ecalj::cf*)zM + ims + vcol, (int64_t)ngb * nmt, ngb, zsz, nm * nbatch, nm));
        else
          ECALJ_TRY(gemm(h, z));
      } else {
        ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_T, es * nvmd * nq * bq, (void**)&T));
        GemmCall g;  // T[(a+nv*I), (m,k)] = ppb^T * cphiq, A(r,b) = ppb[b + nv*r]
        g.mode = mode;
        g.areal = true; g.ta = true;
        g.M = (int)nvmd; g.K = nv;
        g.A = h->d_ppb + h->ppb_off[ap]; g.lda = nv;
        g.B = (const char*)cphiq + es * ias; g.ldb = ndima;
        g.C = T; g.ldc = nvmd;
        if (bq == 1 || sq == (int64_t)ndima * nq) {
          g.N = nq * bq;
        } else {
          g.N = nq;
          g.bB = {sq, 0}; g.bC = {nvmd * nq, 0};
          g.batch = bq; g.div = bq;
        }
        ECALJ_TRY(gemm(h, g));
        GemmCall z;  // per (m,k): zmel(ims+I, voff+n, m) = phase * sum_a T[a + nv*I + nvmd*(m + nq*k)] conj(cphim_k(ias+a, n))
        z.mode = mode;
        z.ta = true; z.cb = true;
        z.M = md; z.N = nm; z.K = nv;
        z.A = T; z.lda = nv; z.bA = {nvmd, (bq == 1) ? 0 : nvmd * nq};
        z.B = (const char*)cphim + es * ias; z.ldb = ndima; z.bB = {0, sm};
        z.C = zM + es * (ims + vcol); z.ldc = ngb; z.bC = {(int64_t)ngb * nmt, zsz};
        z.batch = nq * nbatch; z.div = nq;
        z.alpha = ph; z.conjC = conj_now;
        if (mode == ECALJ_PREC_BF16X3)
          ECALJ_TRY(zmel_z_bf3(h, md, nm, nv, (const ecalj::cf*)T, nv, nvmd, (bq == 1) ? 0 : nvmd * nq,
                               (const ecalj::cf*)cphim + ias, ndima, sm, true, ph, conj_now,
                               (ecalj::cf*)zM + ims + vcol, ngb, (int64_t)ngb * nmt, zsz, nq * nbatch, nq));
        else
          ECALJ_TRY(gemm(h, z));
      }
    }
    if (any_core) {  // CoreValence (S/m_zmel.f90:374-390): zmel(ims+I, it, m) = phase sum_b ppb(nc+b, it-ics, I) cphiq(ias+b, m)
      const int nc = h->ncore[a];
      const int64_t ics = ics_v[a];
      const int lo = std::max<int64_t>(rg.nm1c, ics + 1), hi = std::min<int64_t>(rg.nm2c, ics + nc);
      if (hi >= lo) {
        const int ncc = hi - lo + 1;
        GemmCall c;  // A(I, b) = ppbc_ap[b + nv*(it-ics-1) + nv*nc*I]
        c.mode = mode;
        c.areal = true; c.ta = true;
        c.M = md; c.N = nq; c.K = nv;
        c.A = h->d_ppbc + h->ppbc_off[ap] + (int64_t)nv * (lo - ics - 1); c.lda = (int64_t)nv * nc;
        c.bA = {nv, 0};
        c.B = (const char*)cphiq + es * ias; c.ldb = ndima; c.bB = {0, sq};
        c.C = zM + es * (ims + (int64_t)ngb * (lo - rg.nm1)); c.ldc = (int64_t)ngb * nmt; c.bC = {ngb, zsz};
        c.batch = ncc * nbatch; c.div = ncc;
        c.alpha = ph; c.conjC = conj_now;
        ECALJ_TRY(gemm(h, c));
      }
    }
  }
  if (!is_m_basis) {  // zmel = zcousq^H zmel_M  (S/m_zmel.f90:559-560), then conj
    GemmCall e;
    e.mode = mode;
    e.ta = true; e.ca = true;
    e.M = ngb; e.N = (int)((int64_t)nmt * nq * nbatch); e.K = ngb;
    e.A = wide(h) ? (const void*)h->d_zcousq : (const void*)h->d_zcousq32; e.lda = ngb;
    e.B = zM; e.ldb = ngb;
    e.C = zmel; e.ldc = ngb;
    e.conjC = conjg;
    ECALJ_TRY(gemm(h, e));
  }
  ECALJ_TRY(timing_end(h, 2, tev));
  return ECALJ_OK;
}

ecalj_status ecalj_zmel_dev(ecalj_handle h, const double* phase, int ndima, int nm, const void* cphim, int nq,
                            const void* cphiq, int is_m_basis, int zmelconjg, int nbatch, void* zmel) {
  if (!h || !cphim || !cphiq || !zmel) return ECALJ_ERR_ARG;
  return zmel_impl(h, phase, ndima, nm, (const cd*)cphim, (int64_t)ndima * nm, nq, (const cd*)cphiq,
                   (int64_t)ndima * nq, is_m_basis, zmelconjg, nbatch, zmel);
}

ecalj_status ecalj_zmel_strided_dev(ecalj_handle h, const double* phase, int ndima, int nm, const void* cphim,
                                    int64_t cphim_stride, int nq, const void* cphiq, int64_t cphiq_stride,
                                    int is_m_basis, int zmelconjg, int nbatch, void* zmel) {
  if (!h || !cphim || !cphiq || !zmel) return ECALJ_ERR_ARG;
  if (cphim_stride < 0 || cphiq_stride < 0) return ECALJ_ERR_DIM;
  return zmel_impl(h, phase, ndima, nm, (const cd*)cphim, cphim_stride, nq, (const cd*)cphiq, cphiq_stride,
                   is_m_basis, zmelconjg, nbatch, zmel);
}

// Middle states ns1..ns2 over core states 1..nctot then valence bands
// (S/m_zmel.f90:199-214): the reference's nm1c:nm2c and nm1v:nm2v.
static void zmel_ranges(int ns1, int ns2, int nctot, int* nm1c, int* nm2c, int* nm1v, int* nm2v) {
  if (ns1 > nctot) { *nm1c = 0; *nm2c = -1; *nm1v = ns1; *nm2v = ns2; }
  else if (ns2 > nctot) { *nm1c = ns1; *nm2c = nctot; *nm1v = nctot + 1; *nm2v = ns2; }
  else { *nm1c = ns1; *nm2c = ns2; *nm1v = 0; *nm2v = -1; }
}

ecalj_status ecalj_zmel_init_dev(ecalj_handle h, const double* phase, const int* iatomp, int ndima, int ns1,
                                 int ns2, int nctot, const void* cphim, int nq, const void* cphiq, int is_m_basis,
                                 int zmelconjg, int nbatch, void* zmel) {
  if (!h || !cphiq || !zmel) return ECALJ_ERR_ARG;
  if (ns1 < 1 || ns2 < ns1 - 1 || nctot < 0 || nbatch < 1) return ECALJ_ERR_DIM;
  int nm1c, nm2c, nm1v, nm2v;
  zmel_ranges(ns1, ns2, nctot, &nm1c, &nm2c, &nm1v, &nm2v);
  const int nmv = nm2v >= nm1v ? nm2v - nm1v + 1 : 0;
  if (nmv > 0 && !cphim) return ECALJ_ERR_ARG;
  if (nm2c >= nm1c) {  // core states must be those of ecalj_set_ppb_full
    int nct = 0;
    for (int c : h->ncore) nct += c;
    if (nctot != nct || !h->d_ppbc) return ECALJ_ERR_STATE;
  }
  std::vector<int> map;
  ZmelRange rg;
  if (iatomp) {
    map.resize(h->natom);
    for (int a = 0; a < h->natom; ++a) {
      if (iatomp[a] < 1 || iatomp[a] > h->natom) return ECALJ_ERR_ARG;
      map[a] = iatomp[a] - 1;
    }
    rg.iatomp = map.data();
  }
  rg.nmt = ns2 - ns1 + 1;
  rg.voff = nmv > 0 ? nm1v - ns1 : 0;
  rg.nm1c = nm1c;
  rg.nm2c = nm2c;
  rg.nm1 = ns1;
  rg.nctot = nctot;
  return zmel_impl(h, phase, ndima, nmv, (const cd*)cphim, (int64_t)ndima * nmv, nq, (const cd*)cphiq,
                   (int64_t)ndima * nq, is_m_basis, zmelconjg, nbatch, zmel, rg);
}

ecalj_status ecalj_zmel_init(ecalj_handle h, const double* phase, const int* iatomp, int ndima, int ns1, int ns2,
                             int nctot, const double* cphim, int nq, const double* cphiq, int is_m_basis,
                             int zmelconjg, int nbatch, double* zmel) {
  if (!h || !cphiq || !zmel || nbatch < 1 || ns2 < ns1 - 1 || ns1 < 1) return ECALJ_ERR_ARG;
  int nm1c, nm2c, nm1v, nm2v;
  zmel_ranges(ns1, ns2, nctot, &nm1c, &nm2c, &nm1v, &nm2v);
  const int nmv = nm2v >= nm1v ? nm2v - nm1v + 1 : 0;
  const int nmt = ns2 - ns1 + 1;
  const size_t bm = sizeof(cd) * (size_t)ndima * nmv * nbatch, bq = sizeof(cd) * (size_t)ndima * nq * nbatch;
  const int64_t nz = (int64_t)h->ngb_ppb * nmt * nq * nbatch;
  DevBuf dm, dq;
  DevC dz;
  ECALJ_TRY(dm.put(h, nmv > 0 ? cphim : nullptr, bm));
  ECALJ_TRY(dq.put(h, cphiq, bq));
  ECALJ_TRY(dz.put(h, nullptr, nz));
  ECALJ_TRY(ecalj_zmel_init_dev(h, phase, iatomp, ndima, ns1, ns2, nctot, dm.p, nq, dq.p, is_m_basis, zmelconjg,
                                nbatch, dz.buf.p));
  return dz.get(h, zmel, nz);
}

ecalj_status ecalj_zmel(ecalj_handle h, const double* phase, int ndima, int nm, const double* cphim, int nq,
                        const double* cphiq, int is_m_basis, int zmelconjg, int nbatch, double* zmel) {
  if (!h || !cphim || !cphiq || !zmel || nbatch < 1) return ECALJ_ERR_ARG;
  const size_t bm = sizeof(cd) * (size_t)ndima * nm * nbatch, bq = sizeof(cd) * (size_t)ndima * nq * nbatch;
  const int64_t nz = (int64_t)h->ngb_ppb * nm * nq * nbatch;
  DevBuf dm, dq;
  DevC dz;
  ECALJ_TRY(dm.put(h, cphim, bm));
  ECALJ_TRY(dq.put(h, cphiq, bq));
  ECALJ_TRY(dz.put(h, nullptr, nz));
  ECALJ_TRY(ecalj_zmel_dev(h, phase, ndima, nm, dm.p, nq, dq.p, is_m_basis, zmelconjg, nbatch, dz.buf.p));
  return dz.get(h, zmel, nz);
}

// ------------------------------------------------------------------ x0gemm
// Bin lists of x0gemm (S/x0gemm.f90:36-64), built over all k of the batch,
// then the per-bin gather-Herk.
static ecalj_status x0_impl(ecalj_ctx* h, void* rcxq, int npr, int ipr_col, int npr_col, int nwhis, int npm, int nk,
                            const int* ns1, const int* ns2, const void* zmel, int64_t zstride, int nbb, int nqtot,
                            const int* kmin, const int* kmax, int ncoun, const int* iwini, const int* iwend,
                            const int* itc, const int* itpc, const int* jpmc, const int* icouini, int ncount,
                            const double* whwc) {
  if (npr < 0 || npr_col < 0 || ipr_col < 1 || ipr_col - 1 + npr_col > nbb || npr > nbb || nwhis < 1 ||
      (npm != 1 && npm != 2) || nk < 1)
    return ECALJ_ERR_DIM;
  const int nbin = nwhis * npm;
  std::vector<int64_t> cnt(nbin + 1, 0);
  auto visit = [&](auto&& f) -> ecalj_status {
    for (int k = 0; k < nk; ++k) {
      const int nm = ns2[k] - ns1[k] + 1;
      if (kmin[k] < 1 || kmax[k] > ncoun) return ECALJ_ERR_DIM;
      for (int ic = kmin[k]; ic <= kmax[k]; ++ic) {
        const int it = itc[ic - 1], itp = itpc[ic - 1], jpm = jpmc[ic - 1];
        if (it < ns1[k] || it > ns2[k]) continue;
        if (itp < 1 || itp > nqtot || jpm < 1 || jpm > npm) return ECALJ_ERR_DIM;
        const int64_t col = zstride * k + (int64_t)nbb * ((it - ns1[k]) + (int64_t)nm * (itp - 1));
        for (int iw = iwini[ic - 1]; iw <= iwend[ic - 1]; ++iw) {
          if (iw < 1 || iw > nwhis) return ECALJ_ERR_DIM;
          const int64_t wi = iw - iwini[ic - 1] + icouini[ic - 1] - 1;
          if (wi < 0 || wi >= ncount) return ECALJ_ERR_DIM;
          f((iw - 1) + nwhis * (jpm - 1), col, whwc[wi]);
        }
      }
    }
    return ECALJ_OK;
  };
  ECALJ_TRY(visit([&](int b, int64_t, double) { cnt[b + 1]++; }));
  for (int b = 0; b < nbin; ++b) cnt[b + 1] += cnt[b];
  const int64_t tot = cnt[nbin];
  std::vector<int64_t> col(std::max<int64_t>(tot, 1));
  std::vector<double> wt(std::max<int64_t>(tot, 1));
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  ECALJ_TRY(visit([&](int b, int64_t c, double w) {
    const int64_t p = fill[b]++;
    col[p] = c;
    wt[p] = w;
  }));
  // work items; square output -> upper tiles + mirror
  const int mirror = (ipr_col == 1 && npr_col == npr) ? 1 : 0;
  if (h->precision == ECALJ_PREC_BF16X3) {  // bins gathered into bf16 planes, grouped GEMM on the Sigma_c core
    std::vector<int> slice(nbin);
    for (int b = 0; b < nbin; ++b) {
      const int iw = b % nwhis + 1, jpm = b / nwhis + 1;
      slice[b] = iw * (3 - 2 * jpm) - (1 - npm) * nwhis;
    }
    return x0_bf3(h, (const ecalj::cf*)zmel, npr, npr_col, ipr_col, nbin, cnt.data(), col.data(), wt.data(),
                  slice.data(), (ecalj::cf*)rcxq, (int64_t)npr * npr_col, mirror);
  }
  const int tm = (npr + 63) / 64, tn = (npr_col + 63) / 64;
  struct Item { int list, slice, ti, tj; };
  std::vector<Item> items;
  std::vector<int64_t> kk;
  for (int b = 0; b < nbin; ++b) {
    const int64_t K = cnt[b + 1] - cnt[b];
    if (K == 0) continue;
    const int iw = b % nwhis + 1, jpm = b / nwhis + 1;
    const int slice = iw * (3 - 2 * jpm) - (1 - npm) * nwhis;
    for (int ti = 0; ti < tm; ++ti)
      for (int tj = mirror ? ti : 0; tj < tn; ++tj) {
        items.push_back({b, slice, ti, tj});
        kk.push_back(K);
      }
  }
  std::vector<int> order(items.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return kk[x] > kk[y]; });
  std::vector<Item> sorted(items.size());
  for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];
  // one upload: [bstart | col | wt | items]
  const size_t nb1 = nbin + 1;
  std::vector<char> blob(sizeof(int64_t) * (nb1 + col.size()) + sizeof(double) * wt.size() +
                         sizeof(Item) * std::max<size_t>(sorted.size(), 1));
  char* q = blob.data();
  memcpy(q, cnt.data(), sizeof(int64_t) * nb1);
  memcpy(q + sizeof(int64_t) * nb1, col.data(), sizeof(int64_t) * col.size());
  memcpy(q + sizeof(int64_t) * (nb1 + col.size()), wt.data(), sizeof(double) * wt.size());
  memcpy(q + sizeof(int64_t) * (nb1 + col.size()) + sizeof(double) * wt.size(), sorted.data(),
         sizeof(Item) * sorted.size());
  char* d;
  ECALJ_TRY(upload(h, ecalj_ctx::S_LIST, blob, &d));
  const int64_t* d_b = (const int64_t*)d;
  const int64_t* d_col = d_b + nb1;
  const double* d_w = (const double*)(d_col + col.size());
  const void* d_items = (const char*)d_w + sizeof(double) * wt.size();
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 1, &tev));
  ECALJ_TRY(x0_herk(h, h->precision, zmel, d_col, d_w, d_b, d_items, (int)sorted.size(), rcxq,
                    (int64_t)npr * npr_col, npr, npr_col, ipr_col, mirror));
  ECALJ_TRY(timing_end(h, 1, tev));
  return ECALJ_OK;
}

ecalj_status ecalj_x0gemm_batched_dev(ecalj_handle h, void* rcxq, int npr, int ipr_col, int npr_col, int nwhis,
                                      int npm, int nk, const int* ns1, const int* ns2, const void* zmel,
                                      int64_t zmel_stride, int nbb, int nqtot, const int* icounkmin,
                                      const int* icounkmax, int ncoun, const int* iwini, const int* iwend,
                                      const int* itc, const int* itpc, const int* jpmc, const int* icouini,
                                      int ncount, const double* whwc) {
  if (!h || !rcxq || !zmel || !ns1 || !ns2 || !icounkmin || !icounkmax || !iwini || !iwend || !itc || !itpc ||
      !jpmc || !icouini || !whwc)
    return ECALJ_ERR_ARG;
  return x0_impl(h, rcxq, npr, ipr_col, npr_col, nwhis, npm, nk, ns1, ns2, zmel, zmel_stride, nbb, nqtot,
                 icounkmin, icounkmax, ncoun, iwini, iwend, itc, itpc, jpmc, icouini, ncount, whwc);
}

ecalj_status ecalj_x0gemm(ecalj_handle h, double* rcxq, int npr, int ipr_col, int npr_col, int nwhis, int npm,
                          int ns1, int ns2, const double* zmel, int nbb, int nqtot, int icoun_lo, int icoun_hi,
                          const int* iwini, const int* iwend, const int* itc, const int* itpc, const int* jpmc,
                          const int* icouini, const double* whwc) {
  if (!h || !rcxq || !zmel) return ECALJ_ERR_ARG;
  if (icoun_hi < icoun_lo) return ECALJ_OK;
  const int nm = ns2 - ns1 + 1;
  if (nm < 1) return ECALJ_ERR_DIM;
  int ncount = 0;  // the caller's whwc covers the indices its lists name
  for (int ic = icoun_lo; ic <= icoun_hi; ++ic)
    ncount = std::max(ncount, icouini[ic - 1] + iwend[ic - 1] - iwini[ic - 1]);
  const int64_t nslice = (npm == 1) ? (int64_t)nwhis + 1 : 2 * (int64_t)nwhis + 1;
  const int64_t nr = (int64_t)npr * npr_col * nslice;
  const int64_t nz = (int64_t)nbb * nm * nqtot;
  DevC dr, dz;
  ECALJ_TRY(dr.put(h, rcxq, nr));
  ECALJ_TRY(dz.put(h, zmel, nz));
  ECALJ_TRY(ecalj_x0gemm_batched_dev(h, dr.buf.p, npr, ipr_col, npr_col, nwhis, npm, 1, &ns1, &ns2, dz.buf.p, 0, nbb,
                                     nqtot, &icoun_lo, &icoun_hi, icoun_hi, iwini, iwend, itc, itpc, jpmc, icouini,
                                     ncount, whwc));
  return dr.get(h, rcxq, nr);
}

// ------------------------------------------------------------------ dpsion
ecalj_status ecalj_dpsion_imag_dev(ecalj_handle h, void* rcxq, int npr, int npr_col, int nwhis, int niw,
                                   const double* frhis, const double* freqi, double ecut, int scale_rcxq,
                                   void* zxqi) {
  if (!h || !rcxq || !frhis || !freqi || !zxqi) return ECALJ_ERR_ARG;
  if (npr < 0 || npr_col < 0 || nwhis < 1 || niw < 1) return ECALJ_ERR_DIM;
  std::vector<double> coef, scale;
  hilbert_coef(nwhis, niw, frhis, freqi, ecut, coef, scale);
  std::vector<double> blob(coef);
  blob.insert(blob.end(), scale.begin(), scale.end());
  double* d;
  ECALJ_TRY(upload(h, ecalj_ctx::S_COEF, blob, &d));
  const int64_t n = (int64_t)npr * npr_col;
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 4, &tev));
  ECALJ_TRY(hilbert(h, wide(h), rcxq, n, nwhis, niw, hilbert_ldc(niw), d, scale_rcxq ? d + coef.size() : nullptr,
                    nullptr, (cd*)zxqi, n));
  ECALJ_TRY(timing_end(h, 4, tev));
  return ECALJ_OK;
}

ecalj_status ecalj_dpsion_imag(ecalj_handle h, double* rcxq, int npr, int npr_col, int nwhis, int niw,
                               const double* frhis, const double* freqi, double ecut, int scale_rcxq, double* zxqi) {
  if (!h || !rcxq || !zxqi) return ECALJ_ERR_ARG;
  const int64_t nr = (int64_t)npr * npr_col * (nwhis + 1);
  const size_t bx = sizeof(cd) * (size_t)npr * npr_col * niw;
  DevC dr;
  DevBuf dx;
  ECALJ_TRY(dr.put(h, rcxq, nr));
  ECALJ_TRY(dx.alloc(bx));
  ECALJ_TRY(ecalj_dpsion_imag_dev(h, dr.buf.p, npr, npr_col, nwhis, niw, frhis, freqi, ecut, scale_rcxq, dx.p));
  if (scale_rcxq) ECALJ_TRY(dr.get(h, rcxq, nr));
  ECALJ_HIP(hipMemcpyAsync(zxqi, dx.p, bx, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

ecalj_status ecalj_dpsion_real_dev(ecalj_handle h, const void* rcxq, int npr, int npr_col, int nwhis, int nw,
                                   const double* frhis, const double* freqr, double ecut, int rcxq_scaled,
                                   void* zxq) {
  if (!h || !rcxq || !frhis || !freqr || !zxq) return ECALJ_ERR_ARG;
  if (npr < 0 || npr_col < 0 || nwhis < 1 || nw < 0) return ECALJ_ERR_DIM;
  if (nwhis <= nw) return ECALJ_ERR_DIM;  // S/dpsion5.f90:203
  if (freqr[0] != 0.0) return ECALJ_ERR_ARG;  // S/dpsion5.f90:204
  std::vector<double> coef, dscale;
  hilbert_coef_real(nwhis, nw, frhis, freqr, ecut, rcxq_scaled != 0, coef, dscale);
  std::vector<double> blob(coef);
  blob.insert(blob.end(), dscale.begin(), dscale.end());
  double* d;
  ECALJ_TRY(upload(h, ecalj_ctx::S_COEF, blob, &d));
  const int64_t n = (int64_t)npr * npr_col;
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 4, &tev));
  ECALJ_TRY(hilbert(h, wide(h), const_cast<void*>(rcxq), n, nwhis, nw + 1, hilbert_ldc(nw + 1), d, nullptr,
                    d + coef.size(), (cd*)zxq, n));
  ECALJ_TRY(timing_end(h, 4, tev));
  return ECALJ_OK;
}

ecalj_status ecalj_dpsion_real(ecalj_handle h, const double* rcxq, int npr, int npr_col, int nwhis, int nw,
                               const double* frhis, const double* freqr, double ecut, int rcxq_scaled, double* zxq) {
  if (!h || !rcxq || !zxq) return ECALJ_ERR_ARG;
  const int64_t nr = (int64_t)npr * npr_col * (nwhis + 1);
  const size_t bx = sizeof(cd) * (size_t)npr * npr_col * (nw + 1);
  DevC dr;
  DevBuf dx;
  ECALJ_TRY(dr.put(h, rcxq, nr));
  ECALJ_TRY(dx.alloc(bx));
  ECALJ_TRY(ecalj_dpsion_real_dev(h, dr.buf.p, npr, npr_col, nwhis, nw, frhis, freqr, ecut, rcxq_scaled, dx.p));
  ECALJ_HIP(hipMemcpyAsync(zxq, dx.p, bx, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

// ------------------------------------------------------------------ composed chi0
ecalj_status ecalj_chi0_freq_dev(ecalj_handle h, const void* z, const void* w, int ni, int npairs, int nf,
                                 void* chi) {
  if (!h || !z || !w || !chi) return ECALJ_ERR_ARG;
  if (ni < 0 || npairs < 0 || nf < 0) return ECALJ_ERR_DIM;
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 1, &tev));
  ECALJ_TRY(chi0_freq(h, h->precision, z, (const double*)w, ni, npairs, nf, (cd*)chi));
  ECALJ_TRY(timing_end(h, 1, tev));
  return ECALJ_OK;
}

ecalj_status ecalj_chi0_freq(ecalj_handle h, const double* z, const double* w, int ni, int npairs, int nf,
                             double* chi) {
  if (!h || !z || !w || !chi) return ECALJ_ERR_ARG;
  DevC dz;
  DevBuf dw, dc;
  ECALJ_TRY(dz.put(h, z, (int64_t)ni * npairs));
  ECALJ_TRY(dw.put(h, w, sizeof(double) * (size_t)nf * npairs));
  const size_t bc = sizeof(cd) * (size_t)ni * ni * nf;
  ECALJ_TRY(dc.alloc(bc));
  ECALJ_TRY(ecalj_chi0_freq_dev(h, dz.buf.p, dw.p, ni, npairs, nf, dc.p));
  ECALJ_HIP(hipMemcpyAsync(chi, dc.p, bc, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

// ------------------------------------------------------------------ W build
// WVIllwI (imaginary axis, epsilon Hermitian positive definite: in-order
// elimination) and WVRllwR (real axis, epsilon general: row exchanges).
// FP64 form: every step in complex128.  Mixed forms: the reference's __MP
// split (S/m_kind.f90:3-6, S/m_llw.f90:31-39,62-64): the M <-> E transforms
// on the precision's GEMM core with complex64 chi_E and W_E (cmm_d), epsilon
// and its inverse in complex128 (epstilde is complex(8) in every build).
static ecalj_status wc_build_impl(ecalj_handle h, const void* chi, int ngb, int nw, int nspin, int is_x0_m_basis,
                                  int is_wc_m_basis, void* wc, int ldw, bool pivot) {
  if (!h || !chi || !wc) return ECALJ_ERR_ARG;
  if (ngb < 1 || nw < 1 || ldw < ngb || (nspin != 1 && nspin != 2)) return ECALJ_ERR_DIM;
  if (!h->d_vcousq || h->ngb != ngb) return ECALJ_ERR_STATE;
  const int n = ngb;
  const int64_t nn = (int64_t)n * n;
  const double fac = (nspin == 1) ? 2.0 : 1.0;  // S/m_llw.f90:299-303
  cd *eps, *tmp;
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_EPS, sizeof(cd) * nn * nw, (void**)&eps));
  ECALJ_TRY(slot_ensure(h, ecalj_ctx::S_TMP, sizeof(cd) * nn * nw, (void**)&tmp));
  int tev = -1;
  ECALJ_TRY(timing_begin(h, 3, &tev));
  ECALJ_TRY(eps_build(h, chi, n, nw, fac, is_x0_m_basis != 0, 0.0, eps, tmp));  // (S/m_llw.f90:336-348)
  const int* perm = nullptr;
  ECALJ_TRY(gj_inverse(h, eps, n, n, nn, nw, pivot ? &perm : nullptr));  // zminv (S/m_llw.f90:143,350)
  ECALJ_TRY(w_form(h, eps, perm, n, nw, is_wc_m_basis != 0, (cd*)wc, ldw, tmp));  // (S/m_llw.f90:353-367)
  ECALJ_TRY(timing_end(h, 3, tev));
  return ECALJ_OK;
}

ecalj_status ecalj_wc_build_dev(ecalj_handle h, const void* chi, int ngb, int nw, int nspin, int is_x0_m_basis,
                                int is_wc_m_basis, void* wc, int ldw) {
  return wc_build_impl(h, chi, ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, wc, ldw, false);
}

ecalj_status ecalj_wc_build_real_dev(ecalj_handle h, const void* chi, int ngb, int nw, int nspin, int is_x0_m_basis,
                                     int is_wc_m_basis, void* wc, int ldw) {
  return wc_build_impl(h, chi, ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, wc, ldw, true);
}

ecalj_status ecalj_wc_build_real(ecalj_handle h, const double* chi, int ngb, int nw, int nspin, int is_x0_m_basis,
                                 int is_wc_m_basis, double* wc, int ldw) {
  if (!h || !chi || !wc) return ECALJ_ERR_ARG;
  DevBuf dc, dw;
  ECALJ_TRY(dc.put(h, chi, sizeof(cd) * (size_t)ngb * ngb * nw));
  const size_t bw = sizeof(cd) * (size_t)ldw * ngb * nw;
  ECALJ_TRY(dw.put(h, wc, bw));
  ECALJ_TRY(ecalj_wc_build_real_dev(h, dc.p, ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, dw.p, ldw));
  ECALJ_HIP(hipMemcpyAsync(wc, dw.p, bw, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return check_flag(h);
}

// zminv_d (S/m_lapack.f90:108-176), batched: in-place inverse with partial pivoting.
ecalj_status ecalj_zminv_dev(ecalj_handle h, void* a, int n, int lda, int nbatch, int64_t stride) {
  if (!h || !a) return ECALJ_ERR_ARG;
  if (n < 0 || lda < std::max(n, 1) || nbatch < 0 || (nbatch > 1 && stride < (int64_t)lda * n))
    return ECALJ_ERR_DIM;
  if (n == 0 || nbatch == 0) return ECALJ_OK;
  const int* perm = nullptr;
  ECALJ_TRY(gj_inverse(h, (cd*)a, n, lda, stride, nbatch, &perm));
  ECALJ_TRY(gj_unscramble(h, (cd*)a, n, lda, stride, nbatch, perm));
  return ECALJ_OK;
}

// ecalj_zminv_dev on the device's fixed scratch array: the same kernels, no
// allocation, no host synchronisation.  Calls that overlap on one device race
// on the array.
ecalj_status ecalj_zminv_noalloc_dev(ecalj_handle h, void* a, int n, int lda, int nbatch, int64_t stride) {
  if (!h || !a) return ECALJ_ERR_ARG;
  if (n < 0 || lda < std::max(n, 1) || nbatch < 0 || (nbatch > 1 && stride < (int64_t)lda * n))
    return ECALJ_ERR_DIM;
  if (n == 0 || nbatch == 0) return ECALJ_OK;
  if (n > 8192 || nbatch > 65535) return ECALJ_ERR_DIM;  // limits of gj_inverse
  if (zminv_scratch_bytes(n, nbatch) > (size_t)ECALJ_NOALLOC_SCRATCH_BYTES) return ECALJ_ERR_SCRATCH;
  void* base = scratch_ptr(h->device);
  if (!base || h->arena) return ECALJ_ERR_STATE;
  h->arena = (unsigned char*)base;
  h->arena_cap = ECALJ_NOALLOC_SCRATCH_BYTES;
  h->arena_used = 0;
  std::fill(h->arena_len, h->arena_len + ecalj_ctx::S_NSLOT, (size_t)0);
  const int* perm = nullptr;
  ecalj_status st = gj_inverse(h, (cd*)a, n, lda, stride, nbatch, &perm);
  if (st == ECALJ_OK) st = gj_unscramble(h, (cd*)a, n, lda, stride, nbatch, perm);
  h->arena = nullptr;
  h->arena_cap = h->arena_used = 0;
  return st;
}

ecalj_status ecalj_zminv(ecalj_handle h, double* a, int n, int lda) {
  if (!h || !a) return ECALJ_ERR_ARG;
  if (n < 0 || lda < std::max(n, 1)) return ECALJ_ERR_DIM;
  const size_t bytes = sizeof(cd) * (size_t)lda * n;
  DevBuf d;
  ECALJ_TRY(d.put(h, a, bytes));
  ECALJ_TRY(ecalj_zminv_dev(h, d.p, n, lda, 1, (int64_t)lda * n));
  ECALJ_HIP(hipMemcpyAsync(a, d.p, bytes, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return check_flag(h);
}

ecalj_status ecalj_wc_build(ecalj_handle h, const double* chi, int ngb, int nw, int nspin, int is_x0_m_basis,
                            int is_wc_m_basis, double* wc, int ldw) {
  if (!h || !chi || !wc) return ECALJ_ERR_ARG;
  DevBuf dc, dw;
  ECALJ_TRY(dc.put(h, chi, sizeof(cd) * (size_t)ngb * ngb * nw));
  const size_t bw = sizeof(cd) * (size_t)ldw * ngb * nw;
  ECALJ_TRY(dw.put(h, wc, bw));
  ECALJ_TRY(ecalj_wc_build_dev(h, dc.p, ngb, nw, nspin, is_x0_m_basis, is_wc_m_basis, dw.p, ldw));
  ECALJ_HIP(hipMemcpyAsync(wc, dw.p, bw, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return check_flag(h);
}

// ------------------------------------------------------------------ Sigma_c
ecalj_status ecalj_sigc_weights_dev(ecalj_handle h, int nbatch, int niw, int ns1, int ns2, int ntq, int nctot,
                                    const double* omega, const double* ekc, double esmr, double ua,
                                    const double* wkkr, const double* freqx, const double* wwx,
                                    const double* expa, void* wgtim) {
  if (!h || !omega || !ekc || !wkkr || !freqx || !wwx || !expa || !wgtim) return ECALJ_ERR_ARG;
  if (niw < 1 || ns2 < ns1 || ntq < 0 || nbatch < 1) return ECALJ_ERR_DIM;
  return sig_weights(h, nbatch, niw, ns2 - ns1 + 1, ntq, ns1, nctot, omega, ekc, esmr, ua, wkkr, freqx, wwx, expa,
                     (double*)wgtim);
}

ecalj_status ecalj_sigc_weights(ecalj_handle h, int niw, int ns1, int ns2, int ntq, int nctot, const double* omega,
                                const double* ekc, double esmr, double ua, double wkkr, const double* freqx,
                                const double* wwx, const double* expa, double* wgtim) {
  if (!h || !omega || !ekc || !freqx || !wwx || !expa || !wgtim) return ECALJ_ERR_ARG;
  const int nm = ns2 - ns1 + 1;
  if (nm < 1) return ECALJ_ERR_DIM;
  DevBuf d_om, d_ek, d_wk, d_x, d_w, d_e, d_g;
  ECALJ_TRY(d_om.put(h, omega, sizeof(double) * ntq));
  ECALJ_TRY(d_ek.put(h, ekc, sizeof(double) * nm));
  ECALJ_TRY(d_wk.put(h, &wkkr, sizeof(double)));
  ECALJ_TRY(d_x.put(h, freqx, sizeof(double) * niw));
  ECALJ_TRY(d_w.put(h, wwx, sizeof(double) * niw));
  ECALJ_TRY(d_e.put(h, expa, sizeof(double) * niw));
  const size_t bg = sizeof(double) * (size_t)(niw + 1) * nm * ntq;
  ECALJ_TRY(d_g.alloc(bg));
  ECALJ_TRY(ecalj_sigc_weights_dev(h, 1, niw, ns1, ns2, ntq, nctot, (double*)d_om.p, (double*)d_ek.p, esmr, ua,
                                   (double*)d_wk.p, (double*)d_x.p, (double*)d_w.p, (double*)d_e.p, d_g.p));
  ECALJ_HIP(hipMemcpyAsync(wgtim, d_g.p, bg, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

ecalj_status ecalj_sigc_dev(ecalj_handle h, int nbatch, const void* zmel, int ngb, int nm, int ntq, int niw,
                            const void* wc, int ldw, const void* wgtim, void* zsec, int ldz) {
  if (!h || !zmel || !wc || !wgtim || !zsec) return ECALJ_ERR_ARG;
  if (ngb < 1 || nm < 0 || ntq < 0 || niw < 0 || ldw < ngb || ldz < ntq || nbatch < 1) return ECALJ_ERR_DIM;
  return sig_contract(h, h->precision, nbatch, zmel, ngb, nm, ntq, niw, (const cd*)wc, ldw, (const double*)wgtim,
                      (cd*)zsec, ldz);
}

ecalj_status ecalj_sigc_grouped_dev(ecalj_handle h, int nk, const int* nm, const void* zmel, int ngb, int ntq,
                                    int niw, const void* wc, int ldw, const void* wgtim, void* zsec, int ldz) {
  if (!h || !nm || !zmel || !wc || !wgtim || !zsec) return ECALJ_ERR_ARG;
  if (ngb < 1 || ntq < 0 || niw < 0 || ldw < ngb || ldz < ntq || nk < 1) return ECALJ_ERR_DIM;
  if (ntq == 0) return ECALJ_OK;
  return sig_contract_grouped(h, h->precision, nk, nm, zmel, ngb, ntq, niw, (const cd*)wc, ldw,
                              (const double*)wgtim, (cd*)zsec, ldz);
}

ecalj_status ecalj_sigc_real_dev(ecalj_handle h, int nbatch, const void* zmel, int ngb, int nm, int ntq,
                                 const void* wr, int ldw, int64_t wr_stride, int nw, const double* freq_r,
                                 const double* omega, const double* ekc, double ef, double esmr, int nctot, int ns2r,
                                 const double* wkkr, void* zsec, int ldz) {
  if (!h || !zmel || !wr || !freq_r || !omega || !ekc || !wkkr || !zsec) return ECALJ_ERR_ARG;
  if (ngb < 1 || nm < 1 || ntq < 0 || nbatch < 1 || ldw < ngb || ldz < ntq || nw < 2 || ns2r < 0 || ns2r > nm ||
      nctot < 0 || wr_stride < 0)
    return ECALJ_ERR_DIM;
  if (ntq == 0) return ECALJ_OK;
  RealSpec rs;
  rs.wr = (const cd*)wr; rs.ldw = ldw; rs.wstride = wr_stride; rs.nw = nw;
  rs.freq_r = freq_r; rs.omega = omega; rs.ekc = ekc; rs.wkkr = wkkr;
  rs.ef = ef; rs.esmr = esmr; rs.nctot = nctot; rs.ns2r = ns2r;
  return sig_real(h, h->precision, nbatch, zmel, ngb, nm, ntq, rs, (cd*)zsec, ldz);
}

ecalj_status ecalj_sigc_full_dev(ecalj_handle h, int nbatch, const void* zmel, int ngb, int nm, int ntq, int niw,
                                 const void* wc, int ldw, const void* wgtim, const void* wr, int ldwr,
                                 int64_t wr_stride, int nw, const double* freq_r, const double* omega,
                                 const double* ekc, double ef, double esmr, int nctot, int ns2r, const double* wkkr,
                                 void* zsec, int ldz) {
  if (!h || !zmel || !wc || !wgtim || !wr || !freq_r || !omega || !ekc || !wkkr || !zsec) return ECALJ_ERR_ARG;
  if (ngb < 1 || nm < 1 || ntq < 0 || niw < 0 || nbatch < 1 || ldw < ngb || ldwr < ngb || ldz < ntq || nw < 2 ||
      ns2r < 0 || ns2r > nm || nctot < 0 || wr_stride < 0)
    return ECALJ_ERR_DIM;
  if (ntq == 0) return ECALJ_OK;
  RealSpec rs;
  rs.wr = (const cd*)wr; rs.ldw = ldwr; rs.wstride = wr_stride; rs.nw = nw;
  rs.freq_r = freq_r; rs.omega = omega; rs.ekc = ekc; rs.wkkr = wkkr;
  rs.ef = ef; rs.esmr = esmr; rs.nctot = nctot; rs.ns2r = ns2r;
  return sig_contract(h, h->precision, nbatch, zmel, ngb, nm, ntq, niw, (const cd*)wc, ldw, (const double*)wgtim,
                      (cd*)zsec, ldz, &rs);
}

ecalj_status ecalj_sigc_real(ecalj_handle h, const double* zmel, int ngb, int nm, int ntq, const double* wr, int ldw,
                             int nw, const double* freq_r, const double* omega, const double* ekc, double ef,
                             double esmr, int nctot, int ns2r, double wkkr, double* zsec, int ldz) {
  if (!h || !zmel || !wr || !zsec) return ECALJ_ERR_ARG;
  DevC dz;
  DevBuf dw, ds;
  ECALJ_TRY(dz.put(h, zmel, (int64_t)ngb * nm * ntq));
  ECALJ_TRY(dw.put(h, wr, sizeof(cd) * (size_t)ldw * ngb * (nw + 1)));
  const size_t bs = sizeof(cd) * (size_t)ldz * ntq;
  ECALJ_TRY(ds.put(h, zsec, bs));
  ECALJ_TRY(ecalj_sigc_real_dev(h, 1, dz.buf.p, ngb, nm, ntq, dw.p, ldw, 0, nw, freq_r, omega, ekc, ef, esmr, nctot,
                                ns2r, &wkkr, ds.p, ldz));
  ECALJ_HIP(hipMemcpyAsync(zsec, ds.p, bs, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

ecalj_status ecalj_sigc(ecalj_handle h, const double* zmel, int ngb, int nm, int ntq, int niw, const double* wc,
                        int ldw, const double* wgtim, double* zsec, int ldz) {
  if (!h || !zmel || !wc || !wgtim || !zsec) return ECALJ_ERR_ARG;
  DevC dz;
  DevBuf dw, dg, ds;
  ECALJ_TRY(dz.put(h, zmel, (int64_t)ngb * nm * ntq));
  ECALJ_TRY(dw.put(h, wc, sizeof(cd) * (size_t)ldw * ngb * (niw + 1)));
  ECALJ_TRY(dg.put(h, wgtim, sizeof(double) * (size_t)(niw + 1) * nm * ntq));
  const size_t bs = sizeof(cd) * (size_t)ldz * ntq;
  ECALJ_TRY(ds.put(h, zsec, bs));
  ECALJ_TRY(ecalj_sigc_dev(h, 1, dz.buf.p, ngb, nm, ntq, niw, dw.p, ldw, dg.p, ds.p, ldz));
  ECALJ_HIP(hipMemcpyAsync(zsec, ds.p, bs, hipMemcpyDeviceToHost, h->stream));
  ECALJ_HIP(hipStreamSynchronize(h->stream));
  return ECALJ_OK;
}

// ------------------------------------------------------------------ QP update
ecalj_status ecalj_qp_energy(int ntq, const double* elda, const double* sex, const double* sexcore,
                             const double* sec, int ldsec, const double* vxc, double* eqp) {
  if (!elda || !sex || !sexcore || !sec || !vxc || !eqp) return ECALJ_ERR_ARG;
  if (ntq < 0 || ldsec < ntq) return ECALJ_ERR_DIM;
  const double hartree = 2.0 * 13.605693122994;  // hartree = 2*rydberg()
  for (int i = 0; i < ntq; ++i) {
    const double rsec = sec[2 * (i + (int64_t)ldsec * i)] * hartree;
    eqp[i] = elda[i] + (sex[i] * hartree + sexcore[i] * hartree + rsec - vxc[i]);
  }
  return ECALJ_OK;
}

ecalj_status ecalj_sigma_herm(int ntq, const double* sex2, const double* sexcore2, const double* sec,
                              const double* vxcmat, double* se) {
  if (!sex2 || !sexcore2 || !sec || !vxcmat || !se) return ECALJ_ERR_ARG;
  if (ntq < 0) return ECALJ_ERR_DIM;
  for (int j = 0; j < ntq; ++j)
    for (int i = 0; i < ntq; ++i) {
      const int64_t ij = 2 * (i + (int64_t)ntq * j), ji = 2 * (j + (int64_t)ntq * i);
      se[ij] = sex2[ij] + sexcore2[ij] + 0.5 * (sec[ij] + sec[ji]) - vxcmat[ij];
      se[ij + 1] = sex2[ij + 1] + sexcore2[ij + 1] + 0.5 * (sec[ij + 1] - sec[ji + 1]) - vxcmat[ij + 1];
    }
  return ECALJ_OK;
}

}  // extern "C"
