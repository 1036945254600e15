// ctx.hpp -- handle state and host-side launch helpers.
#pragma once
#include <array>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"

struct RotState;  // eigenfunction-rotation tables (rotwave.hip)

struct ecalj_ctx {
  int device = 0;
  int ncu = 256;                 // compute units of the device (persistent grids)
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int precision = ECALJ_PREC_F64;  // F64, F32 or BF16X3 (BF16X3_F8 is BF16X3 + sigc_fp8)
  bool sigc_fp8 = false;           // Sigma_c contraction with fp8 cross terms

  // set_ppb state: per-atom blocks
  int natom = 0;
  std::vector<int> nv, mdim;
  std::vector<int64_t> ppb_off;   // element offset of each atom block
  int ndima_used = 0, ngb_ppb = 0;
  double* d_ppb = nullptr;        // original layout ppb(i1,i2,I) per atom
  double* d_ppb2 = nullptr;       // relaid: P2[(i1 + nv*I) + nv*mdim*i2]
  // core blocks of ecalj_set_ppb_full: per atom (nv, ncore, mdim) = ppb(nc+b, c, I)
  double* d_ppbc = nullptr;
  std::vector<int64_t> ppbc_off;
  std::vector<int> ncore;         // core functions (nlnmc) per atom; empty or zeros: none
  // Coulomb state
  int ngb = 0;
  double* d_vcousq = nullptr;     // sqrt(vcoud)
  ecalj::cd* d_zcousq = nullptr;  // ngb x ngb
  ecalj::cf* d_zcousq32 = nullptr;  // complex64 copy for the mixed-precision M->E transform
  // set_ppovl state (m_read_ppovl of the current q): the interstitial step
  // function's Fourier table ggg with its index cube, and the combined-G list
  // with its index cube (host copies for argument checks)
  ecalj::cd* d_ggg = nullptr;
  int* d_igggi = nullptr;
  int ggg_lo[3] = {0, 0, 0}, ggg_n[3] = {0, 0, 0};
  int* d_nvgcgp2 = nullptr;
  int ngcgp = 0;
  int* d_igcgp2i = nullptr;
  int gcp_lo[3] = {0, 0, 0}, gcp_n[3] = {0, 0, 0};
  std::vector<int> h_igcgp2i;
  // m_hamindex and @MNLA_CPHI state of readcphi / readgeig
  RotState* rot = nullptr;

  // grow-only scratch slots
  enum { S_T = 0, S_WH, S_CZ, S_PART, S_LIST, S_GJP, S_GJC, S_EPS, S_TMP, S_TMP2, S_COEF, S_STAGE,
         S_CVT1, S_CVT2, S_MLP, S_TAIL, S_GTAB, S_RTAB, S_RSLOT, S_KST, S_GJPIV, S_GJD, S_GJOC, S_GJR, S_IPWT, S_IPW1, S_IPW2, S_IPW3, S_IPWB, S_GAM, S_ROT1, S_ROT2, S_ROT3, S_QA, S_QB, S_QW, S_GJW, S_X0A, S_X0B, S_ZA, S_ZB, S_ZPA, S_ZPB,
         S_EIG1, S_EIG2, S_EIG3, S_EIG4, S_EIG5, S_EIG6, S_NSLOT };
  void* slot[S_NSLOT] = {};
  size_t slot_bytes[S_NSLOT] = {};
  // fixed-arena mode (the entry points that allocate nothing): while arena is
  // set, slot_ensure carves each slot out of [arena, arena + arena_cap) by a
  // bump pointer instead of allocating; the slot table above is untouched
  unsigned char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  size_t arena_off[S_NSLOT] = {}, arena_len[S_NSLOT] = {};
  int* d_flag = nullptr;          // singular-pivot flag
  // pinned host staging per slot for asynchronous uploads; the event marks
  // the end of the last copy out of it
  void* pin[S_NSLOT] = {};
  size_t pin_bytes[S_NSLOT] = {};
  hipEvent_t pin_ev[S_NSLOT] = {};

  // HIP graph capture of the handle's stream (ecalj_graph_begin/end): scratch
  // slots may not grow, uploads stage through fresh pinned buffers that the
  // graph keeps, kernel timing is off, the tuner uses cached choices only
  bool capturing = false;
  // bumped whenever the handle frees or moves device memory that captured
  // kernels may reference (scratch regrowth, set_ppb, set_coulomb, ...);
  // a graph captured at another generation is refused at launch
  uint64_t state_gen = 0;
  uint64_t id = 0;                // unique per handle of the process
  hipEvent_t ev_join = nullptr;   // ecalj_stream_wait: last point of this stream another waited for
  bool timing_before_capture = false;
  std::vector<void*> cap_pinned;

  // kernel-variant tuner: 0 = measure each variant on the first call of a shape and
  // keep the fastest; k > 0 = always variant k (1-based)
  int tune_mode = 0;
  std::map<std::array<int64_t, 5>, int> tuned;
  std::string tune_log;

  // kernel timing
  bool timing = false;
  struct EvPair { hipEvent_t a, b; };
  std::vector<EvPair> ev_pool[5];
  int ev_used[5] = {0, 0, 0, 0, 0};
};

namespace ecalj {

ecalj_status slot_ensure(ecalj_ctx* h, int slot, size_t bytes, void** out);
void rot_free(ecalj_ctx* h);
// Device buffer owned by one host-pointer entry point (freed on return).
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  ecalj_status alloc(size_t bytes) {
    if (bytes == 0) bytes = 16;
    return hipMalloc(&p, bytes) == hipSuccess ? ECALJ_OK : ECALJ_ERR_NOMEM;
  }
  ecalj_status put(ecalj_ctx* h, const void* src, size_t bytes) {
    ECALJ_TRY(alloc(bytes));
    if (src && bytes) ECALJ_HIP(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, h->stream));
    return ECALJ_OK;
  }
};
// Waits for the stream; ECALJ_ERR_SINGULAR (and the flag cleared) when a
// kernel met a zero pivot since the last check.
ecalj_status check_flag(ecalj_ctx* h);
// Copies a host table to scratch slot `slot` on the handle's stream (pinned
// staging; capture-safe); *out is its device address.
ecalj_status upload_blob(ecalj_ctx* h, int slot, const std::vector<unsigned char>& v, unsigned char** out);

// ---- kernel-variant tuner
enum TuneOp { TUNE_SIGC_BF3 = 0, TUNE_SIGC_BF3_GROUPED = 1 };
typedef std::array<int64_t, 5> TuneKey;
bool tune_lookup(ecalj_ctx* h, const TuneKey& k, int* var);
void tune_store(ecalj_ctx* h, const TuneKey& k, int var, const char* op, const std::vector<const char*>& names,
                const std::vector<double>& ms);
// Times f() (a launch sequence on h->stream) with HIP events; waits for it.
template <class F>
ecalj_status tune_time(ecalj_ctx* h, F&& f, double* ms) {
  hipEvent_t a, b;
  ECALJ_HIP(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return ECALJ_ERR_HIP;
  }
  ecalj_status st = ECALJ_OK;
  if (hipEventRecord(a, h->stream) != hipSuccess) st = ECALJ_ERR_HIP;
  if (st == ECALJ_OK) st = f();
  if (st == ECALJ_OK && hipEventRecord(b, h->stream) != hipSuccess) st = ECALJ_ERR_HIP;
  if (st == ECALJ_OK && hipEventSynchronize(b) != hipSuccess) st = ECALJ_ERR_HIP;
  float t = 0.f;
  if (st == ECALJ_OK && hipEventElapsedTime(&t, a, b) != hipSuccess) st = ECALJ_ERR_HIP;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = t;
  return st;
}
// Event bracketing for the kernel classes of ecalj_kernel_timing.
ecalj_status timing_begin(ecalj_ctx* h, int kclass, int* idx);
ecalj_status timing_end(ecalj_ctx* h, int kclass, int idx);

// Generic complex GEMM (FP64 MFMA).  ta/tb: operand stored transposed
// (op = T); ca/cb: operand conjugated.  op 'C' of BLAS = ta + ca.  A may be real.
struct GemmCall {
  int mode = 0;  // M64 (complex128), M32 / MBF3 (complex64)
  bool ta = false, tb = false, ca = false, cb = false;
  bool areal = false;
  int M = 0, N = 0, K = 0;
  const void* A = nullptr; int64_t lda = 0; Batch2 bA{0, 0};
  const void* B = nullptr; int64_t ldb = 0; Batch2 bB{0, 0};
  void* C = nullptr; int64_t ldc = 0; Batch2 bC{0, 0};
  int batch = 1, div = 1;
  cd alpha{1.0, 0.0}, beta{0.0, 0.0};
  int conjC = 0;
  int epi = 0;
  const double* svec = nullptr;
  int gj_k0 = 0, gj_k1 = 0;
};
ecalj_status gemm(ecalj_ctx* h, const GemmCall& c);

}  // namespace ecalj
