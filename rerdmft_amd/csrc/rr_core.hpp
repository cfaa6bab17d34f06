// Internal interface shared by the translation units of librerdmft_gpu (not installed; the public
// interface is include/rerdmft_gpu.h). Each translation unit has its own kernels; what is shared
// here are the context, error handling, workspace management and the host-side GEMM / small-matrix
// drivers.
#pragma once

#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "rr_gemm.hpp"
#include "../../include/rerdmft_gpu.h"

namespace rr {


// ================================================================================================
// Context, errors, workspace
// ================================================================================================
struct Tensor {
  size_t naux = 0, n = 0, ld = 0;
  double* f64 = nullptr;
  float* f32 = nullptr;
  bf16_t* b16 = nullptr;
  bool symmetric = true;
  size_t naux_nz = SIZE_MAX;  // aux rows from this index on are all zero (SIZE_MAX: not known)
};

// Aux rows that can hold non-zero values (at least 1 when the tensor has rows).
inline size_t live_rows(const Tensor& t) { return t.naux ? std::max<size_t>(1, std::min(t.naux, t.naux_nz)) : 0; }

constexpr int kSlots = 16;

// A set of 4c RKB Cholesky vectors (RkbCholesky, src/C4_DHF/RkbCholesky.h): per vector L the five
// nl x nl components large (promoted to complex), small[0], small[1], small_cross[0],
// small_cross[1], complex interleaved, rows padded to ld = roundup(nl, 8). `nat` holds them as
// given ([comp][L][row][ld]); `tr` holds the first three transposed.
constexpr int kRkbSets = 4;
struct RkbChol {
  size_t nl = 0, nchol = 0, ld = 0;
  double* nat = nullptr;
  double* tr = nullptr;
  // max |B - B^H| over the large and the two small components, relative to max |B| (measured at
  // upload): the factor-form exchange needs Hermitian flavor vectors
  double herm_dev = 0.0;
};

// A captured HIP graph together with the shape it was recorded for and the workspace generation
// its buffers belong to.
struct GraphPlan {
  std::vector<long long> key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  unsigned long long gen = 0;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  std::string err;
  Tensor slot[kSlots];
  std::map<int, std::pair<void*, size_t>> ws;
  int tile_override = -1;
  int lds_override = -1;  // -1 auto, -2 generic kernel only, >= 0 forced LDS variant
  bool autotune = true;
  int last_ntiles = 0;  // tiles per batch of the most recent GEMM launch
  int num_cu = 256;     // compute units of the device (persistent kernels launch one workgroup each)
  std::map<std::vector<long long>, std::pair<int, int>> tuned;  // shape key -> (generic cfg, LDS variant)
  float last_ms[8] = {0};
  RkbChol rkb[kRkbSets];
  void* zpage = nullptr;  // 256 zero bytes (source of out-of-range chunks in direct-to-LDS GEMMs)
  unsigned long long ws_generation = 0;  // bumped whenever a workspace buffer is (re)allocated
  GraphPlan multi_plan;                  // rr_ri_jk_multi
  // Second stream for work that can overlap the exchange products (the Coulomb stream reads B while
  // the rank-k update reads W); null in contexts without one. Default priority, like `stream`.
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // Occupied-orbital recovery from a density: 0 spectral projector (eigensolver when it does not
  // converge), 1 eigensolver only; counts of the recoveries each method finished.
  int recovery_method = 0;
  long long n_projector = 0, n_eigensolver = 0;
  // Allocation-free calls: while `arena` is set, workspace requests are carved from it in request
  // order (256-byte aligned, a region per id and size) instead of allocated; a request past
  // arena_size is an error. Set only for the duration of such a call (ArenaScope).
  char* arena = nullptr;
  size_t arena_size = 0, arena_used = 0;
  std::map<int, std::pair<void*, size_t>> arena_ws;
};

// Routes the context's workspace requests to a caller-provided device region for one call.
struct ArenaScope {
  Context& c;
  ArenaScope(Context& cc, void* base, size_t size) : c(cc) {
    c.arena = static_cast<char*>(base), c.arena_size = size, c.arena_used = 0;
    c.arena_ws.clear();
  }
  ~ArenaScope() {
    c.arena = nullptr, c.arena_size = 0, c.arena_used = 0;
    c.arena_ws.clear();
  }
};

struct Error {
  int code;
  std::string msg;
};

#define RR_HIP(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw Error{RR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)};             \
  } while (0)

#define RR_REQUIRE(cond, code, msg) \
  do {                              \
    if (!(cond)) throw Error{code, msg}; \
  } while (0)

inline size_t roundup(size_t x, size_t m) { return (x + m - 1) / m * m; }
// Row stride of a stored 3-index tensor: a multiple of 32 elements (so that every 64-byte bf16
// chunk a GEMM stage loads lies inside one 128-byte cache line) when that pads by at most 1/16,
// else a multiple of 8 (16-byte vectors).
inline size_t tensor_ld(size_t n) {
  const size_t m = n > 0 ? n : 1, a = roundup(m, 32);
  return (a - m) * 16 <= m ? a : roundup(m, 8);
}


// Named grow-only device workspace buffers.
void* wsbuf(Context& c, int id, size_t bytes);

enum WsId {
  WS_STAGE = 1, WS_DT, WS_GPART, WS_GAMMA, WS_JPART, WS_LT, WS_W, WS_KPART, WS_OUT_F, WS_OUT_J, WS_OUT_K,
  WS_H, WS_S, WS_X, WS_P, WS_T1, WS_T2, WS_T3, WS_T4, WS_EIGW, WS_EIGE, WS_INFO, WS_COCC, WS_W2, WS_CA, WS_CB,
  WS_MO_A, WS_MO_B, WS_MO_S1, WS_MO_OUT, WS_DENS2, WS_TMP8, WS_TMP9
};
// Translation units other than rr_lib.hip number their workspace buffers from here.
constexpr int WS_SIGMA_BASE = 100000;

// True when p is device (or managed) memory visible to the current device.
bool is_device_ptr(const void* p);

// Copy host-or-device array into a device workspace buffer (returns device pointer).
template <typename T>
T* to_device(Context& c, int id, const T* src, size_t count) {
  T* d = static_cast<T*>(wsbuf(c, id, count * sizeof(T)));
  if (count) RR_HIP(hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyDefault, c.stream));
  return d;
}
template <typename T>
void from_device(Context& c, T* dst, const T* src, size_t count) {
  if (count) RR_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDefault, c.stream));
}

inline int grid_for(size_t total, int block = 256, int maxg = 65536) {
  size_t g = (total + block - 1) / block;
  if (g < 1) g = 1;
  return (int)std::min<size_t>(g, maxg);
}

enum Prec { P64 = 0, P32 = 1, P16 = 2 };

struct GemmSpec {
  int prec = P64;      // input type
  int out = P64;       // output type
  bool cplx = false;
  int cfg = 1;
  int lds = -1;        // >= 0: LDS-staged kernel variant (real only), see kLds
};

// Tiled MFMA GEMM C = alpha A B^T + beta C on the context stream (rr_gemm.hpp semantics).
void run_gemm(Context& c, const GemmSpec& sp, GemmArgs g);
// As run_gemm, with the per-shape tile autotuner for large products.
void gemm_auto(Context& c, GemmSpec sp, const GemmArgs& g);
size_t prec_size(int p);
struct Tensor;
const void* tensor_data(const Tensor& t, int prec);
// Small FP64 (or complex FP64, interleaved) row-major product C = alpha op(A) op(B) + beta C,
// op in {'N', 'T', 'C'}; pointers are device memory.
void dgemm_small(Context& c, bool cplx, char opa, char opb, size_t m, size_t n, size_t k, const double* A, size_t lda,
                 const double* B, size_t ldb, double* C, size_t ldc, double alpha = 1.0, double beta = 0.0);
// Hermitian eigensolve in place (row j of a = eigenvector j), ascending eigenvalues in w.
void eig_device(Context& c, bool cplx, size_t n, double* a, double* w);
// Occupied set of a density: the eigenvectors of M = scale X^H (S P S) X with eigenvalue above 1/2 as
// n_occ rows of length nx in *rows (device), with their weights in occ_host (eigenvalues, or ones for
// the rows of a weighted factor F with F F^H = sum lambda v v^H when `weighted`). All pointers device.
size_t recover_occupied(Context& c, bool cplx, size_t n, size_t nx, const double* dS, const double* dP, const double* dX,
                        double scale, const double** rows, std::vector<double>& occ_host, bool weighted);
Tensor& get_slot(Context& c, int slot);
void check_prec(int prec);
// FP64 MO tensors into device memory: buildRiMoEri (complex, (n_aux, nt^2)) and the embedded
// NON_REL spin-orbital tensor (real, (n_aux, (2n)^2)).
void mo_tensor_rkb(Context& c, int slot_ll, int slot_ss, size_t nl, size_t ns, const double* dC_ukb, size_t nt,
                   double* dout);
void mo_tensor_nonrel_embedded(Context& c, int slot, size_t n, const double* dC_spatial, double* dout);
size_t slot_naux(Context& c, int slot);
// A whole batch of factor-form J/K builds (one per tensor, D = L L^T) as a few grouped launches on
// the context stream, with results identical to rr_ri_jk_factor (D = L L^T) per tensor. L_m is read from
// lin + loff[m] (n x r row-major), J_m and K_m are written to jout / kout + ooff[m].
struct BatchPlan;
BatchPlan* batch_plan_create(Context& c, const std::vector<Tensor>& t, int prec, const std::vector<size_t>& r,
                             const double* lin, const std::vector<size_t>& loff, double* jout, double* kout,
                             const std::vector<size_t>& ooff);
void batch_plan_enqueue(Context& c, const BatchPlan* p);
void batch_plan_free(BatchPlan* p);
// Tensor storage outside the numbered slots (batches of molecules).
void tensor_free(Tensor& t);
void tensor_store(Context& c, Tensor& t, const double* eri3, size_t naux, size_t n, unsigned copies);
// Switches the per-phase HIP-event timing on or off; returns the previous state.
bool set_phase_timing(bool on);

// Output helper: device pointer if the caller's buffer is device memory, else a workspace buffer
// that is copied back by finish_out().
struct OutBuf {
  double* dev = nullptr;
  double* user = nullptr;
  size_t count = 0;
  bool copy = false;
};
OutBuf out_buf(Context& c, int id, double* user, size_t count);
void finish_out(Context& c, const OutBuf& o);
// Input helper: the caller's pointer if it is device memory, else a device copy in workspace `id`.
const double* in_buf(Context& c, int id, const double* user, size_t count);

}  // namespace rr

struct rr_context {
  rr::Context ctx;
};

namespace rr {
// Runs f(ctx) with the device current; maps exceptions to status codes and keeps the message.
template <typename F>
int guarded(rr_handle h, F&& f) {
  if (!h) return RR_ERR_ARG;
  try {
    RR_HIP(hipSetDevice(h->ctx.device));
    f(h->ctx);
    RR_HIP(hipStreamSynchronize(h->ctx.stream));
    h->ctx.err.clear();
    return RR_OK;
  } catch (const Error& e) {
    h->ctx.err = e.msg;
    (void)hipStreamSynchronize(h->ctx.stream);
    return e.code;
  } catch (const std::exception& e) {
    h->ctx.err = e.what();
    return RR_ERR_INTERNAL;
  }
}
}  // namespace rr
